"""Writes the golden vectors under tests/golden/ from the reference's own host code.

Runs only where the reference checkout is present: it needs
oracle/_ref/ref_streams, which `make -C oracle ref` compiles from the
reference's A/src/host/stream_gen.cpp and A/src/host/graph_loader.cpp.

Outputs (committed):
  graphs/*.mtx     small input graphs (hand-made edge cases plus seeded random ones)
  streams.json     for every graph: the reference loader's CSR hash, the init
                   stream hash, and churn-stream hashes over seeds x sizes x
                   insert percentages, each with its first and last records.
"""
import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_streams")
GRAPHS = os.path.join(HERE, "graphs")

SEEDS = (42, 7)
SIZES = (1000, 100000)
PCTS = (20, 50, 80)
KEEP = 16  # records kept at each end of a stream


def write_mtx(path, n, edges, banner="%%MatrixMarket matrix coordinate pattern symmetric",
              comments=(), weights=None, rows=None, cols=None):
    with open(path, "w") as f:
        f.write(banner + "\n")
        for c in comments:
            f.write("%" + c + "\n")
        f.write(f"{rows or n} {cols or n} {len(edges)}\n")
        for i, (u, v) in enumerate(edges):
            if weights is None:
                f.write(f"{u + 1} {v + 1}\n")
            else:
                f.write(f"{u + 1} {v + 1} {weights[i]}\n")


def make_graphs():
    os.makedirs(GRAPHS, exist_ok=True)
    out = {}

    def add(name, *a, **k):
        p = os.path.join(GRAPHS, name + ".mtx")
        write_mtx(p, *a, **k)
        out[name] = p

    add("k4", 4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
    add("c100", 100, [(i, (i + 1) % 100) for i in range(100)])
    add("star6", 6, [(0, i) for i in range(1, 6)])
    add("triangle", 3, [(0, 1), (1, 2), (0, 2)])
    # general + weights + self loops + duplicates in both directions + comments
    add("weighted_general", 9,
        [(0, 1), (1, 0), (2, 2), (3, 4), (4, 3), (3, 4), (5, 8), (8, 7), (7, 7), (6, 1)],
        banner="%%MatrixMarket matrix coordinate real general",
        comments=("a comment", " another"), weights=[1.5, 2, -3, 4e-3, 5, 6, 7, 8, 9, 10])
    # pattern with duplicates, self loops, and trailing isolated vertices (cols > max id)
    add("pattern_dups", 12, [(0, 3), (3, 0), (0, 3), (1, 1), (2, 5), (5, 2), (4, 6), (6, 9)],
        rows=10, cols=12)

    rng = np.random.default_rng(20241)
    # Erdos-Renyi, n=2000, 8000 sampled pairs
    n = 2000
    e = rng.integers(0, n, size=(8000, 2))
    add("er2k", n, [(int(a), int(b)) for a, b in e])
    # heavy-tailed (Chung-Lu style) n=4096 so that some degrees exceed 64
    n = 4096
    w = (np.arange(n) + 8.0) ** (-1 / 1.3)
    p = w / w.sum()
    a = rng.choice(n, size=40000, p=p)
    b = rng.choice(n, size=40000, p=p)
    perm = rng.permutation(n)
    add("chunglu4k", n, [(int(perm[x]), int(perm[y])) for x, y in zip(a, b)])
    # R-MAT scale 10, edge factor 16, (a, b, c, d) = (.57, .19, .19, .05)
    n, m = 1 << 10, 16 << 10
    src = np.zeros(m, dtype=np.int64)
    dst = np.zeros(m, dtype=np.int64)
    for lvl in range(10):
        r = rng.random(m)
        bit_s = (r >= 0.57 + 0.19).astype(np.int64)          # quadrants c, d
        bit_d = (((r >= 0.57) & (r < 0.76)) | (r >= 0.95)).astype(np.int64)  # quadrants b, d
        src |= bit_s << lvl
        dst |= bit_d << lvl
    add("rmat_s10", n, [(int(x), int(y)) for x, y in zip(src, dst)])
    # Delaunay triangulation of 1024 uniform points (edge set from scipy/Qhull)
    from scipy.spatial import Delaunay
    pts = rng.random((1024, 2))
    tri = Delaunay(pts)
    es = set()
    for s in tri.simplices:
        for i in range(3):
            x, y = int(s[i]), int(s[(i + 1) % 3])
            es.add((min(x, y), max(x, y)))
    add("delaunay1k", 1024, sorted(es))
    return out


def sha(b):
    return hashlib.sha256(b).hexdigest()


def run(*args):
    subprocess.run([REF_BIN, *args], check=True, stdout=subprocess.DEVNULL)


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build oracle/_ref/ref_streams first: make -C oracle ref")
    graphs = make_graphs()
    result = {"keep": KEEP, "seeds": SEEDS, "sizes": SIZES, "pcts": PCTS, "graphs": {}}
    with tempfile.TemporaryDirectory() as td:
        for name, path in sorted(graphs.items()):
            csr = os.path.join(td, name + ".csr")
            run("mtx", path, csr)
            raw = open(csr, "rb").read()
            n, nnz = struct.unpack_from("<IQ", raw, 4)
            body = raw[16:]
            g = {"n": n, "nnz": nnz, "csr_sha256": sha(body)}
            rp = np.frombuffer(body, dtype=np.uint32, count=n + 1)
            deg = np.frombuffer(body, dtype=np.uint32, offset=4 * (n + 1 + nnz), count=n)
            g["max_deg"] = int(deg.max()) if n else 0
            g["deg_sha256"] = sha(deg.tobytes())
            assert int(rp[-1]) == nnz
            pairs = os.path.join(td, name + ".pairs")
            run("init", csr, pairs)
            pb = open(pairs, "rb").read()
            cnt = struct.unpack_from("<Q", pb, 0)[0]
            arr = np.frombuffer(pb, dtype=np.uint32, offset=8).reshape(-1, 2)
            g["init"] = {"count": cnt, "sha256": sha(pb[8:]),
                         "head": arr[:KEEP].tolist(), "tail": arr[-KEEP:].tolist()}
            g["bench"] = {}
            for seed in SEEDS:
                for size in SIZES:
                    for pct in PCTS:
                        out = os.path.join(td, "u.bin")
                        run("bench", csr, str(size), str(pct), str(seed), out)
                        ub = open(out, "rb").read()
                        cnt = struct.unpack_from("<Q", ub, 0)[0]
                        rec = np.frombuffer(ub, dtype=np.uint32, offset=8).reshape(-1, 4)
                        trip = np.stack([rec[:, 0], rec[:, 1], rec[:, 2] & 0xFF], axis=1)
                        g["bench"][f"{seed}/{size}/{pct}"] = {
                            "count": cnt, "sha256": sha(ub[8:]),
                            "head": trip[:KEEP].tolist(), "tail": trip[-KEEP:].tolist()}
            result["graphs"][name] = g
            print(name, n, nnz, file=sys.stderr)
    with open(os.path.join(HERE, "streams.json"), "w") as f:
        json.dump(result, f, indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
