"""The synthetic multi-graph workloads of the batched (C4) and served (C5)
patterns, shared by bench.py and the tests so both run the same shapes.

C4: 256 graphs, half Delaunay triangulations of 2^U[10,16] uniform points
(log-uniform), half R-MAT scale U[10,16] with edge factor U[8,40]; graph i is
seeded 42 + i and gets a churn stream of U[1K, 64K] updates at U[30,70]%
insert, with E_g = min(stream, 4096).

C5: 64 tenant graphs cycling through the C1-C4 shapes at tenant scale
(Delaunay 2^U[10,14], R-MAT scale U[10,14] ef U[8,40], Chung-Lu 2^U[12,15]
vertices at average degree 16, R-MAT scale 12 ef 48), seeded 42 + i, plus one
com-Orkut-shape Chung-Lu graph (3,072,441 vertices, 117,185,083 edges).
"""
import numpy as np

ORKUT = dict(n=3072441, m=117185083, gamma=2.3, max_degree=33313.0)


def batch_graphs(S, count=256, seed=42):
    """[(kind, graph, churn length, insert %)] of the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i < count // 2:
            g = S.gen_delaunay(int(round(2 ** rng.uniform(10, 16))), seed=seed + i)
            kind = "delaunay"
        else:
            g = S.gen_rmat(int(rng.integers(10, 17)), float(rng.uniform(8, 40)), seed=seed + i)
            kind = "rmat"
        out.append((kind, g, int(rng.integers(1000, 64001)), int(rng.integers(30, 71))))
    return out


def batch_epoch_sizes(spec):
    """E_g = min(stream, 4096) per graph of batch_graphs()."""
    return [min(L, 4096) for _, _, L, _ in spec]


def tenants(S, count=64, seed=42, orkut=True):
    """[(kind, graph)] of the served workload: `count` tenants of the C1-C4
    shapes, then (orkut=True) the com-Orkut-shape graph."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = ("delaunay", "rmat", "chung_lu", "rmat_dense")[i % 4]
        if kind == "delaunay":
            g = S.gen_delaunay(int(2 ** rng.uniform(10, 14)), seed=seed + i)
        elif kind == "rmat":
            g = S.gen_rmat(int(rng.integers(10, 15)), float(rng.uniform(8, 40)), seed=seed + i)
        elif kind == "chung_lu":
            n = int(2 ** rng.uniform(12, 15))
            g = S.gen_chung_lu(n, n * 8, 2.3, 4.0 * n ** 0.5, seed=seed + i, permute=True)
        else:
            g = S.gen_rmat(12, 48.0, seed=seed + i)
        out.append((kind, g))
    if orkut:
        out.append(("orkut", S.gen_chung_lu(ORKUT["n"], ORKUT["m"], ORKUT["gamma"], ORKUT["max_degree"], seed=seed,
                                            permute=True)))
    return out


def chunked_launches(streams, chunk):
    """Launch j holds records [j*chunk, (j+1)*chunk) of every stream that has them."""
    allr = np.concatenate(streams)
    launch = np.concatenate([np.arange(len(s), dtype=np.int64) // chunk for s in streams])
    order = np.argsort(launch, kind="stable")
    bounds = np.concatenate([[0], np.cumsum(np.bincount(launch))])
    allr = allr[order]
    return [allr[bounds[j]:bounds[j + 1]] for j in range(len(bounds) - 1)]
