"""The C ABI library loads and exports what include/streamgc.h declares; the
product's host pipeline (C++17) reproduces the reference's bytes.  No GPU."""
import hashlib
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S
from streamgc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GRAPHS = ["k4", "c100", "star6", "triangle", "weighted_general", "pattern_dups", "er2k", "chunglu4k",
          "rmat_s10", "delaunay1k"]


def sha(b):
    return hashlib.sha256(b).hexdigest()


def declared_functions():
    text = open(os.path.join(ROOT, "include", "streamgc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sgc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers every declared function
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_record_layout():
    assert S.EDGE_UPDATE.itemsize == 16
    r = S.make_updates([1], [2], [1])
    raw = r.tobytes()
    assert raw[:8] == (1).to_bytes(4, "little") + (2).to_bytes(4, "little") and raw[8] == 1
    assert raw[9:] == bytes(7)


def test_default_config_matches_reference():
    c = _lib.Config()
    _lib.lib().sgc_default_config(c)
    want = S.StreamGCConfig().to_c()
    for k, _ in _lib.Config._fields_:
        assert getattr(c, k) == getattr(want, k), k
    assert (c.epoch_size, c.batch_size, c.histogram_threshold, c.slack_pct, c.min_row_capacity,
            c.overflow_capacity) == (32768, 32768, 64, 300, 8, 1 << 18)


def test_version_and_device_probe_do_not_fail():
    assert b"gfx950" in _lib.lib().sgc_version()
    assert S.device_count() >= 0


@pytest.mark.skipif(S.device_count() > 0, reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    g = S.csr_from_edges(3, [(0, 1)])
    with pytest.raises(S.SgcError, match="no HIP device"):
        S.GPUPartition(g)


@pytest.mark.parametrize("name", GRAPHS)
def test_loader_matches_reference(name, golden, graph_path):
    g = S.load_graph_csr(graph_path(name))
    want = golden["graphs"][name]
    assert g.n == want["n"] and g.num_edges == want["nnz"]
    assert sha(g.row_ptr.tobytes() + g.col_idx.tobytes() + g.degree.tobytes()) == want["csr_sha256"]


@pytest.mark.parametrize("name", GRAPHS)
def test_init_stream_matches_reference(name, golden, graph_path):
    g = S.load_graph_csr(graph_path(name))
    r = S.generate_init_stream(g)
    pairs = np.stack([r["u"], r["v"]], axis=1).astype(np.uint32)
    assert len(r) == golden["graphs"][name]["init"]["count"]
    assert sha(pairs.tobytes()) == golden["graphs"][name]["init"]["sha256"]
    assert np.all(r["type"] == 0)


@pytest.mark.parametrize("name", GRAPHS)
def test_churn_stream_matches_reference(name, golden, graph_path):
    g = S.load_graph_csr(graph_path(name))
    for key, want in golden["graphs"][name]["bench"].items():
        seed, size, pct = (int(x) for x in key.split("/"))
        r = S.generate_bench_stream_edges(g, size, pct, seed)
        assert sha(r.tobytes()) == want["sha256"], key


def test_churn_stream_prefix_is_size_independent(graph_path):
    g = S.load_graph_csr(graph_path("er2k"))
    a = S.generate_bench_stream_edges(g, 5000, 50, 3)
    b = S.generate_bench_stream_edges(g, 12000, 50, 3)
    assert np.array_equal(a, b[:5000])


def test_static_greedy_matches_oracle(graph_path):
    for name in ["k4", "c100", "er2k", "chunglu4k", "rmat_s10", "delaunay1k"]:
        g = S.load_graph_csr(graph_path(name))
        assert S.static_greedy_coloring(g) == O.static_greedy(O.Csr(g.n, g.row_ptr, g.col_idx, g.degree))


def test_edge_intervals_give_the_live_set(graph_path):
    g = S.load_graph_csr(graph_path("er2k"))
    r = S.generate_bench_stream_edges(g, 4000, 50, 9)
    eu, ev, born, died = S.edge_intervals(g, r)
    og = O.Csr(g.n, g.row_ptr, g.col_idx, g.degree)
    for p in (0, 1, 777, 2500, 4000):
        alive = (born <= p) & (p < died)
        got = set(zip(eu[alive].tolist(), ev[alive].tolist()))
        lu, lv = O.live_edges_after(og, r["u"][:p], r["v"][:p], r["type"][:p])
        assert got == set(zip(lu.tolist(), lv.tolist())), p


# Every ctypes structure of the binding against the header's own layout: a
# program compiled from include/streamgc.h prints sizeof and the offset of
# every field, and the binding must agree field for field.
_STRUCTS = {"sgc_config": "Config", "sgc_epoch_stats": "EpochStats", "sgc_totals": "Totals",
            "sgc_tuning": "Tuning", "sgc_pool_config": "PoolConfig", "sgc_graph_info": "GraphInfo",
            "sgc_batch_config": "BatchConfig", "sgc_batch_stats": "BatchStats",
            "sgc_server_config": "ServerConfig", "sgc_server_stats": "ServerStats",
            "sgc_trace_item": "TraceItem", "sgc_replay_stats": "ReplayStats", "sgc_run_meta": "RunMeta"}


def test_ctypes_structures_match_the_header(tmp_path):
    import ctypes
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    lines = ["#include <cstdio>", "#include <cstddef>", '#include "streamgc.h"', "int main() {"]
    for cname, pyname in _STRUCTS.items():
        lines.append(f'  std::printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in getattr(_lib, pyname)._fields_:
            lines.append(f'  std::printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cxx, "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = {}
    for line in out:
        if line:
            c, f, v = line.split()
            want[(c, f)] = int(v)
    for cname, pyname in _STRUCTS.items():
        cls = getattr(_lib, pyname)
        assert ctypes.sizeof(cls) == want[(cname, "size")], cname
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == want[(cname, f)], (cname, f)
