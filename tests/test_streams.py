"""The other update streams of the reference's README.md:83-91 (no reference
code; the table there specifies them): uniform, hold-out and dense.  Checked
against what the table says, with independent numpy set arithmetic."""
import numpy as np

import streamgc_amd as S


def _edges(g):
    u = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.row_ptr.astype(np.int64)))
    v = g.col_idx.astype(np.uint64)
    keep = u < v
    return set(((u[keep] << np.uint64(32)) | v[keep]).tolist())


def _keys(rec):
    u = rec["u"].astype(np.uint64)
    v = rec["v"].astype(np.uint64)
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    return ((lo << np.uint64(32)) | hi).tolist()


def test_uniform_inserts_new_non_adjacent_pairs(graph_path):
    g = S.load_graph_csr(graph_path("er2k"))
    st = S.uniform_stream(g, 5000, seed=3)
    assert len(st) == 5000 and (st["type"] == 0).all()
    assert (st["u"] != st["v"]).all() and (st["u"] < g.n).all() and (st["v"] < g.n).all()
    k = _keys(st)
    assert len(set(k)) == len(k)                       # each pair new
    assert not (set(k) & _edges(g))                    # non-adjacent in the graph
    assert np.array_equal(S.uniform_stream(g, 5000, seed=3), st)
    assert not np.array_equal(S.uniform_stream(g, 5000, seed=4), st)
    # uniform over pairs: both endpoints spread over the whole id range
    ends = np.concatenate([st["u"], st["v"]])
    counts = np.bincount(ends * 8 // g.n, minlength=8)
    assert counts.min() > 0.8 * counts.mean()


def test_holdout_splits_the_edges(graph_path):
    g = S.load_graph_csr(graph_path("chunglu4k"))
    all_e = _edges(g)
    base, st = S.holdout_stream(g, pct=10, seed=7)
    assert base.n == g.n
    assert len(st) == (len(all_e) * 10 + 50) // 100
    assert (st["type"] == 0).all()
    held = _keys(st)
    assert len(set(held)) == len(held)
    assert set(held) | _edges(base) == all_e and not (set(held) & _edges(base))
    assert base.num_edges == 2 * (len(all_e) - len(held))
    b2, s2 = S.holdout_stream(g, pct=10, seed=7)
    assert np.array_equal(s2, st) and np.array_equal(b2.col_idx, base.col_idx)
    # in random order, not CSR order
    assert held != sorted(held)
    _, s0 = S.holdout_stream(g, pct=0, seed=7)
    assert len(s0) == 0


def test_dense_stream_stays_in_the_top_cores(graph_path):
    g = S.load_graph_csr(graph_path("rmat_s10"))
    core = S.core_numbers(g)
    cmin = int(np.percentile(core, 90, method="lower"))
    st = S.dense_stream(g, 4000, seed=2)
    assert (core[st["u"]] >= cmin).all() and (core[st["v"]] >= cmin).all()
    ins = int((st["type"] == 0).sum())
    assert 0.4 < ins / len(st) < 0.6                   # one insert per delete on average
    grow = S.dense_growth_stream(g, 4000, seed=2)
    assert 0.6 < (grow["type"] == 0).mean() < 0.73    # two inserts per delete
