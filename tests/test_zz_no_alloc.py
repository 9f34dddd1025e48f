"""sgc_pool_query_noalloc: the allocation-free query entry point.

It stages ids and colours in a static device array instead of the pool's
growable buffers, so its answers must equal sgc_pool_query's on the same
published snapshots: on two calls in a row, the second with fewer ids (the
array keeps the first call's contents, which must not leak into the second
answer), on ids past the graph's end (colour 0), before any publish (zeros),
at exactly the maximum size, and with a clear error beyond it.  Both entry
points are called through the C ABI (ctypes).
"""
import ctypes as C

import numpy as np
import pytest

import streamgc_amd as S
from streamgc_amd import _lib

pytestmark = pytest.mark.gpu

MAX = 65536          # SGC_QUERY_NOALLOC_MAX in include/streamgc.h
SGC_ERR_RANGE = -4


def _call(fn, pool, gid, ids):
    ids = np.ascontiguousarray(ids, np.uint32)
    out = np.full(len(ids), 0xBEEF, np.uint16)    # a value no colour here takes
    rc = fn(pool.h, gid, ids, len(ids), out)
    return rc, out


@pytest.fixture(scope="module")
def pool(graph_path):
    p = S.GraphPool(max_graphs=4, batch_capacity=1 << 14, repair_every_epoch=True, exact_histogram=True)
    gs = [S.load_graph_csr(graph_path(n)) for n in ("chunglu4k", "delaunay1k")]
    gids = [p.add(g, S.StreamGCConfig(epoch_size=2048, batch_size=2048)) for g in gs]
    for gid, g in zip(gids, gs):
        p.submit(S.tag_graph(S.generate_init_stream(g), gid))
    p.finish(gids)
    for k, (gid, g) in enumerate(zip(gids, gs)):
        p.submit(S.tag_graph(S.generate_bench_stream_edges(g, 3000, 50, 5 + k), gid))
    p.finish(gids)
    yield p, gids, gs
    p.close()


def test_noalloc_query_equals_the_allocating_query_on_two_calls(pool):
    p, gids, gs = pool
    lib = _lib.lib()
    rng = np.random.default_rng(7)
    big = rng.integers(0, gs[0].n + 100, 50000).astype(np.uint32)   # some ids past the end
    small = rng.integers(0, gs[1].n, 1000).astype(np.uint32)
    for gid, ids in ((gids[0], big), (gids[1], small)):
        rc_a, want = _call(lib.sgc_pool_query, p, gid, ids)
        rc_b, got = _call(lib.sgc_pool_query_noalloc, p, gid, ids)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(got, want), gid
        assert np.all(got[ids >= p.graphs[gid].n] == 0)
        snap = p.snapshot(gid)
        inb = ids < p.graphs[gid].n
        assert np.array_equal(got[inb], snap[ids[inb]])


def test_noalloc_query_at_the_maximum_and_beyond(pool):
    p, gids, gs = pool
    lib = _lib.lib()
    ids = (np.arange(MAX, dtype=np.uint64) % gs[1].n).astype(np.uint32)
    rc_a, want = _call(lib.sgc_pool_query, p, gids[1], ids)
    rc_b, got = _call(lib.sgc_pool_query_noalloc, p, gids[1], ids)
    assert rc_a == 0 and rc_b == 0 and np.array_equal(got, want)
    over = np.zeros(MAX + 1, np.uint32)
    rc, out = _call(lib.sgc_pool_query_noalloc, p, gids[1], over)
    assert rc == SGC_ERR_RANGE
    assert np.all(out == 0xBEEF)                    # nothing written
    msg = lib.sgc_pool_last_error(p.h).decode()
    assert "65536" in msg and "query_noalloc" in msg, msg
    rc, _ = _call(lib.sgc_pool_query_noalloc, p, 99, ids[:4])   # no such graph
    assert rc != 0
    rc, out = _call(lib.sgc_pool_query_noalloc, p, gids[0], ids[:0])
    assert rc == 0 and len(out) == 0


def test_noalloc_query_before_any_publish_is_zero(graph_path):
    g = S.load_graph_csr(graph_path("k4"))
    p = S.GraphPool(max_graphs=2, batch_capacity=256)
    gid = p.add(g, S.StreamGCConfig(epoch_size=64, batch_size=64))
    lib = _lib.lib()
    ids = np.arange(g.n, dtype=np.uint32)
    rc_a, want = _call(lib.sgc_pool_query, p, gid, ids)
    rc_b, got = _call(lib.sgc_pool_query_noalloc, p, gid, ids)
    assert rc_a == 0 and rc_b == 0
    assert np.array_equal(got, want) and np.all(got == 0)
    p.close()
