"""Pools: many graphs on one GPU, one launch per phase for all of them.

Deterministic pools must give every graph exactly what it gets alone (and
what the oracle gets); parallel pools must keep every graph proper; the HBM
budget must refuse what does not fit.
"""
import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S

pytestmark = pytest.mark.gpu

NAMES = ["er2k", "chunglu4k", "rmat_s10", "delaunay1k", "c100", "k4", "star6", "weighted_general"]


def _mixed(streams, seed):
    """Interleave per-graph streams at random, keeping each graph's own order."""
    rng = np.random.default_rng(seed)
    owners = np.concatenate([np.full(len(s), g, np.int64) for g, s in streams.items()])
    rng.shuffle(owners)
    pos = {g: 0 for g in streams}
    out = np.zeros(len(owners), dtype=S.EDGE_UPDATE)
    for i, g in enumerate(owners.tolist()):
        out[i] = streams[g][pos[g]]
        pos[g] += 1
    return out


def _graphs(graph_path):
    return {k: S.load_graph_csr(graph_path(n)) for k, n in enumerate(NAMES)}


@pytest.mark.parametrize("repair", [False, True])
def test_deterministic_pool_equals_each_graph_alone(repair, graph_path):
    graphs = _graphs(graph_path)
    E = 160
    cfg = S.StreamGCConfig(epoch_size=E, batch_size=E)
    pool = S.GraphPool(max_graphs=16, batch_capacity=4096, deterministic=True, repair_every_epoch=repair,
                       exact_histogram=repair)
    gid = {k: pool.add(g, cfg) for k, g in graphs.items()}
    init = {gid[k]: S.tag_graph(S.generate_init_stream(g), gid[k]) for k, g in graphs.items()}
    pool.submit(_mixed(init, 1))
    pool.finish()
    bench = {gid[k]: S.tag_graph(S.generate_bench_stream_edges(g, 700, 50, 11 + k), gid[k])
             for k, g in graphs.items()}
    pool.submit(_mixed(bench, 2))
    for k, g in graphs.items():
        solo = S.GPUPartition(g, S.StreamGCConfig(epoch_size=E, batch_size=E, deterministic=True,
                                                  repair_every_epoch=repair, exact_histogram=repair))
        solo.process_stream(S.generate_init_stream(g))
        solo.finish()
        solo.process_stream(S.generate_bench_stream_edges(g, 700, 50, 11 + k))
        assert np.array_equal(pool.words(gid[k]), solo.words()), NAMES[k]
        assert np.array_equal(pool.snapshot(gid[k]), solo.snapshot()), NAMES[k]
        flags = (O.FLAG_REPAIR_EVERY_EPOCH | O.FLAG_EXACT_HISTOGRAM) if repair else 0
        orc = O.Engine(O.Csr(g.n, g.row_ptr, g.col_idx, g.degree),
                       O.default_config(epoch_size=E, batch_size=E, flags=flags))
        i0 = S.generate_init_stream(g)
        orc.submit(i0["u"], i0["v"], i0["type"])
        orc.finish()
        b0 = S.generate_bench_stream_edges(g, 700, 50, 11 + k)
        orc.submit(b0["u"], b0["v"], b0["type"])
        assert np.array_equal(pool.words(gid[k]), orc.words()), NAMES[k]


def test_parallel_pool_keeps_every_graph_proper(graph_path):
    graphs = _graphs(graph_path)
    E = 256
    cfg = S.StreamGCConfig(epoch_size=E, batch_size=E)
    pool = S.GraphPool(max_graphs=16, batch_capacity=8192, repair_every_epoch=True, exact_histogram=True)
    gid = {k: pool.add(g, cfg) for k, g in graphs.items()}
    init = {gid[k]: S.tag_graph(S.generate_init_stream(g), gid[k]) for k, g in graphs.items()}
    pool.submit(_mixed(init, 3))
    pool.finish()
    streams = {}
    for k, g in graphs.items():
        streams[k] = S.generate_bench_stream_edges(g, 1024, 50, 5 + k)
        pool.history_reserve(gid[k], 8)
    st = pool.submit(_mixed({gid[k]: S.tag_graph(s, gid[k]) for k, s in streams.items()}, 4))
    assert st and all(e["t2_remaining"] == 0 and e["snapshot_published"] for e in st)
    for k, g in graphs.items():
        n_snap = pool.history_count(gid[k])
        assert n_snap == 4, NAMES[k]
        bad = pool.validate_history(gid[k], S.edge_intervals(g, streams[k]), [E * (j + 1) for j in range(n_snap)])
        assert bad.tolist() == [0] * n_snap, NAMES[k]
        ids = np.arange(g.n, dtype=np.uint32)
        assert np.array_equal(pool.query(gid[k], ids), pool.snapshot(gid[k]))


def test_admission_respects_the_hbm_budget(graph_path):
    g = S.load_graph_csr(graph_path("chunglu4k"))
    cfg = S.StreamGCConfig()
    need = S.GraphPool.graph_bytes(g, cfg)
    pool = S.GraphPool(max_graphs=8, batch_capacity=4096, hbm_budget=int(need * 2.5))
    a = pool.add(g, cfg)
    b = pool.add(g, cfg)
    with pytest.raises(S.SgcNoMemory):
        pool.add(g, cfg)
    used, budget, k = pool.usage()
    assert k == 2 and used == 2 * need and used <= budget
    pool.remove(a)
    c = pool.add(g, cfg)
    assert pool.usage()[2] == 2 and c == a
    # the re-admitted graph works
    pool.submit(S.tag_graph(S.generate_init_stream(g), c))
    st = pool.finish([c, b])
    assert st[0]["t2_remaining"] == 0
