"""HIP engine vs the oracle, through the C ABI.

Deterministic mode (one wavefront, stream order, repair in vertex order) must
be bit-exact with the oracle: packed words (colour + version), every slack
slot, bitmaps, histograms, snapshots and per-epoch statistics.  The parallel
mode is nondeterministic like the reference: every repaired colouring must be
proper against the live edge set, and the maximum colour close to the
oracle's (the colour-count definition of A/src/streamgc.cu:549-553).
"""
import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S

pytestmark = pytest.mark.gpu

GRAPHS = ["k4", "c100", "star6", "triangle", "weighted_general", "pattern_dups", "er2k", "chunglu4k",
          "rmat_s10", "delaunay1k"]


def _pair(g, **cfg):
    sc = S.StreamGCConfig(**cfg)
    oc = O.default_config(epoch_size=sc.epoch_size, batch_size=sc.batch_size,
                          histogram_threshold=sc.histogram_threshold, slack_pct=sc.slack_pct,
                          min_row_capacity=sc.min_row_capacity, overflow_capacity=sc.overflow_capacity,
                          max_repair_iters=sc.max_repair_iters, plateau_iters=sc.plateau_iters,
                          tier2_passes=sc.tier2_passes,
                          flags=(O.FLAG_REPAIR_EVERY_EPOCH if sc.repair_every_epoch else 0)
                          | (O.FLAG_EXACT_HISTOGRAM if sc.exact_histogram else 0)
                          | sum(S._lib.ABLATIONS[a] for a in sc.ablations))
    og = O.Csr(g.n, g.row_ptr, g.col_idx, g.degree)
    return S.GPUPartition(g, sc), O.Engine(og, oc), og


def _same_state(gpu, orc, where):
    gw, ow = gpu.words(), orc.words()
    if not np.array_equal(gw, ow):
        bad = np.nonzero(gw != ow)[0]
        raise AssertionError(f"{where}: {len(bad)} words differ, first {bad[:8].tolist()} "
                             f"gpu {gw[bad[:8]].tolist()} oracle {ow[bad[:8]].tolist()}")
    grp, gcol = gpu.csr()
    orp, ocol = orc.csr()
    assert np.array_equal(grp, orp), where
    if not np.array_equal(gcol, ocol):
        bad = np.nonzero(gcol != ocol)[0]
        raise AssertionError(f"{where}: {len(bad)} slots differ, first {bad[:8].tolist()}")
    assert np.array_equal(gpu.bitmaps(), orc.bitmaps()), where
    assert np.array_equal(gpu.hist(), orc.hist()), where
    assert gpu.overflow_count() == orc.overflow_count(), where


def _records(u, v, t=None):
    return S.make_updates(u, v, t)


@pytest.mark.parametrize("mode", ["reference_order", "repair_every_epoch"])
@pytest.mark.parametrize("name", GRAPHS)
def test_deterministic_bit_exact(name, mode, graph_path):
    g = S.load_graph_csr(graph_path(name))
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=96, batch_size=96,
                         repair_every_epoch=(mode == "repair_every_epoch"))
    init = S.generate_init_stream(g)
    ge = gpu.process_stream(init)
    oe = orc.submit(init["u"], init["v"], init["type"])
    assert ge == oe
    _same_state(gpu, orc, "init stream")
    gs, os_ = gpu.finish(), orc.finish()
    assert gs == os_
    _same_state(gpu, orc, "init repair")
    assert np.array_equal(gpu.snapshot(), orc.snapshot())
    bench = S.generate_bench_stream_edges(g, 1500, 50, 42)
    ge = gpu.process_stream(bench)
    oe = orc.submit(bench["u"], bench["v"], bench["type"])
    assert ge == oe
    _same_state(gpu, orc, "bench stream")
    assert np.array_equal(gpu.snapshot(), orc.snapshot())
    gs, os_ = gpu.finish(), orc.finish()
    assert gs == os_
    _same_state(gpu, orc, "bench repair")
    assert np.array_equal(gpu.snapshot(), orc.snapshot())


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("threshold", [0, 4, 12])
def test_deterministic_histogram_path_bit_exact(threshold, exact, graph_path):
    for name in ["er2k", "delaunay1k", "rmat_s10"]:
        g = S.load_graph_csr(graph_path(name))
        gpu, orc, og = _pair(g, deterministic=True, epoch_size=200, batch_size=200, histogram_threshold=threshold,
                             exact_histogram=exact, repair_every_epoch=exact)
        init = S.generate_init_stream(g)
        assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
        assert gpu.finish() == orc.finish()
        bench = S.generate_bench_stream_edges(g, 2000, 60, 7)
        assert gpu.process_stream(bench) == orc.submit(bench["u"], bench["v"], bench["type"])
        assert gpu.finish() == orc.finish()
        _same_state(gpu, orc, f"{name} threshold {threshold}")


@pytest.mark.parametrize("sweeps,t2", [("tracked", "device"), ("tracked", "host"), ("full", "host")])
@pytest.mark.parametrize("repair", [False, True])
@pytest.mark.parametrize("slack,min_cap,cap", [(0, 1, 8), (0, 2, 4096), (50, 2, 64)])
def test_deterministic_overflow_and_tier2_bit_exact(slack, min_cap, cap, repair, sweeps, t2, graph_path):
    # overflowed edges leave the two rows disagreeing (sticky endpoints of the
    # dirty tracking); both sweep modes, and Tier-2 on the device or on the
    # host, must match the oracle's full sweeps and host Tier-2
    g = S.load_graph_csr(graph_path("er2k"))
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=128, batch_size=128, slack_pct=slack,
                         min_row_capacity=min_cap, overflow_capacity=cap, repair_every_epoch=repair,
                         full_sweeps=(sweeps == "full"), host_tier2=(t2 == "host"))
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    bench = S.generate_bench_stream_edges(g, 3000, 80, 5)
    ge = gpu.process_stream(bench)
    oe = orc.submit(bench["u"], bench["v"], bench["type"])
    assert ge == oe
    assert any(e["overflow_count"] for e in oe)
    fs = gpu.finish()
    assert fs == orc.finish()
    _same_state(gpu, orc, "overflow")


def test_find_color_vectors_on_device():
    # A/test/unit/test_color_select.cu vectors, on the device's selection code
    g = S.csr_from_edges(4, [(0, 1), (0, 2), (0, 3)])
    p = S.GPUPartition(g)
    for bm, want in [(0, 1), (0x1, 2), (0b10111, 4), (0b11, 3)]:
        b = p.bitmaps()
        b[0] = bm
        p.set_bitmaps(b)
        assert p.find_color(0) == want
    w = p.words()
    w[1], w[2], w[3] = 1, 2, 4
    p.set_words(w)
    b = p.bitmaps()
    b[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    p.set_bitmaps(b)
    assert p.find_color(0) == 3
    h = S.csr_from_edges(2, [(0, 1)])
    q = S.GPUPartition(h, S.StreamGCConfig(histogram_threshold=0))
    hist = q.hist()
    assert hist.shape == (2, 1024)
    for bins, want in [({}, 1), ({0: 3, 1: 2, 2: 1, 3: 5, 4: 4}, 6), ({0: 2, 1: 1, 2: 0, 3: 3}, 3)]:
        hist[:] = 0
        for k, c in bins.items():
            hist[0, k] = c
        q.set_hist(hist)
        assert q.find_color(0) == want
    hist[0, :] = 5
    hist[0, 700] = 2
    hist[0, 900] = 2
    q.set_hist(hist)
    assert q.find_color(0) == 701


def test_find_color_saturated_row_returns_256():
    n = 257
    g = S.csr_from_edges(n, [(0, i) for i in range(1, n)])
    p = S.GPUPartition(g, S.StreamGCConfig(histogram_threshold=1 << 30))
    w = p.words()
    w[1:256] = np.arange(1, 256)
    w[256] = 300
    p.set_words(w)
    b = p.bitmaps()
    b[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    p.set_bitmaps(b)
    assert p.find_color(0) == 256


def _proper_against_live(og, colors, u=None, v=None, t=None):
    if u is None:
        lu, lv = og.edges()
    else:
        lu, lv = O.live_edges_after(og, u, v, t)
    return O.validate_edges(og.n, lu, lv, colors)


@pytest.mark.parametrize("t1", ["parallel", "serial"])
@pytest.mark.parametrize("repair", [False, True])
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("name", ["er2k", "chunglu4k", "rmat_s10", "delaunay1k"])
def test_parallel_repair_is_proper_and_close(name, exact, repair, t1, graph_path):
    # batches of 64 on graphs of 1K-4K vertices: 3-12% of the vertices are
    # endpoints of one batch, 2-8x the endpoint density of a 32K batch on the
    # 4M-vertex power-law graph
    g = S.load_graph_csr(graph_path(name))
    gpu, orc, og = _pair(g, epoch_size=256, batch_size=64, exact_histogram=exact, repair_every_epoch=repair)
    # small graphs repair in vertex order by default; "parallel" forces the racy sweep
    gpu.set_tuning(serial_repair=0 if t1 == "parallel" else 1 << 20,
                   serial_conflicts=0 if t1 == "parallel" else 0xFFFFFFFF)
    init = S.generate_init_stream(g)
    gpu.process_stream(init)
    orc.submit(init["u"], init["v"], init["type"])
    gs, os_ = gpu.finish(), orc.finish()
    assert _proper_against_live(og, gpu.colors()) == (0, 0)
    bench = S.generate_bench_stream_edges(g, 20000, 50, 42)
    gpu.process_stream(bench)
    orc.submit(bench["u"], bench["v"], bench["type"])
    gs, os_ = gpu.finish(), orc.finish()
    assert gs["t2_remaining"] == 0
    assert _proper_against_live(og, gpu.colors(), bench["u"], bench["v"], bench["type"]) == (0, 0)
    assert _proper_against_live(og, gpu.snapshot(), bench["u"], bench["v"], bench["type"]) == (0, 0)
    # Concurrent updates colour adjacent vertices at the same time, and the
    # conflicts they leave cost colours that the sequential replay does not
    # spend; at this density up to ~20%.  This bound only catches gross
    # regressions.
    assert gs["max_color"] <= os_["max_color"] * 1.25 + 2, (gs["max_color"], os_["max_color"])
    assert gpu.validate() == (0, 0, gs["max_color"])


@pytest.mark.parametrize("deterministic", [True, False])
def test_every_published_snapshot_is_proper(deterministic, graph_path):
    g = S.load_graph_csr(graph_path("chunglu4k"))
    cfg = dict(epoch_size=512, batch_size=512, repair_every_epoch=True, deterministic=deterministic)
    gpu, orc, og = _pair(g, **cfg)
    init = S.generate_init_stream(g)
    gpu.process_stream(init)
    gpu.finish()
    bench = S.generate_bench_stream_edges(g, 512 * 12, 50, 3)
    gpu.history_reserve(16)
    epochs = gpu.process_stream(bench)
    assert len(epochs) == 12 and all(e["snapshot_published"] and e["t2_remaining"] == 0 for e in epochs)
    iv = S.edge_intervals(g, bench)
    pos = [512 * (k + 1) for k in range(12)]
    bad = gpu.validate_history(iv, pos)
    assert bad.tolist() == [0] * 12
    # the validator is not vacuous: an all-zero colouring of the same history fails it
    assert gpu.history_count() == 12


def test_validator_detects_conflicts(graph_path):
    g = S.load_graph_csr(graph_path("er2k"))
    p = S.GPUPartition(g, S.StreamGCConfig(repair_every_epoch=True, epoch_size=64, batch_size=64))
    p.history_reserve(2)
    p.set_words(np.full(g.n, 1, np.uint32))     # every vertex colour 1: all edges conflict
    p.epoch()                                   # repair fixes the live colouring, snapshot published
    viol, unc, mx = p.validate()
    assert viol == 0
    p2 = S.GPUPartition(g, S.StreamGCConfig())
    p2.history_reserve(1)
    p2.set_words(np.full(g.n, 1, np.uint32))
    p2.epoch()                                  # reference order: snapshot first, no repair
    iv = S.edge_intervals(g, S.make_updates([], []))
    assert p2.validate_history(iv, [0])[0] == g.num_edges // 2
    assert p2.validate(snapshot=True)[0] == g.num_edges


def test_query_reads_published_snapshot(graph_path):
    g = S.load_graph_csr(graph_path("delaunay1k"))
    p = S.GPUPartition(g, S.StreamGCConfig(epoch_size=256, batch_size=256, repair_every_epoch=True))
    p.process_stream(S.generate_init_stream(g))
    p.finish()
    ids = np.arange(g.n, dtype=np.uint32)[::-1].copy()
    assert np.array_equal(p.query(ids), p.snapshot()[ids])
    p.process_stream(S.generate_bench_stream_edges(g, 300, 50, 1))   # one epoch publishes
    snap = p.snapshot()
    assert np.array_equal(p.query(ids), snap[ids])
    assert p.query(np.array([g.n + 5], np.uint32)).tolist() == [0]


def test_apply_then_epoch_equals_submit(graph_path):
    g = S.load_graph_csr(graph_path("rmat_s10"))
    a, _, _ = _pair(g, deterministic=True, epoch_size=300, batch_size=300)
    b, _, _ = _pair(g, deterministic=True, epoch_size=300, batch_size=300)
    init = S.generate_init_stream(g)
    a.process_stream(init)
    for k in range(0, len(init), 300):
        b.apply(init[k:k + 300])
        if k + 300 <= len(init):
            b.epoch()
    assert np.array_equal(a.words(), b.words())


def test_device_resident_stream_equals_host_stream(graph_path):
    g = S.load_graph_csr(graph_path("chunglu4k"))
    a, _, _ = _pair(g, deterministic=True, epoch_size=1000, batch_size=1000)
    b, _, _ = _pair(g, deterministic=True, epoch_size=1000, batch_size=1000)
    init = S.generate_init_stream(g)
    a.process_stream(init)
    ds = S.DeviceStream(init)
    b.process_stream(ds)
    assert np.array_equal(a.words(), b.words())
    assert np.array_equal(a.csr()[1], b.csr()[1])


@pytest.mark.parametrize("name", ["delaunay1k", "rmat_s10", "chunglu4k"])
def test_switching_sweep_mode_mid_stream_stays_bit_exact(name, graph_path):
    g = S.load_graph_csr(graph_path(name))
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=100, batch_size=100, repair_every_epoch=True,
                         exact_histogram=True)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    bench = S.generate_bench_stream_edges(g, 2400, 60, 11)
    parts = np.array_split(np.arange(len(bench)), 4)
    for k, idx in enumerate(parts):
        gpu.set_mode(full_sweeps=(k % 2 == 1))
        chunk = bench[idx]
        assert gpu.process_stream(chunk) == orc.submit(chunk["u"], chunk["v"], chunk["type"])
        _same_state(gpu, orc, f"{name} part {k}")
    assert gpu.finish() == orc.finish()
    assert np.array_equal(gpu.snapshot(), orc.snapshot())


def test_tracked_sweeps_visit_fewer_rows(graph_path):
    g = S.load_graph_csr(graph_path("delaunay1k"))
    rows = {}
    for mode in ("tracked", "full"):
        p = S.GPUPartition(g, S.StreamGCConfig(epoch_size=64, batch_size=64, repair_every_epoch=True,
                                                full_sweeps=(mode == "full")))
        init = S.generate_init_stream(g)
        p.process_stream(init)
        p.finish()
        before = p.totals()
        p.process_stream(S.generate_bench_stream_edges(g, 640, 40, 3))
        after = p.totals()
        rows[mode] = {k: after[k] - before[k] for k in after if k.startswith("rows_")}
        violations, uncoloured, _ = p.validate()
        assert violations == 0 and uncoloured == 0
    full, tracked = rows["full"], rows["tracked"]
    for r in (full, tracked):   # Tier-1 rows of both kinds of sweep
        r["rows_tier1"] += r.pop("rows_tier1_serial")
    assert full["rows_compact"] >= 10 * g.n and full["rows_tier1"] >= 10 * g.n
    assert tracked["rows_tier1"] < full["rows_tier1"] / 2
    assert tracked["rows_compact"] < full["rows_compact"] / 2
    # a clean Tier-1 sweep on rows without overflow proves the colouring
    # proper, so the Tier-2 check is skipped in both modes
    assert full["rows_t2_check"] == 0 and tracked["rows_t2_check"] == 0


@pytest.mark.parametrize("overflow", [False, True])
@pytest.mark.parametrize("name", ["er2k", "delaunay1k", "rmat_s10", "chunglu4k"])
def test_long_rows_swept_by_blocks_stay_bit_exact(name, overflow, graph_path):
    # long_row=16 sends most rows of these graphs down the block-per-row
    # paths of compaction and rebuild; rows that overflow keep the Tier-2
    # check running (a clean Tier-1 proves nothing once an edge sits in one
    # row only), which exercises its block path too
    g = S.load_graph_csr(graph_path(name))
    extra = dict(slack_pct=0, min_row_capacity=2, overflow_capacity=4096) if overflow else {}
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=120, batch_size=120, repair_every_epoch=True,
                         exact_histogram=True, histogram_threshold=8, **extra)
    gpu.set_tuning(long_row=16, sweep_blocks=7)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    bench = S.generate_bench_stream_edges(g, 2400, 50, 13)
    assert gpu.process_stream(bench) == orc.submit(bench["u"], bench["v"], bench["type"])
    _same_state(gpu, orc, f"{name} long rows")
    assert gpu.finish() == orc.finish()
    assert np.array_equal(gpu.snapshot(), orc.snapshot())
    t = gpu.totals()
    assert t["rows_compact"] > 0
    if overflow:
        assert gpu.overflow_count() >= 0 and t["rows_t2_check"] > 0


@pytest.mark.parametrize("long_row", [16, 1 << 30])
def test_parallel_long_row_repair_is_proper(long_row, graph_path):
    g = S.load_graph_csr(graph_path("chunglu4k"))
    gpu, orc, og = _pair(g, epoch_size=256, batch_size=256, exact_histogram=True, repair_every_epoch=True,
                         histogram_threshold=8)
    gpu.set_tuning(long_row=long_row, serial_repair=0)
    init = S.generate_init_stream(g)
    gpu.process_stream(init)
    gpu.finish()
    bench = S.generate_bench_stream_edges(g, 256 * 20, 50, 9)
    gpu.history_reserve(24)
    epochs = gpu.process_stream(bench)
    assert all(e["t2_remaining"] == 0 for e in epochs)
    iv = S.edge_intervals(g, bench)
    bad = gpu.validate_history(iv, [256 * (k + 1) for k in range(gpu.history_count())])
    assert int(bad.sum()) == 0
    assert gpu.validate()[:2] == (0, 0)


@pytest.mark.parametrize("fresh", [0, 1, 2])
@pytest.mark.parametrize("name", ["rmat_s10", "chunglu4k"])
def test_parallel_selection_modes_repair_to_proper(name, fresh, graph_path):
    # fresh 0: the reference's selection (bitmaps rebuilt before every sweep);
    # 1 / 2: exact neighbour colours for bitmap rows / all rows
    g = S.load_graph_csr(graph_path(name))
    gpu, orc, og = _pair(g, epoch_size=256, batch_size=128, exact_histogram=True, repair_every_epoch=True,
                         histogram_threshold=16)
    gpu.set_tuning(fresh_repair=fresh, serial_repair=0)
    init = S.generate_init_stream(g)
    gpu.process_stream(init)
    gpu.finish()
    bench = S.generate_bench_stream_edges(g, 256 * 16, 50, 21)
    epochs = gpu.process_stream(bench)
    assert all(e["t2_remaining"] == 0 for e in epochs)
    assert _proper_against_live(og, gpu.snapshot(), bench["u"], bench["v"], bench["type"]) == (0, 0)
    orc.submit(init["u"], init["v"], init["type"])
    orc.finish()
    orc.submit(bench["u"], bench["v"], bench["type"])
    assert gpu.validate()[2] <= orc.finish()["max_color"] * 1.25 + 2


@pytest.mark.parametrize("kind", ["uniform", "holdout", "dense", "dense_growth"])
def test_other_stream_types_bit_exact(kind, graph_path):
    # the README's other stream types through the deterministic engine, against the oracle
    g = S.load_graph_csr(graph_path("rmat_s10" if kind.startswith("dense") else "er2k"))
    if kind == "holdout":
        g, stream = S.holdout_stream(g, 10, 5)
    elif kind == "uniform":
        stream = S.uniform_stream(g, 3000, seed=5)
    elif kind == "dense":
        stream = S.dense_stream(g, 3000, seed=5)
    else:
        stream = S.dense_growth_stream(g, 3000, seed=5)
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=128, batch_size=128, repair_every_epoch=True)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    assert gpu.process_stream(stream) == orc.submit(stream["u"], stream["v"], stream["type"])
    assert gpu.finish() == orc.finish()
    _same_state(gpu, orc, kind)
    assert gpu.validate(snapshot=True)[:2] == (0, 0)


@pytest.mark.parametrize("repair", [False, True])
@pytest.mark.parametrize("ablation", ["force_bitmap", "force_histogram", "no_overflow", "live_degree"])
def test_ablations_bit_exact(ablation, repair, graph_path):
    # the reference's compile-time ablations as runtime flags, against the oracle's
    overflow = dict(slack_pct=0, min_row_capacity=2, overflow_capacity=64) if ablation == "no_overflow" else {}
    g = S.load_graph_csr(graph_path("er2k" if overflow else "chunglu4k"))
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=128, batch_size=128, repair_every_epoch=repair,
                         exact_histogram=repair, ablations=(ablation,), **overflow)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    bench = S.generate_bench_stream_edges(g, 2000, 70 if overflow else 50, 5)
    assert gpu.process_stream(bench) == orc.submit(bench["u"], bench["v"], bench["type"])
    assert gpu.finish() == orc.finish()
    _same_state(gpu, orc, ablation)
    if overflow:
        assert gpu.overflow_count() == 0 and gpu.totals()["overflow_events"] > 0


def test_parallel_same_edge_groups_keep_the_oracles_slots(graph_path):
    # One parallel launch applies each edge's updates in stream order on one
    # wavefront; a delete followed by the same edge's re-insert claims the
    # slot the delete's scan found first free.  With pairwise disjoint edges
    # every row's result is independent of the wavefront order, so the rows
    # must equal the oracle's slot for slot (holes and fill included).
    g = S.load_graph_csr(graph_path("chunglu4k"))
    gpu, orc, og = _pair(g, epoch_size=1 << 20, batch_size=1 << 20)
    rp, col = g.row_ptr, g.col_idx
    hub = int(np.argmax(g.degree))
    nbrs = [int(x) for x in col[rp[hub]:rp[hub + 1]]]
    assert len(nbrs) >= 40
    # holes early in the hub's row
    holes = S.make_updates([hub] * 3, [nbrs[2], nbrs[5], nbrs[9]], [1, 1, 1])
    gpu.apply(holes)
    orc.submit(holes["u"], holes["v"], holes["type"])
    used = {hub, nbrs[2], nbrs[5], nbrs[9]}
    # groups: the hub's edge deep in its row (re-insert goes to the first hole),
    # other edges deleted and re-inserted, an edge inserted then deleted, and
    # an edge deleted, re-inserted and deleted again
    groups = [[(hub, nbrs[30], 1), (hub, nbrs[30], 0)]]
    used |= {nbrs[30]}
    rng = np.random.default_rng(7)
    order = rng.permutation(g.n)
    for u in order.tolist():
        if len(groups) >= 60:
            break
        row = [int(x) for x in col[rp[u]:rp[u + 1]]]
        free = [w for w in row if w not in used and w != u]
        if u in used or not free:
            continue
        w = free[len(free) // 2]
        kind = len(groups) % 3
        if kind == 0:
            groups.append([(u, w, 1), (u, w, 0)])
        elif kind == 1:
            groups.append([(u, w, 1), (u, w, 0), (u, w, 1)])
        else:
            x = next((y for y in order.tolist() if y not in used and y not in (u, w) and y not in row), None)
            if x is None:
                continue
            groups.append([(u, x, 0), (u, x, 1)])
            used.add(x)
        used |= {u, w}
    # interleave the groups at random, keeping each group's order
    owner = np.concatenate([np.full(len(q), k) for k, q in enumerate(groups)])
    rng.shuffle(owner)
    pos = [0] * len(groups)
    recs = []
    for k in owner.tolist():
        recs.append(groups[k][pos[k]])
        pos[k] += 1
    u, v, t = zip(*recs)
    batch = S.make_updates(list(u), list(v), list(t))
    gpu.apply(batch)
    orc.submit(batch["u"], batch["v"], batch["type"])
    grp, gcol = gpu.csr()
    orp, ocol = orc.csr()
    assert np.array_equal(grp, orp)
    assert np.array_equal(gcol, ocol), np.nonzero(gcol != ocol)[0][:8]


@pytest.mark.parametrize("name", ["er2k", "chunglu4k", "rmat_s10"])
@pytest.mark.parametrize("repair", [False, True])
def test_duplicate_inserts_bit_exact(name, repair, graph_path):
    """Inserts of edges that are already live, after deletes have left holes
    in front of their copies: the reference's first-free-or-equal scan stores
    a second copy (A/src/kernels/local_update.cu:12-40), and later deletes
    remove one copy per row, so rows disagree.  The stream generators never
    do this; here it is done on purpose, and the deterministic engine must
    stay bit-exact with the oracle through it (words, every slot, bitmaps,
    histograms, snapshots, epoch statistics)."""
    g = S.load_graph_csr(graph_path(name))
    gpu, orc, og = _pair(g, deterministic=True, epoch_size=256, batch_size=256, repair_every_epoch=repair)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    assert gpu.finish() == orc.finish()
    churn = S.generate_bench_stream_edges(g, 1500, 40, 5)     # deletes leave holes
    assert gpu.process_stream(churn) == orc.submit(churn["u"], churn["v"], churn["type"])
    rng = np.random.default_rng(11)
    lu, lv = O.live_edges_after(og, np.concatenate([init["u"], churn["u"]]),
                                np.concatenate([init["v"], churn["v"]]),
                                np.concatenate([init["type"], churn["type"]]))
    pick = rng.choice(len(lu), size=min(300, len(lu)), replace=False)
    swap = rng.random(len(pick)) < 0.5                          # either endpoint first
    du = np.where(swap, lv[pick], lu[pick]).astype(np.uint32)
    dv = np.where(swap, lu[pick], lv[pick]).astype(np.uint32)
    dup = S.make_updates(du, dv, np.zeros(len(pick), np.uint8))   # second copies
    more = S.generate_bench_stream_edges(g, 800, 50, 6)
    gone = S.make_updates(du[::2], dv[::2], np.ones(len(du[::2]), np.uint8))   # one copy of half of them
    for k, part in enumerate((dup, more, gone)):
        assert gpu.process_stream(part) == orc.submit(part["u"], part["v"], part["type"]), (name, k)
        _same_state(gpu, orc, f"{name} duplicates part {k}")
    assert gpu.finish() == orc.finish()
    assert np.array_equal(gpu.snapshot(), orc.snapshot())
