"""Full-scale workloads against the oracle's numbers (tests/golden/scale_reference.json).

C1 (Delaunay 2^17, 100,000 churn updates): the deterministic replay is
bit-exact with the oracle -- every word, slot and epoch statistic -- and its
colour count is the reference's.  C2 (Chung-Lu com-LiveJournal shape) and C3
(R-MAT scale 20, dense growth inside the top k-cores): the parallel engine
publishes proper snapshots and its colour count stays close to the
sequential replay's (exact histograms).
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "scale_reference.json")))
E = 32768


def test_c1_deterministic_replay_is_bit_exact_at_full_scale():
    g = S.gen_delaunay(1 << 17, seed=42)
    ref = REF["c1/reference"]
    assert (g.n, g.num_edges // 2) == (ref["vertices"], ref["edges"])
    gpu = S.GPUPartition(g, S.StreamGCConfig(epoch_size=E, batch_size=E, deterministic=True))
    og = O.Csr(g.n, g.row_ptr, g.col_idx, g.degree)
    orc = O.Engine(og, O.default_config(), threads=os.cpu_count() or 1)
    init = S.generate_init_stream(g)
    assert gpu.process_stream(init) == orc.submit(init["u"], init["v"], init["type"])
    fin = gpu.finish()
    assert fin == orc.finish()
    assert fin["max_color"] == ref["init_max_color"]
    gpu.set_mode(repair_every_epoch=True)
    orc.set_flags(O.FLAG_REPAIR_EVERY_EPOCH)
    bench = S.generate_bench_stream_edges(g, ref["stream"], 50, 42)
    ge = gpu.process_stream(bench)
    oe = orc.submit(bench["u"], bench["v"], bench["type"])
    assert ge == oe
    assert [e["max_color"] for e in ge] == ref["epoch_max_color"]
    last = gpu.finish()
    assert last == orc.finish()
    assert last["max_color"] == ref["final_max_color"]
    assert np.array_equal(gpu.words(), orc.words())
    assert np.array_equal(gpu.csr()[1], orc.csr()[1])


def test_c2_parallel_colouring_is_proper_and_close_to_the_replay():
    g = S.gen_chung_lu(3997962, 34681189, 2.3, 15000.0, seed=42, permute=True)
    ref = REF["c2/exact"]
    assert (g.n, g.num_edges // 2) == (ref["vertices"], ref["edges"])
    gpu = S.GPUPartition(g, S.StreamGCConfig(epoch_size=E, batch_size=E, exact_histogram=True))
    init = S.generate_init_stream(g)
    ds = S.DeviceStream(init)
    for k in range(0, len(init), E):           # launches of 256, epochs every E (as bench.py)
        for j in range(k, min(k + E, len(init)), 256):
            gpu.apply_device(ds.at(j), min(256, len(init) - j))
        if k + E <= len(init):
            gpu.epoch()
    ds.close()
    fin = gpu.finish()
    assert fin["t2_remaining"] == 0
    gpu.set_mode(repair_every_epoch=True)
    bench = S.generate_bench_stream_edges(g, 8 * E, 50, 42)
    gpu.history_reserve(8)
    epochs = gpu.process_stream(bench)
    assert len(epochs) == 8 and all(e["t2_remaining"] == 0 for e in epochs)
    bad = gpu.validate_history(S.edge_intervals(g, bench), [E * (k + 1) for k in range(8)])
    assert bad.tolist() == [0] * 8
    # Epoch k's colour count against the sequential replay's after the same
    # epoch, within 2% or two colours, whichever is more.  The replay selects
    # from exact histograms, as this engine does; the reference's own
    # semantics (drifting histograms) uses far more colours on the same stream
    # (REF["c2/reference"]).  The two colours: the init phase's first
    # colourings race in launches of 256, and its maximum alone has ranged
    # 79-85 over 23 runs of this workload (83 for the replay); eight epochs of
    # churn seldom touch the few vertices holding the maximum.
    got = [e["max_color"] for e in epochs]
    exact = ref["epoch_max_color"][:8]
    drift = REF["c2/reference"]["epoch_max_color"][:8]
    print(f"C2 colours per epoch: GPU {got}, replay (exact histograms) {exact}, replay (reference semantics) {drift}")
    for k, (c, r) in enumerate(zip(got, exact)):
        assert c <= max(r * 1.02, r + 2), (k, got, exact)
        assert c < drift[k], (k, got, drift)


def _parallel_run(g, stream, ref, tol, tol_each):
    gpu = S.GPUPartition(g, S.StreamGCConfig(epoch_size=E, batch_size=E, exact_histogram=True))
    init = S.generate_init_stream(g)
    ds = S.DeviceStream(init)
    for k in range(0, len(init), E):           # launches of 256, epochs every E (as bench.py)
        for j in range(k, min(k + E, len(init)), 256):
            gpu.apply_device(ds.at(j), min(256, len(init) - j))
        if k + E <= len(init):
            gpu.epoch()
    ds.close()
    assert gpu.finish()["t2_remaining"] == 0
    gpu.set_mode(repair_every_epoch=True)
    epochs_n = len(stream) // E
    gpu.history_reserve(epochs_n)
    epochs = gpu.process_stream(stream[:epochs_n * E])
    assert len(epochs) == epochs_n and all(e["t2_remaining"] == 0 for e in epochs)
    bad = gpu.validate_history(S.edge_intervals(g, stream), [E * (k + 1) for k in range(epochs_n)])
    assert bad.tolist() == [0] * epochs_n
    got = [e["max_color"] for e in epochs]
    exact = ref["epoch_max_color"][:epochs_n]
    print(f"colours per epoch: GPU {got}, replay (exact histograms) {exact}")
    assert got[-1] <= exact[-1] * tol, (got, exact)
    for k, (c, r) in enumerate(zip(got, exact)):
        assert c <= r * tol_each, (k, got, exact)
    return got[-1]


def test_c3_dense_growth_is_proper_and_within_two_percent_of_the_replay():
    g = S.gen_rmat(20, 48.0, seed=42)
    ref = REF["c3/exact"]
    assert (g.n, g.num_edges // 2) == (ref["vertices"], ref["edges"])
    stream = S.dense_growth_stream(g, ref["stream"], core_pct=90, ins_per_del=2, seed=42)
    # The colour count after the last full epoch within 2% of the replay's
    # after the same epoch.  The early epochs carry the init phase's colouring,
    # whose concurrent first colourings leave a few more colours than the
    # sequential replay (0.5-2.2% over at epochs 0-2 across runs): each epoch
    # within 3%.
    final = _parallel_run(g, stream, ref, 1.02, 1.03)
    assert final < REF["c3/reference"]["epoch_max_color"][len(stream) // E - 1]
