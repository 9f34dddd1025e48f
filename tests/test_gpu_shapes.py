"""The batched (C4) and served (C5) workloads at their configured shapes
(streamgc_amd/workloads.py, the shapes bench.py measures).

C4: 256 graphs (Delaunay 2^10-2^16 points, R-MAT scale 10-16, edge factor
8-40), each with its own churn of 1K-64K updates at 30-70% insert and E_g =
min(stream, 4096), replayed as captured rounds with the autotuner.  Every
published snapshot of every graph is checked against the live edge set its
stream implies at that point.  A graph whose rows never lost an edge (the
engine's count of half-edges lost to a full overflow buffer or a failed
drain, sgc_graph_info.lost) must have no violation at all; a graph that lost
some may show violations only on edges it no longer holds, so its count is
bounded by the half-edges it lost, and both are reported.  A sample of the
graphs also runs deterministically and must equal the oracle alone.

C5: the 64 tenants plus the com-Orkut-shape graph resident in one pool,
served from an open-loop replayed trace of U[1, 8192]-update requests with
queries; every published snapshot of every tenant is checked against the
tenant's stream at the snapshot's watermark.
"""
import numpy as np
import pytest

import oracle_lib as O
import streamgc_amd as S
from streamgc_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    spec = W.batch_graphs(S, 256)
    churn = [S.generate_bench_stream_edges(g, L, pct, 42 + i) for i, (_, g, L, pct) in enumerate(spec)]
    return spec, churn


def test_c4_256_graphs_every_snapshot_validated(c4):
    spec, churn = c4
    E = W.batch_epoch_sizes(spec)
    pool = S.GraphPool(max_graphs=len(spec), batch_capacity=1 << 20, repair_every_epoch=True, exact_histogram=True)
    gids = [pool.add(g, S.StreamGCConfig(epoch_size=e, batch_size=e)) for (_, g, _, _), e in zip(spec, E)]
    init = [S.tag_graph(S.generate_init_stream(g), gid) for (_, g, _, _), gid in zip(spec, gids)]
    for launch in W.chunked_launches(init, 256):
        pool.apply(launch)
    pool.finish()
    epochs = [-(-L // e) for (_, _, L, _), e in zip(spec, E)]
    for gid, k in zip(gids, epochs):
        pool.history_reserve(gid, k)
    mixed = np.concatenate([S.tag_graph(x, gid) for x, gid in zip(churn, gids)])
    rounds = pool.batch_prepare(mixed, use_graph=True, autotune_rounds=1)
    st = pool.batch_run()
    pool.batch_release()
    assert st["rounds"] == rounds == max(epochs)
    # the autotuner timed every candidate on a round that ran captured to the end
    assert all(n >= 1 for n in st["candidate_samples"]) and all(v > 0 for v in st["candidate_ns_per_update"]), st
    clean = lossy = snaps = 0
    viol_lossy = lost_total = 0
    for (kind, g, L, _), gid, x, k, e in zip(spec, gids, churn, epochs, E):
        assert pool.history_count(gid) == k
        assert pool.validate(gid)[:2] == (0, 0)                 # proper on the rows it holds
        bad = pool.validate_history(gid, S.edge_intervals(g, x), [min(e * (j + 1), L) for j in range(k)])
        snaps += k
        lost = pool.info(gid)["lost"]
        if lost == 0:
            clean += 1
            assert bad.tolist() == [0] * k, (gid, kind, g.n)
        else:
            lossy += 1
            lost_total += lost
            viol_lossy += int(bad.sum())
            assert int(bad.max()) <= lost, (gid, kind, g.n, bad.tolist(), lost)
    print(f"C4: {len(spec)} graphs, {rounds} rounds ({st['host_rounds']} finished by the host), {snaps} snapshots "
          f"validated; {clean} graphs lost no edge (0 violations), {lossy} lost {lost_total} half-edges to full rows "
          f"and show {viol_lossy} violating edge-snapshots on them; {st['ms']:.1f} ms")
    assert clean > len(spec) // 2
    pool.close()


def test_c4_sample_deterministic_equals_the_oracle_alone(c4):
    """Six of the C4 graphs (the three smallest of each kind) in one deterministic
    pool, captured rounds: each graph bit-exact with the oracle running it alone."""
    spec, churn = c4
    E = W.batch_epoch_sizes(spec)
    order = sorted(range(len(spec)), key=lambda i: spec[i][1].n)
    pick = [i for i in order if spec[i][0] == "delaunay"][:3] + [i for i in order if spec[i][0] == "rmat"][:3]
    pool = S.GraphPool(max_graphs=8, batch_capacity=1 << 18, deterministic=True, repair_every_epoch=True,
                       exact_histogram=True)
    cfg = {i: S.StreamGCConfig(epoch_size=E[i], batch_size=E[i]) for i in pick}
    gids = {i: pool.add(spec[i][1], cfg[i]) for i in pick}
    init = np.concatenate([S.tag_graph(S.generate_init_stream(spec[i][1]), gids[i]) for i in pick])
    pool.submit(init, max_stats=0)
    pool.finish()
    sub = {i: churn[i][:6000] for i in pick}
    mixed = np.concatenate([S.tag_graph(sub[i], gids[i]) for i in pick])
    pool.batch_prepare(mixed, use_graph=True)
    pool.batch_run()
    pool.batch_release()
    for i in pick:
        g = spec[i][1]
        oc = O.default_config(epoch_size=E[i], batch_size=E[i], flags=O.FLAG_REPAIR_EVERY_EPOCH | O.FLAG_EXACT_HISTOGRAM)
        orc = O.Engine(O.Csr(g.n, g.row_ptr, g.col_idx, g.degree), oc)
        ini = S.generate_init_stream(g)
        orc.submit(ini["u"], ini["v"], ini["type"])
        orc.finish()
        orc.submit(sub[i]["u"], sub[i]["v"], sub[i]["type"])
        orc.finish()
        assert np.array_equal(pool.words(gids[i]), orc.words()), (i, spec[i][0], g.n)
        assert np.array_equal(pool.csr(gids[i])[1], orc.csr()[1]), (i, spec[i][0], g.n)
        assert np.array_equal(pool.snapshot(gids[i]), orc.snapshot()), (i, spec[i][0], g.n)
    pool.close()


def test_c5_tenants_and_orkut_served_every_snapshot_validated():
    tenants = W.tenants(S, 64)
    T = len(tenants)
    E = 4096
    cfg = S.StreamGCConfig(epoch_size=1 << 30, batch_size=1 << 16, exact_histogram=True, repair_every_epoch=True)
    pool = S.GraphPool(max_graphs=T, batch_capacity=1 << 16, repair_every_epoch=True, exact_histogram=True)
    gids = [pool.add(g, cfg) for _, g in tenants]
    for gid, (_, g) in zip(gids, tenants):
        pool.submit(S.tag_graph(S.generate_init_stream(g), gid), max_stats=0)
    pool.finish()
    for gid in gids:
        pool.history_reserve(gid, 128)
    rng = np.random.default_rng(3)
    nreq, rate = 1200, 20e6
    picks = rng.integers(0, T, nreq)
    sizes = rng.integers(1, 8193, nreq)
    need = np.bincount(picks, weights=sizes, minlength=T).astype(np.int64)
    streams = [S.generate_bench_stream_edges(g, int(need[k]) + 1, 50, 42 + k)[:int(need[k])]
               for k, (_, g) in enumerate(tenants)]
    head = np.zeros(T, np.int64)
    at = np.cumsum(rng.exponential(4096.5 / rate, nreq))
    items = np.zeros(nreq + nreq // 10, dtype=S.TRACE_ITEM)
    parts, qids = [], []
    off = qoff = n_items = 0
    for j, (k, n) in enumerate(zip(picks.tolist(), sizes.tolist())):
        parts.append(streams[k][head[k]:head[k] + n])
        head[k] += n
        items[n_items] = (int(at[j] * 1e9), off, gids[k], n, 0, 0)
        off += n
        n_items += 1
        if j % 10 == 9:
            ids = rng.integers(0, tenants[k][1].n, int(rng.integers(1, 1025))).astype(np.uint32)
            items[n_items] = (int(at[j] * 1e9), qoff, gids[k], len(ids), 1, 0)
            qids.append(ids)
            qoff += len(ids)
            n_items += 1
    srv = S.GraphServer(pool, max_batch=1 << 16, max_delay_us=2000)
    for gid in gids:
        srv.attach(gid, E)
    rs = srv.replay(items[:n_items], np.concatenate(parts), np.concatenate(qids))
    assert rs["requests"] == nreq and rs["queries"] == len(qids)
    lat = srv.latencies("update")
    assert len(lat) == nreq
    snaps = 0
    for k, (kind, g) in enumerate(tenants):
        wm = srv.watermarks(gids[k])
        hc = pool.history_count(gids[k])
        assert len(wm) == hc and (hc == 0 or wm[-1] == need[k]), (k, kind, len(wm), hc)
        assert pool.info(gids[k])["lost"] == 0, (k, kind)
        bad = pool.validate_history(gids[k], S.edge_intervals(g, streams[k]), wm[:hc])
        assert bad.tolist() == [0] * hc, (k, kind, bad.tolist())
        assert pool.validate(gids[k], snapshot=True)[:2] == (0, 0), (k, kind)
        snaps += hc
    srv.close()
    print(f"C5: {T} tenants, {nreq} requests ({int(need.sum())} updates) at {rate / 1e6:.0f}M/s offered, "
          f"{snaps} snapshots validated; p99 update-to-snapshot {np.percentile(lat, 99) / 1e3:.2f} ms")
    pool.close()
