"""The served pattern: a persistent request loop over a pool (through the C ABI).

Each tenant's result must be the same alone and in the mix: bit-exact in
deterministic mode (one wavefront per graph replays its updates in stream
order, and the loop cuts every graph's updates at its own epoch boundaries);
in parallel mode every published colouring is proper.
"""
import threading

import numpy as np
import pytest

import streamgc_amd as S

pytestmark = pytest.mark.gpu

TENANTS = ["er2k", "chunglu4k", "rmat_s10", "delaunay1k", "c100", "star6"]


def _streams(g, seed):
    init = S.generate_init_stream(g)
    churn = S.generate_bench_stream_edges(g, 700 + 37 * seed, 50, seed)
    return np.concatenate([init, churn])


def _requests(streams, rng):
    """Interleave per-tenant streams into requests of random size, each
    tenant's own order kept."""
    heads = {t: 0 for t in streams}
    out = []
    while heads:
        t = rng.choice(sorted(heads))
        s = streams[t]
        k = int(rng.integers(1, 300))
        out.append((t, s[heads[t]:heads[t] + k]))
        heads[t] += k
        if heads[t] >= len(s):
            del heads[t]
    return out


@pytest.mark.parametrize("max_batch,mode", [(64, S.GraphServer.BATCHED), (4096, S.GraphServer.BATCHED),
                                            (4096, S.GraphServer.PLAIN)])
def test_served_mix_equals_each_graph_alone(max_batch, mode, graph_path):
    cfg = S.StreamGCConfig(epoch_size=160, batch_size=160, deterministic=True, repair_every_epoch=True,
                           exact_histogram=True)
    graphs = {name: S.load_graph_csr(graph_path(name)) for name in TENANTS}
    pool = S.GraphPool(max_graphs=8, batch_capacity=8192, deterministic=True, repair_every_epoch=True,
                       exact_histogram=True)
    srv = S.GraphServer(pool, max_batch=max_batch, mode=mode)
    gid = {name: srv.add(g, cfg) for name, g in graphs.items()}
    streams = {name: _streams(g, k + 1) for k, (name, g) in enumerate(graphs.items())}
    tickets = [srv.submit(gid[t], recs) for t, recs in _requests(streams, np.random.default_rng(5))]
    srv.flush()
    for tk in tickets:
        srv.wait(tk, 1000)
    st = srv.stats()
    assert st["requests_done"] == len(tickets) and st["queued_updates"] == 0
    assert st["updates_applied"] == sum(len(s) for s in streams.values())
    if mode == S.GraphServer.PLAIN:              # one launch per request (more where one crosses an epoch)
        assert st["batches"] >= len(tickets) and st["epoch_steps"] == st["graph_epochs"]
    for name, g in graphs.items():
        alone = S.GPUPartition(g, cfg)
        alone.process_stream(streams[name])
        if len(streams[name]) % cfg.epoch_size:
            alone.epoch()                       # the flush publishes the partial last epoch
        assert np.array_equal(pool.words(gid[name]), alone.words()), name
        assert np.array_equal(pool.snapshot(gid[name]), alone.snapshot()), name
        ids = np.arange(g.n, dtype=np.uint32)
        assert np.array_equal(srv.query(gid[name], ids), alone.snapshot()), name
    lat = srv.latencies("update")
    assert len(lat) == len(tickets) and np.all(lat > 0)
    srv.close()


def test_served_parallel_producers_keep_every_tenant_proper(graph_path):
    cfg = S.StreamGCConfig(epoch_size=256, batch_size=256, repair_every_epoch=True, exact_histogram=True)
    graphs = {name: S.load_graph_csr(graph_path(name)) for name in TENANTS}
    pool = S.GraphPool(max_graphs=8, batch_capacity=16384, repair_every_epoch=True, exact_histogram=True)
    srv = S.GraphServer(pool, max_batch=16384, max_delay_us=3000)
    gid = {name: srv.add(g, cfg) for name, g in graphs.items()}
    streams = {name: _streams(g, k + 3) for k, (name, g) in enumerate(graphs.items())}
    errors, tickets = [], []

    def producer(names, seed):
        try:
            rng = np.random.default_rng(seed)
            for t, recs in _requests({n: streams[n] for n in names}, rng):
                tickets.append(srv.submit(gid[t], recs))
                if rng.random() < 0.1:
                    ids = rng.integers(0, graphs[t].n, 64).astype(np.uint32)
                    srv.query(gid[t], ids)
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=producer, args=(TENANTS[k::3], k)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for tk in tickets:
        srv.wait(tk, 60000)
    srv.flush()
    for name, g in graphs.items():
        viol, unc, _ = pool.validate(gid[name])
        assert (viol, unc) == (0, 0), name
        viol, unc, _ = pool.validate(gid[name], snapshot=True)
        assert (viol, unc) == (0, 0), name
    st = srv.stats()
    assert st["requests_done"] == len(tickets)
    assert st["batches"] < len(tickets)            # requests shared launches
    assert len(srv.latencies("query")) == st["queries"] > 0
    srv.close()


def test_served_admission_waits_for_memory(graph_path):
    g = S.load_graph_csr(graph_path("er2k"))
    cfg = S.StreamGCConfig(epoch_size=128, batch_size=128, repair_every_epoch=True)
    need = S.GraphPool.graph_bytes(g, cfg)
    pool = S.GraphPool(max_graphs=4, batch_capacity=4096, repair_every_epoch=True, hbm_budget=2 * need + need // 2)
    srv = S.GraphServer(pool, max_batch=4096)
    a = srv.add(g, cfg)
    b = srv.add(g, cfg)
    t3 = srv.admit(g, cfg)
    with pytest.raises(S.SgcError):
        srv.admitted(t3, timeout_ms=300)           # over budget: still waiting
    used, budget, n = pool.usage()
    assert n == 2 and used <= budget
    tk = srv.submit(a, S.generate_init_stream(g)[:300])
    srv.remove(b)
    c = srv.admitted(t3, timeout_ms=30000)         # admitted once b's memory is free
    srv.flush()                                    # the last 44 updates fill no epoch: publish them now
    srv.wait(tk, 30000)
    used, budget, n = pool.usage()
    assert n == 2 and used <= budget and c == b
    srv.close()


def test_served_max_delay_publishes_partial_epochs(graph_path):
    g = S.load_graph_csr(graph_path("delaunay1k"))
    cfg = S.StreamGCConfig(epoch_size=100000, batch_size=4096, repair_every_epoch=True)
    pool = S.GraphPool(max_graphs=2, batch_capacity=4096, repair_every_epoch=True)
    srv = S.GraphServer(pool, max_delay_us=2000)
    gid = srv.add(g, cfg)
    init = S.generate_init_stream(g)
    for k in range(0, 600, 50):
        srv.wait(srv.submit(gid, init[k:k + 50]), 10000)   # far below one epoch, still published
    lat = srv.latencies("update")
    assert len(lat) == 12 and np.max(lat) < 2e6
    assert srv.stats()["graph_epochs"] >= 1
    srv.close()


@pytest.mark.parametrize("mode", [S.GraphServer.BATCHED, S.GraphServer.PLAIN])
def test_replayed_trace_equals_each_graph_alone(mode, graph_path):
    """An open-loop trace replayed by the library (updates at their scheduled
    times, queries from a second thread) gives every tenant the same result as
    alone, and one latency per request and per query."""
    cfg = S.StreamGCConfig(epoch_size=160, batch_size=160, deterministic=True, repair_every_epoch=True,
                           exact_histogram=True)
    graphs = {name: S.load_graph_csr(graph_path(name)) for name in TENANTS}
    pool = S.GraphPool(max_graphs=8, batch_capacity=8192, deterministic=True, repair_every_epoch=True,
                       exact_histogram=True)
    srv = S.GraphServer(pool, max_batch=4096, mode=mode)
    gid = {name: srv.add(g, cfg) for name, g in graphs.items()}
    streams = {name: _streams(g, k + 11) for k, (name, g) in enumerate(graphs.items())}
    rng = np.random.default_rng(9)
    reqs = _requests(streams, rng)
    items = np.zeros(len(reqs) + len(reqs) // 5 + 1, dtype=S.TRACE_ITEM)
    parts, qids = [], []
    off = qoff = k = 0
    at = 0
    for j, (t, recs) in enumerate(reqs):
        at += int(rng.exponential(150e3))          # ns
        items[k] = (at, off, gid[t], len(recs), 0, 0)
        parts.append(recs)
        off += len(recs)
        k += 1
        if j % 5 == 4:
            ids = rng.integers(0, graphs[t].n, int(rng.integers(1, 200))).astype(np.uint32)
            items[k] = (at, qoff, gid[t], len(ids), 1, 0)
            qids.append(ids)
            qoff += len(ids)
            k += 1
    items = items[:k]
    rs = srv.replay(items, np.concatenate(parts), np.concatenate(qids), timeout_ms=120000)
    assert rs["requests"] == len(reqs) and rs["queries"] == len(qids)
    assert rs["updates"] == sum(len(s) for s in streams.values())
    assert rs["span_us"] * 1e3 >= at
    lat = srv.latencies("update")
    assert len(lat) == len(reqs) and np.all(lat > 0)
    assert len(srv.latencies("query")) == len(qids)
    for name, g in graphs.items():
        alone = S.GPUPartition(g, cfg)
        alone.process_stream(streams[name])
        if len(streams[name]) % cfg.epoch_size:
            alone.epoch()
        assert np.array_equal(pool.words(gid[name]), alone.words()), name
        assert np.array_equal(pool.snapshot(gid[name]), alone.snapshot()), name
    srv.close()
