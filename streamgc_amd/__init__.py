"""MI355X-native StreamGC engine: streaming graph colouring on gfx950.

Python mirror of the reference's host interface for the hot path, over the
C ABI of libstreamgc_amd.so (include/streamgc.h).  Names follow the
reference: StreamGCConfig (A/include/streamgc.cuh:113-135), EdgeUpdate
(A/include/types.h:8-21), GPUPartition (A/include/streamgc.cuh:26-110) with
the phases of main() (A/src/streamgc.cu:809-1053, 1243-1393, 1057-1240),
load_graph_csr (A/src/host/graph_loader.cpp:29), generate_init_stream and
generate_bench_stream_edges (A/src/host/stream_gen.cpp:30,65).
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import EDGE_UPDATE, TRACE_ITEM, SgcError, check, lib

__all__ = ["StreamGCConfig", "default_config", "CSRGraph", "GPUPartition", "DeviceStream", "EDGE_UPDATE", "TRACE_ITEM",
           "load_graph_csr", "csr_from_edges", "csr_from_edge_arrays", "gen_chung_lu", "gen_rmat",
           "gen_delaunay", "core_numbers", "dense_growth_stream", "dense_stream", "uniform_stream", "holdout_stream",
           "generate_init_stream", "generate_bench_stream_edges",
           "static_greedy_coloring", "check_coloring_valid", "edge_intervals", "make_updates", "SgcError",
           "SgcNoMemory", "device_count", "GraphPool", "GraphServer", "tag_graph", "write_coloring",
           "write_coloring_meta", "read_coloring", "microbench_trc"]

from ._lib import SgcNoMemory  # noqa: E402

UPDATE_ADD = 0
UPDATE_DEL = 1


@dataclass
class StreamGCConfig:
    """Engine configuration (the reference's defaults)."""
    epoch_size: int = 32768
    batch_size: int = 32768
    histogram_threshold: int = 64
    slack_pct: int = 300
    min_row_capacity: int = 8
    overflow_capacity: int = 1 << 18
    max_repair_iters: int = 20
    plateau_iters: int = 3
    tier2_passes: int = 30
    deterministic: bool = False
    repair_every_epoch: bool = False
    exact_histogram: bool = False
    full_sweeps: bool = False
    host_tier2: bool = False       # Tier-2 on the host, as the reference places it (same results)
    ablations: tuple = ()          # the reference's ablations: "force_bitmap", "force_histogram",
                                   # "no_overflow", "live_degree"

    def to_c(self):
        flags = (_lib.SGC_DETERMINISTIC if self.deterministic else 0) | \
                (_lib.SGC_REPAIR_EVERY_EPOCH if self.repair_every_epoch else 0) | \
                (_lib.SGC_EXACT_HISTOGRAM if self.exact_histogram else 0) | \
                (_lib.SGC_FULL_SWEEPS if self.full_sweeps else 0) | \
                (_lib.SGC_HOST_TIER2 if self.host_tier2 else 0)
        for a in self.ablations:
            flags |= _lib.ABLATIONS[a]
        return _lib.Config(self.epoch_size, self.batch_size, self.histogram_threshold, self.slack_pct,
                           self.min_row_capacity, self.overflow_capacity, self.max_repair_iters,
                           self.plateau_iters, self.tier2_passes, flags)


def make_tuning(**knobs):
    """sgc_tuning from the defaults with `knobs` overridden."""
    t = _lib.Tuning()
    lib().sgc_default_tuning(C.byref(t))
    for k, v in knobs.items():
        if k not in ("long_row", "sweep_blocks", "apply_blocks", "serial_repair", "serial_conflicts", "fresh_repair"):
            raise TypeError(f"unknown tuning knob {k}")
        setattr(t, k, int(v))
    return t


def default_config(**kw):
    return StreamGCConfig(**kw)


def device_count():
    return lib().sgc_device_count()


def microbench_trc(waves=1, shared=False, iters=1_000_000, runs=11, device=0):
    """T_rc (A/benchmark/microbench/t_rc_atomic.cu:18-62): dependent version-bumping CAS on a vertex word.

    waves wavefronts, lane 0 of each looping `iters` times on its own word
    (shared=False, a cache line apart) or on one shared word.  Returns the
    median kernel time per iteration in ns and the first and last
    wavefront's word after the last run."""
    ns = C.c_double()
    fin = (C.c_uint32 * 2)()
    check(lib().sgc_microbench_trc(device, waves, 1 if shared else 0, iters, runs, C.byref(ns), fin))
    return {"ns": ns.value, "final": (fin[0], fin[1])}


class CSRGraph:
    """Symmetric CSR with sorted rows (A/src/host/graph_loader.cpp:20-26)."""

    def __init__(self, num_vertices, row_ptr, col_idx, degree=None):
        self.num_vertices = int(num_vertices)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        if degree is None:
            degree = np.diff(self.row_ptr.astype(np.int64))
        self.degree = np.ascontiguousarray(degree, dtype=np.uint32)

    @property
    def n(self):
        return self.num_vertices

    @property
    def num_edges(self):
        """Directed entries (2 x undirected), as the reference counts them."""
        return int(self.row_ptr[-1])

    def edges(self):
        u = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.row_ptr.astype(np.int64)))
        m = u < self.col_idx
        return u[m], self.col_idx[m]


def csr_from_edges(n, edges):
    """CSR from an edge list with the loader's rules (dedup, no self-loops, symmetric, sorted rows)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if len(e):
        e = e[e[:, 0] != e[:, 1]]
        e = np.unique(np.sort(e, axis=1), axis=0)
    a = np.concatenate([e[:, 0], e[:, 1]])
    b = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    deg = np.bincount(a, minlength=n).astype(np.uint32)
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(deg, out=rp[1:])
    return CSRGraph(n, rp, b.astype(np.uint32), deg)


def _from_handle(h):
    L = lib()
    n, nnz, md = C.c_uint32(), C.c_uint64(), C.c_uint32()
    try:
        check(L.sgc_csr_info(h, C.byref(n), C.byref(nnz), C.byref(md)))
        rp = np.zeros(n.value + 1, np.uint32)
        col = np.zeros(max(nnz.value, 1), np.uint32)
        deg = np.zeros(max(n.value, 1), np.uint32)
        check(L.sgc_csr_export(h, rp.ctypes.data, col.ctypes.data, deg.ctypes.data))
    finally:
        L.sgc_csr_free(h)
    return CSRGraph(n.value, rp, col[:nnz.value], deg[:n.value])


def load_graph_csr(path):
    """= load_graph_csr (A/src/host/graph_loader.cpp:29-169)."""
    h = C.c_void_p()
    check(lib().sgc_csr_load_mtx(path.encode(), C.byref(h)))
    return _from_handle(h)


def csr_from_edge_arrays(n, u, v):
    """Same rules as csr_from_edges, in C++ (large edge lists)."""
    h = C.c_void_p()
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    check(lib().sgc_csr_from_edges(n, u, v, len(u), C.byref(h)))
    return _from_handle(h)


def gen_chung_lu(n, m, gamma=2.3, max_degree=15000.0, seed=42, permute=True):
    """Power-law Chung-Lu graph with exactly m distinct edges."""
    h = C.c_void_p()
    check(lib().sgc_gen_chung_lu(n, m, gamma, max_degree, seed, 1 if permute else 0, C.byref(h)))
    return _from_handle(h)


def gen_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=42):
    """R-MAT graph: edge_factor * 2^scale draws, deduplicated and symmetrised."""
    h = C.c_void_p()
    check(lib().sgc_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)))
    return _from_handle(h)


def gen_delaunay(num_points, seed=42):
    """Delaunay triangulation of num_points uniform points in the unit square."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.random((num_points, 2))
    tri = Delaunay(pts).simplices.astype(np.int64)
    u = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
    v = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
    return csr_from_edge_arrays(num_points, u.astype(np.uint32), v.astype(np.uint32))


def core_numbers(g):
    core = np.zeros(max(g.n, 1), np.uint32)
    check(lib().sgc_core_numbers(g.n, g.row_ptr, g.col_idx, core))
    return core[:g.n]


def dense_growth_stream(g, size, core_pct=90, ins_per_del=2, seed=42, core=None):
    """Dense-growth stream inside the vertices whose core number is at least the core_pct percentile."""
    core = core_numbers(g) if core is None else core
    core_min = int(np.percentile(core, core_pct, method="lower"))
    out = np.zeros(max(size, 1), dtype=EDGE_UPDATE)
    check(lib().sgc_dense_growth_stream(g.n, g.row_ptr, g.col_idx, np.ascontiguousarray(core, np.uint32),
                                        core_min, size, ins_per_del, seed, out.ctypes.data))
    return out[:size]


def dense_stream(g, size, core_pct=90, seed=42, core=None):
    """The README's dense stream: dense growth with one insert per delete."""
    return dense_growth_stream(g, size, core_pct=core_pct, ins_per_del=1, seed=seed, core=core)


def uniform_stream(g, size, seed=42):
    """The README's uniform stream: inserts of uniform non-adjacent vertex pairs, each new."""
    out = np.zeros(max(size, 1), dtype=EDGE_UPDATE)
    check(lib().sgc_uniform_stream(g.n, g.row_ptr, g.col_idx, size, seed, out.ctypes.data))
    return out[:size]


def holdout_stream(g, pct=10, seed=42):
    """The README's hold-out stream: (base graph without pct% of the edges, those edges as inserts)."""
    cnt = C.c_uint64()
    check(lib().sgc_holdout_stream(g.n, g.row_ptr, g.col_idx, pct, seed, None, None, C.byref(cnt)))
    out = np.zeros(max(cnt.value, 1), dtype=EDGE_UPDATE)
    h = C.c_void_p()
    check(lib().sgc_holdout_stream(g.n, g.row_ptr, g.col_idx, pct, seed, C.byref(h), out.ctypes.data,
                                   C.byref(cnt)))
    return _from_handle(h), out[:cnt.value]


def make_updates(u, v, t=None, graph=0):
    r = np.zeros(len(u), dtype=EDGE_UPDATE)
    r["u"] = u
    r["v"] = v
    if t is not None:
        r["type"] = t
    r["graph"] = graph
    return r


def generate_init_stream(g):
    """= generate_init_stream (A/src/host/stream_gen.cpp:30-61): all-ADD records."""
    out = np.zeros(max(g.num_edges // 2, 1), dtype=EDGE_UPDATE)
    cnt = C.c_uint64()
    check(lib().sgc_init_stream(g.n, g.row_ptr, g.col_idx, g.degree, out.ctypes.data, C.byref(cnt)))
    return out[:cnt.value]


def generate_bench_stream_edges(g, stream_size, insert_pct, seed):
    """= generate_bench_stream_edges (A/src/host/stream_gen.cpp:65-177): churn records."""
    out = np.zeros(max(stream_size, 1), dtype=EDGE_UPDATE)
    check(lib().sgc_churn_stream(g.n, g.row_ptr, g.col_idx, stream_size, insert_pct, seed, out.ctypes.data))
    return out[:stream_size]


def static_greedy_coloring(g):
    """= static_greedy_coloring (A/src/streamgc.cu:559-592): max colour of id-order first fit."""
    mx = C.c_uint32()
    check(lib().sgc_static_greedy(g.n, g.row_ptr, g.col_idx, C.byref(mx)))
    return mx.value


def edge_intervals(g, stream):
    """Life intervals [born, died) of every edge over a stream (for the independent validator)."""
    L = lib()
    m = C.c_uint64()
    stream = np.ascontiguousarray(stream, dtype=EDGE_UPDATE)
    check(L.sgc_edge_intervals(g.n, g.row_ptr, g.col_idx, stream.ctypes.data, len(stream), C.byref(m),
                               None, None, None, None))
    eu = np.zeros(max(m.value, 1), np.uint32)
    ev = np.zeros(max(m.value, 1), np.uint32)
    born = np.zeros(max(m.value, 1), np.uint64)
    died = np.zeros(max(m.value, 1), np.uint64)
    check(L.sgc_edge_intervals(g.n, g.row_ptr, g.col_idx, stream.ctypes.data, len(stream), C.byref(m),
                               eu.ctypes.data, ev.ctypes.data, born.ctypes.data, died.ctypes.data))
    k = m.value
    return eu[:k], ev[:k], born[:k], died[:k]


class DeviceStream:
    """An update stream resident in HBM (uploaded once)."""

    def __init__(self, records):
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        self.count = len(records)
        p = C.c_void_p()
        check(lib().sgc_device_upload(records.ctypes.data, records.nbytes, C.byref(p)))
        self.ptr = p.value

    def at(self, offset):
        return self.ptr + 16 * int(offset)

    def close(self):
        if self.ptr:
            lib().sgc_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class GPUPartition:
    """One graph resident on the GPU (A/include/streamgc.cuh:26-110 + allocate_gpu_partition)."""

    def __init__(self, g, config=None):
        self.graph = g
        self.config = config or StreamGCConfig()
        h = C.c_void_p()
        cfg = self.config.to_c()
        check(lib().sgc_create(C.byref(cfg), g.n, g.row_ptr, g.col_idx if g.num_edges else
                               np.zeros(1, np.uint32), g.degree if g.n else np.zeros(1, np.uint32),
                               C.byref(h)))
        self.h = h.value
        self.n = g.n

    def close(self):
        if getattr(self, "h", None):
            lib().sgc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc):
        return check(rc, self.h)

    # -- the hot path
    def process_stream(self, records, max_epochs=4096):
        """= process_stream (A/src/streamgc.cu:809-1053).  records: host array or DeviceStream."""
        st = (_lib.EpochStats * max(max_epochs, 1))()
        ne = C.c_uint32()
        if isinstance(records, DeviceStream):
            self._ck(lib().sgc_submit(self.h, records.ptr, records.count, 1, st, max_epochs, C.byref(ne)))
        else:
            records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
            self._ck(lib().sgc_submit(self.h, records.ctypes.data, len(records), 0, st, max_epochs, C.byref(ne)))
        return [st[i].as_dict() for i in range(min(ne.value, max_epochs))]

    def submit_device(self, ptr, count, max_epochs=0):
        st = (_lib.EpochStats * max(max_epochs, 1))()
        ne = C.c_uint32()
        self._ck(lib().sgc_submit(self.h, ptr, count, 1, st if max_epochs else None, max_epochs, C.byref(ne)))
        return [st[i].as_dict() for i in range(min(ne.value, max_epochs))]

    def batch_prepare(self, records, use_graph=True, t1_sweeps=0, autotune_rounds=0, t2_passes=0):
        """The stream's epochs as captured rounds (see GraphPool.batch_prepare); returns the rounds."""
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        cfg = _lib.BatchConfig()
        lib().sgc_default_batch_config(C.byref(cfg))
        cfg.use_graph, cfg.t1_sweeps, cfg.autotune_rounds, cfg.t2_passes = (1 if use_graph else 0, t1_sweeps,
                                                                             autotune_rounds, t2_passes)
        r = C.c_uint32()
        self._ck(lib().sgc_batch_prepare(self.h, records.ctypes.data, len(records), C.byref(cfg), C.byref(r)))
        return r.value

    def batch_run(self):
        st = _lib.BatchStats()
        self._ck(lib().sgc_batch_run(self.h, C.byref(st)))
        return st.as_dict()

    def batch_release(self):
        self._ck(lib().sgc_batch_release(self.h))

    def apply(self, records):
        """= one local_update_kernel launch per batch, no epoch step."""
        if isinstance(records, DeviceStream):
            self._ck(lib().sgc_apply(self.h, records.ptr, records.count, 1))
        else:
            records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
            self._ck(lib().sgc_apply(self.h, records.ctypes.data, len(records), 0))

    def apply_device(self, ptr, count):
        """One batch already resident on the device (no epoch step)."""
        self._ck(lib().sgc_apply(self.h, ptr, count, 1))

    def set_tuning(self, **knobs):
        """Launch-shape knobs (long_row, sweep_blocks, apply_blocks); results never change."""
        self._ck(lib().sgc_set_tuning(self.h, C.byref(make_tuning(**knobs))))

    def set_mode(self, deterministic=None, repair_every_epoch=None, exact_histogram=None, full_sweeps=None,
                 host_tier2=None):
        """Switch the epoch procedure (and determinism) between phases."""
        if host_tier2 is not None:
            self.config.host_tier2 = bool(host_tier2)
        if deterministic is not None:
            self.config.deterministic = bool(deterministic)
        if repair_every_epoch is not None:
            self.config.repair_every_epoch = bool(repair_every_epoch)
        if exact_histogram is not None:
            self.config.exact_histogram = bool(exact_histogram)
        if full_sweeps is not None:
            self.config.full_sweeps = bool(full_sweeps)
        self._ck(lib().sgc_set_flags(self.h, self.config.to_c().flags))

    def epoch(self):
        st = _lib.EpochStats()
        self._ck(lib().sgc_epoch(self.h, C.byref(st)))
        return st.as_dict()

    def finish(self):
        """= drain_and_compact + run_conflict_repair + cpu_serial_repair."""
        st = _lib.EpochStats()
        self._ck(lib().sgc_finish(self.h, C.byref(st)))
        return st.as_dict()

    def query(self, ids):
        """= point_query over the published snapshot."""
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(len(ids), np.uint16)
        if len(ids):
            self._ck(lib().sgc_query(self.h, ids, len(ids), out))
        return out

    # -- state access
    def words(self):
        w = np.zeros(self.n, np.uint32)
        self._ck(lib().sgc_download_words(self.h, w))
        return w

    def colors(self):
        return (self.words() & 0xFFFF).astype(np.uint16)

    def snapshot(self):
        s = np.zeros(self.n, np.uint16)
        self._ck(lib().sgc_download_snapshot(self.h, s))
        return s

    def csr(self):
        s = C.c_uint64()
        self._ck(lib().sgc_num_slots(self.h, C.byref(s)))
        rp = np.zeros(self.n + 1, np.uint32)
        col = np.zeros(max(s.value, 1), np.uint32)
        self._ck(lib().sgc_download_csr(self.h, rp, col))
        return rp, col[:s.value]

    def bitmaps(self):
        b = np.zeros(self.n, np.uint64)
        self._ck(lib().sgc_download_bitmaps(self.h, b))
        return b

    def hist(self):
        k = C.c_uint32()
        self._ck(lib().sgc_num_hd(self.h, C.byref(k)))
        h = np.zeros(max(k.value * 1024, 1), np.uint32)
        self._ck(lib().sgc_download_hist(self.h, h))
        return h[:k.value * 1024].reshape(k.value, 1024)

    def set_words(self, w):
        self._ck(lib().sgc_upload_words(self.h, np.ascontiguousarray(w, np.uint32)))

    def set_bitmaps(self, b):
        self._ck(lib().sgc_upload_bitmaps(self.h, np.ascontiguousarray(b, np.uint64)))

    def set_hist(self, h):
        self._ck(lib().sgc_upload_hist(self.h, np.ascontiguousarray(h, np.uint32).ravel()))

    def find_color(self, v):
        out = C.c_uint16()
        self._ck(lib().sgc_find_color(self.h, v, C.byref(out)))
        return out.value

    def overflow_count(self):
        c = C.c_uint32()
        self._ck(lib().sgc_overflow_count(self.h, C.byref(c)))
        return c.value

    def totals(self):
        t = _lib.Totals()
        self._ck(lib().sgc_totals_get(self.h, C.byref(t)))
        return t.as_dict()

    def validate(self, snapshot=False):
        """= check_coloring_valid on the engine's rows: (violations, uncoloured-with-edges, max colour)."""
        a, b, m = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._ck(lib().sgc_validate(self.h, 1 if snapshot else 0, C.byref(a), C.byref(b), C.byref(m)))
        return a.value, b.value, m.value

    def history_reserve(self, capacity):
        self._ck(lib().sgc_history_reserve(self.h, capacity))

    def history_count(self):
        c = C.c_uint32()
        self._ck(lib().sgc_history_count(self.h, C.byref(c)))
        return c.value

    def validate_history(self, intervals, positions):
        eu, ev, born, died = intervals
        pos = np.ascontiguousarray(positions, np.uint64)
        bad = np.zeros(max(len(pos), 1), np.uint64)
        self._ck(lib().sgc_validate_history(self.h, eu, ev, born, died, len(eu), pos, len(pos), bad))
        return bad[:len(pos)]

    def kernel_timing(self, enable=True):
        self._ck(lib().sgc_kernel_timing_enable(self.h, 1 if enable else 0))

    def kernel_times(self):
        cap = 64
        stride = 32
        names = C.create_string_buffer(cap * stride)
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        cnt = C.c_uint32()
        self._ck(lib().sgc_kernel_times(self.h, names, stride, ms, launches, cap, C.byref(cnt)))
        out = {}
        for i in range(min(cnt.value, cap)):
            nm = names.raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
            out[nm] = (ms[i], launches[i])
        return out

    def last_timing(self):
        a = (C.c_double * 4)()
        self._ck(lib().sgc_last_timing(self.h, a))
        return {"update_ms": a[0], "epoch_ms": a[1], "tier2_host_ms": a[2], "wall_ms": a[3]}


class GraphPool:
    """Many graphs resident on one GPU; every phase is one launch over all graphs taking part.

    Records carry the graph id in their `graph` field.  Flags are pool-wide.
    """

    def __init__(self, max_graphs=256, batch_capacity=1 << 18, hbm_budget=0, deterministic=False,
                 repair_every_epoch=False, exact_histogram=False, full_sweeps=False, host_tier2=False):
        flags = StreamGCConfig(deterministic=deterministic, repair_every_epoch=repair_every_epoch,
                               exact_histogram=exact_histogram, full_sweeps=full_sweeps,
                               host_tier2=host_tier2).to_c().flags
        cfg = _lib.PoolConfig(max_graphs, batch_capacity, flags, 0, hbm_budget)
        h = C.c_void_p()
        check(lib().sgc_pool_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self.flags = flags
        self.batch_capacity = batch_capacity
        self.graphs = {}

    def close(self):
        if getattr(self, "h", None):
            lib().sgc_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc):
        return check(rc, pool=self.h)

    @staticmethod
    def graph_bytes(g, config=None):
        cfg = (config or StreamGCConfig()).to_c()
        b = C.c_uint64()
        check(lib().sgc_pool_graph_bytes(C.byref(cfg), g.n, g.row_ptr, g.degree, C.byref(b)))
        return b.value

    def add(self, g, config=None):
        """Admit a graph; raises SgcNoMemory when it does not fit."""
        cfg = (config or StreamGCConfig()).to_c()
        gid = C.c_uint32()
        col = g.col_idx if g.num_edges else np.zeros(1, np.uint32)
        self._ck(lib().sgc_pool_add(self.h, C.byref(cfg), g.n, g.row_ptr, col, g.degree, C.byref(gid)))
        self.graphs[gid.value] = g
        return gid.value

    def remove(self, gid):
        self._ck(lib().sgc_pool_remove(self.h, gid))
        self.graphs.pop(gid, None)

    def usage(self):
        u, b, k = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._ck(lib().sgc_pool_usage(self.h, C.byref(u), C.byref(b), C.byref(k)))
        return u.value, b.value, k.value

    def set_tuning(self, **knobs):
        self._ck(lib().sgc_pool_set_tuning(self.h, C.byref(make_tuning(**knobs))))

    def set_mode(self, deterministic=None, repair_every_epoch=None, exact_histogram=None, full_sweeps=None,
                 host_tier2=None):
        c = StreamGCConfig(deterministic=bool(self.flags & _lib.SGC_DETERMINISTIC),
                           repair_every_epoch=bool(self.flags & _lib.SGC_REPAIR_EVERY_EPOCH),
                           exact_histogram=bool(self.flags & _lib.SGC_EXACT_HISTOGRAM),
                           full_sweeps=bool(self.flags & _lib.SGC_FULL_SWEEPS),
                           host_tier2=bool(self.flags & _lib.SGC_HOST_TIER2))
        if host_tier2 is not None:
            c.host_tier2 = host_tier2
        if deterministic is not None:
            c.deterministic = deterministic
        if repair_every_epoch is not None:
            c.repair_every_epoch = repair_every_epoch
        if exact_histogram is not None:
            c.exact_histogram = exact_histogram
        if full_sweeps is not None:
            c.full_sweeps = full_sweeps
        self.flags = c.to_c().flags
        self._ck(lib().sgc_pool_set_flags(self.h, self.flags))

    def apply(self, records):
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        self._ck(lib().sgc_pool_apply(self.h, records.ctypes.data, len(records), 0))

    def submit(self, records, max_stats=65536):
        """process_stream over a mixed stream: batches, then one shared epoch for the due graphs."""
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        st = (_lib.EpochStats * max(max_stats, 1))()
        k = C.c_uint32()
        self._ck(lib().sgc_pool_submit(self.h, records.ctypes.data, len(records), 0, st if max_stats else None,
                                       max_stats, C.byref(k)))
        return [st[i].as_dict() for i in range(min(k.value, max_stats))]

    def set_epoch_size(self, gid, epoch_size):
        """E_g of one graph (takes effect at its next batch)."""
        self._ck(lib().sgc_pool_set_epoch_size(self.h, gid, epoch_size))

    def batch_prepare(self, records, use_graph=True, t1_sweeps=0, autotune_rounds=0, t2_passes=0):
        """Batched pattern: make the round plan of a mixed stream and keep it on the device.

        Round r gives every graph its next min(E_g, left) records, then every
        graph of the round takes an epoch step (its last partial epoch too);
        use_graph replays one captured hipGraph per round.  Returns the rounds.
        """
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        cfg = _lib.BatchConfig()
        lib().sgc_default_batch_config(C.byref(cfg))
        cfg.use_graph = 1 if use_graph else 0
        cfg.t1_sweeps = t1_sweeps
        cfg.autotune_rounds = autotune_rounds
        cfg.t2_passes = t2_passes
        r = C.c_uint32()
        self._ck(lib().sgc_pool_batch_prepare(self.h, records.ctypes.data, len(records), C.byref(cfg), C.byref(r)))
        return r.value

    def batch_run(self):
        st = _lib.BatchStats()
        self._ck(lib().sgc_pool_batch_run(self.h, C.byref(st)))
        return st.as_dict()

    def batch_release(self):
        self._ck(lib().sgc_pool_batch_release(self.h))

    def due(self):
        cap = max(len(self.graphs), 1)
        out = np.zeros(cap, np.uint32)
        k = C.c_uint32()
        self._ck(lib().sgc_pool_due(self.h, out, cap, C.byref(k)))
        return out[:min(k.value, cap)].tolist()

    def _epoch(self, fn, gids):
        gids = np.ascontiguousarray(gids, np.uint32)
        st = (_lib.EpochStats * max(len(gids), 1))()
        self._ck(fn(self.h, gids, len(gids), st))
        return [st[i].as_dict() for i in range(len(gids))]

    def epoch(self, gids):
        return self._epoch(lib().sgc_pool_epoch, gids)

    def finish(self, gids=None):
        return self._epoch(lib().sgc_pool_finish, sorted(self.graphs) if gids is None else gids)

    def query(self, gid, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(len(ids), np.uint16)
        if len(ids):
            self._ck(lib().sgc_pool_query(self.h, gid, ids, len(ids), out))
        return out

    def words(self, gid):
        w = np.zeros(self.graphs[gid].n, np.uint32)
        self._ck(lib().sgc_pool_download_words(self.h, gid, w))
        return w

    def colors(self, gid):
        return (self.words(gid) & 0xFFFF).astype(np.uint16)

    def info(self, gid):
        """Size, epochs, overflow events ever, published buffer and host Tier-2 time of one graph."""
        out = _lib.GraphInfo()
        self._ck(lib().sgc_pool_graph_info(self.h, gid, C.byref(out)))
        return out.as_dict()

    def csr(self, gid):
        """(row_ptr, slots) of one graph's slack CSR (INVALID = free slot)."""
        inf = self.info(gid)
        rp = np.zeros(inf["n"] + 1, np.uint32)
        col = np.zeros(max(inf["slots"], 1), np.uint32)
        self._ck(lib().sgc_pool_download_csr(self.h, gid, rp, col))
        return rp, col[:inf["slots"]]

    def bitmaps(self, gid):
        b = np.zeros(self.info(gid)["n"], np.uint64)
        self._ck(lib().sgc_pool_download_bitmaps(self.h, gid, b))
        return b

    def hist(self, gid):
        h = np.zeros((self.info(gid)["num_hd"], 1024), np.uint32)
        self._ck(lib().sgc_pool_download_hist(self.h, gid, h.ctypes.data if h.size else None))
        return h

    def set_words(self, gid, words):
        self._ck(lib().sgc_pool_upload_words(self.h, gid, np.ascontiguousarray(words, np.uint32)))

    def snapshot(self, gid):
        s = np.zeros(self.graphs[gid].n, np.uint16)
        self._ck(lib().sgc_pool_download_snapshot(self.h, gid, s))
        return s

    def validate(self, gid, snapshot=False):
        a, b, m = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._ck(lib().sgc_pool_validate(self.h, gid, 1 if snapshot else 0, C.byref(a), C.byref(b), C.byref(m)))
        return a.value, b.value, m.value

    def history_reserve(self, gid, capacity):
        self._ck(lib().sgc_pool_history_reserve(self.h, gid, capacity))

    def history_count(self, gid):
        c = C.c_uint32()
        self._ck(lib().sgc_pool_history_count(self.h, gid, C.byref(c)))
        return c.value

    def validate_history(self, gid, intervals, positions):
        eu, ev, born, died = intervals
        pos = np.ascontiguousarray(positions, np.uint64)
        bad = np.zeros(max(len(pos), 1), np.uint64)
        self._ck(lib().sgc_pool_validate_history(self.h, gid, eu, ev, born, died, len(eu), pos, len(pos), bad))
        return bad[:len(pos)]

    def totals(self):
        t = _lib.Totals()
        self._ck(lib().sgc_pool_totals(self.h, C.byref(t)))
        return t.as_dict()

    def kernel_timing(self, enable=True):
        self._ck(lib().sgc_pool_kernel_timing_enable(self.h, 1 if enable else 0))

    def kernel_times(self):
        cap, stride = 64, 32
        names = C.create_string_buffer(cap * stride)
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        cnt = C.c_uint32()
        self._ck(lib().sgc_pool_kernel_times(self.h, names, stride, ms, launches, cap, C.byref(cnt)))
        return {names.raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode(): (ms[i], launches[i])
                for i in range(min(cnt.value, cap))}


def tag_graph(records, gid):
    """Copy of records addressed to graph gid of a pool."""
    r = np.array(records, dtype=EDGE_UPDATE, copy=True)
    r["graph"] = gid
    return r


def check_coloring_valid(partition, snapshot=False):
    """= check_coloring_valid (A/src/streamgc.cu:448-556)."""
    return partition.validate(snapshot)


class GraphServer:
    """The served pattern over a GraphPool: a persistent loop thread takes
    update batches and queries for many resident graphs as they arrive,
    packs all arrived updates into one launch (continuous batching) and runs
    shared epoch steps for the graphs that are due.  While it runs, the pool
    is driven only by the server."""

    BATCHED = 0
    PLAIN = 1

    def __init__(self, pool, max_batch=32768, max_delay_us=0, mode=0):
        """mode GraphServer.PLAIN serves the same requests as the plain per-launch
        engine would (one launch per request, one epoch step per due graph)."""
        cfg = _lib.ServerConfig()
        lib().sgc_default_server_config(C.byref(cfg))
        cfg.max_batch = max_batch
        cfg.max_delay_us = max_delay_us
        cfg.mode = mode
        h = C.c_void_p()
        check(lib().sgc_server_create(pool.h, C.byref(cfg), C.byref(h)))
        self.h = h.value
        self.pool = pool
        self.graphs = {}

    def _ck(self, rc):
        if rc != 0:
            msg = (lib().sgc_server_last_error(self.h) or b"").decode()
            raise (SgcNoMemory if rc == -3 else SgcError)(f"streamgc server error {rc}: {msg}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            rc = lib().sgc_server_destroy(self.h)
            self.h = None
            if rc:
                raise SgcError(f"streamgc server loop ended with error {rc}")

    def __del__(self):
        try:
            self.close()
        except SgcError:
            pass

    def admit(self, g, config=None):
        """Queue a graph for admission; returns a ticket for admitted()."""
        cfg = (config or StreamGCConfig()).to_c()
        t = C.c_uint64()
        col = g.col_idx if g.num_edges else np.zeros(1, np.uint32)
        self._ck(lib().sgc_server_admit(self.h, C.byref(cfg), g.n, g.row_ptr, col, g.degree, C.byref(t)))
        self._pending_graph = getattr(self, "_pending_graph", {})
        self._pending_graph[t.value] = g
        return t.value

    def admitted(self, ticket, timeout_ms=60000):
        gid = C.c_uint32()
        self._ck(lib().sgc_server_admitted(self.h, ticket, timeout_ms, C.byref(gid)))
        g = self._pending_graph.pop(ticket, None)
        self.graphs[gid.value] = g
        if g is not None:
            self.pool.graphs[gid.value] = g
        return gid.value

    def add(self, g, config=None, timeout_ms=60000):
        return self.admitted(self.admit(g, config), timeout_ms)

    def attach(self, gid, epoch_size, timeout_ms=60000):
        """Serve a graph already in the pool (added and initialised through it)."""
        t = C.c_uint64()
        self._ck(lib().sgc_server_attach(self.h, gid, epoch_size, C.byref(t)))
        self._pending_graph = getattr(self, "_pending_graph", {})
        self._pending_graph[t.value] = self.pool.graphs.get(gid)
        return self.admitted(t.value, timeout_ms)

    def remove(self, gid):
        """Remove a graph once its pending updates are published."""
        self._ck(lib().sgc_server_remove(self.h, gid))
        self.graphs.pop(gid, None)
        self.pool.graphs.pop(gid, None)

    def submit(self, gid, records):
        records = np.ascontiguousarray(records, dtype=EDGE_UPDATE)
        t = C.c_uint64()
        self._ck(lib().sgc_server_submit(self.h, gid, records.ctypes.data, len(records), C.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms=60000):
        self._ck(lib().sgc_server_wait(self.h, ticket, timeout_ms))

    def flush(self, timeout_ms=120000):
        self._ck(lib().sgc_server_flush(self.h, timeout_ms))

    def query(self, gid, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(len(ids), np.uint16)
        self._ck(lib().sgc_server_query(self.h, gid, ids, len(ids), out))
        return out

    def stats(self):
        st = _lib.ServerStats()
        self._ck(lib().sgc_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def latencies(self, which="update"):
        """Raw latencies in microseconds ("update": update-to-snapshot per request, "query")."""
        k = 0 if which == "update" else 1
        n = C.c_uint64()
        self._ck(lib().sgc_server_latencies(self.h, k, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        self._ck(lib().sgc_server_latencies(self.h, k, out.ctypes.data, n.value, C.byref(n)))
        return out[:n.value]

    def reset_stats(self):
        self._ck(lib().sgc_server_reset_stats(self.h))

    def watermarks(self, gid):
        """Per published snapshot of gid (publish order): updates submitted through the server it contains."""
        n = C.c_uint64()
        self._ck(lib().sgc_server_watermarks(self.h, gid, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint64)
        self._ck(lib().sgc_server_watermarks(self.h, gid, out.ctypes.data, n.value, C.byref(n)))
        return out[:n.value]

    def replay(self, items, updates, query_ids=None, timeout_ms=600000):
        """Open-loop replay of a trace (TRACE_ITEM records, at_ns after the start):
        updates and queries submitted at their scheduled times, latencies from
        those times.  Returns the replay statistics once everything is published."""
        items = np.ascontiguousarray(items, dtype=TRACE_ITEM)
        updates = np.ascontiguousarray(updates, dtype=EDGE_UPDATE)
        qids = np.ascontiguousarray(query_ids if query_ids is not None else np.zeros(1, np.uint32), np.uint32)
        st = _lib.ReplayStats()
        self._ck(lib().sgc_server_replay(self.h, items.ctypes.data, len(items), updates.ctypes.data,
                                         qids.ctypes.data, timeout_ms, C.byref(st)))
        return st.as_dict()


def write_coloring(path, colors):
    """<path> (raw u16 per vertex) and <path>.txt, as the reference writes them."""
    colors = np.ascontiguousarray(colors, dtype=np.uint16)
    check(lib().sgc_write_coloring(os.fsencode(path), colors, len(colors)))


def read_coloring(path):
    return np.fromfile(path, dtype="<u2")


def write_coloring_meta(path, **fields):
    """<path> (.coloring.meta JSON with the reference's keys); fields of sgc_run_meta."""
    m = _lib.RunMeta()
    for k, v in fields.items():
        if k in ("graph", "hw_tag"):
            v = str(v).encode()
        setattr(m, k, v)
    check(lib().sgc_write_meta(os.fsencode(path), C.byref(m)))
