"""ctypes binding of libstreamgc_amd.so (the C ABI declared in include/streamgc.h).

The library is loaded from this package directory only; there is no fallback:
a missing or unloadable library raises, so no caller can silently run
anything but the HIP engine.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstreamgc_amd.so")

SGC_OK = 0
SGC_DETERMINISTIC = 1 << 0
SGC_REPAIR_EVERY_EPOCH = 1 << 1
SGC_EXACT_HISTOGRAM = 1 << 2
SGC_FULL_SWEEPS = 1 << 3
SGC_HOST_TIER2 = 1 << 4
SGC_ABL_FORCE_BITMAP = 1 << 5
SGC_ABL_FORCE_HISTOGRAM = 1 << 6
SGC_ABL_NO_OVERFLOW = 1 << 7
SGC_ABL_LIVE_DEGREE = 1 << 8
ABLATIONS = {"force_bitmap": SGC_ABL_FORCE_BITMAP, "force_histogram": SGC_ABL_FORCE_HISTOGRAM,
             "no_overflow": SGC_ABL_NO_OVERFLOW, "live_degree": SGC_ABL_LIVE_DEGREE}

# numpy view of the 16-byte update record (A/include/types.h:16-21)
EDGE_UPDATE = np.dtype([("u", "<u4"), ("v", "<u4"), ("type", "u1"), ("pad", "u1", 3), ("graph", "<u4")])
assert EDGE_UPDATE.itemsize == 16
# One arrival of a replayed trace (sgc_trace_item): kind 0 = updates, 1 = query.
TRACE_ITEM = np.dtype([("at_ns", "<u8"), ("offset", "<u8"), ("gid", "<u4"), ("count", "<u4"), ("kind", "<u4"),
                       ("reserved", "<u4")])


class Config(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "epoch_size", "batch_size", "histogram_threshold", "slack_pct", "min_row_capacity",
        "overflow_capacity", "max_repair_iters", "plateau_iters", "tier2_passes", "flags")]


class EpochStats(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "epoch_index", "overflow_count", "overflow_drained", "overflow_failed", "t1_iterations",
        "t1_conflicts_last", "t2_fixed", "t2_remaining", "max_color", "snapshot_published",
        "t2_ran", "reserved")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Totals(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "updates", "scan_slots", "hd_endpoints", "recolors", "recolor_cap_slots", "recolor_live",
        "epochs", "t1_sweeps", "t2_runs", "overflow_events", "live_slots", "grouped")] + \
        [("sweep_rows", C.c_uint64 * 5), ("sweep_slots", C.c_uint64 * 5)]

    SWEEPS = ("compact", "rebuild", "tier1", "t2_check", "tier1_serial")

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("sweep_")}
        for i, name in enumerate(self.SWEEPS):
            d[f"rows_{name}"] = self.sweep_rows[i]
            d[f"slots_{name}"] = self.sweep_slots[i]
        return d


class Tuning(C.Structure):
    _fields_ = [("long_row", C.c_uint32), ("sweep_blocks", C.c_uint32), ("apply_blocks", C.c_uint32),
                ("serial_repair", C.c_uint32), ("serial_conflicts", C.c_uint32), ("fresh_repair", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("max_delay_us", C.c_uint32), ("mode", C.c_uint32),
                ("reserved", C.c_uint32)]


class TraceItem(C.Structure):
    _fields_ = [("at_ns", C.c_uint64), ("offset", C.c_uint64), ("gid", C.c_uint32), ("count", C.c_uint32),
                ("kind", C.c_uint32), ("reserved", C.c_uint32)]


class ReplayStats(C.Structure):
    _fields_ = [("span_us", C.c_double), ("max_late_us", C.c_double), ("mean_late_us", C.c_double)] + \
               [(k, C.c_uint64) for k in ("requests", "updates", "queries")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("updates_applied", "batches", "epoch_steps", "graph_epochs",
                                          "requests_done", "queries", "queued_updates")] + \
               [(k, C.c_double) for k in ("span_us", "busy_us", "intake_us", "pack_us", "apply_us", "epoch_us")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunMeta(C.Structure):
    _fields_ = [("graph", C.c_char_p), ("hw_tag", C.c_char_p), ("num_vertices", C.c_uint32),
                ("num_colors_used", C.c_uint32), ("num_edges", C.c_uint64), ("static_greedy_colors", C.c_uint32),
                ("insert_pct", C.c_uint32), ("stream_size", C.c_uint64), ("seed", C.c_uint64)] + \
               [(k, C.c_double) for k in ("init_time_ms", "bench_time_ms", "bench_stream_ms", "repair_time_ms",
                                          "local_kernel_ms", "gpu_sync_ms", "epoch_maintenance_ms", "memcpy_ms",
                                          "query_latency_us")] + [("query_count", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("use_graph", C.c_uint32), ("t1_sweeps", C.c_uint32), ("autotune_rounds", C.c_uint32),
                ("t2_passes", C.c_uint32)]


BATCH_CANDIDATES = 4


class BatchStats(C.Structure):
    _fields_ = [("rounds", C.c_uint32), ("host_rounds", C.c_uint32), ("candidates", C.c_uint32),
                ("chosen", C.c_uint32), ("updates", C.c_uint64), ("ms", C.c_double),
                ("candidate_ns_per_update", C.c_double * BATCH_CANDIDATES),
                ("candidate", Tuning * BATCH_CANDIDATES),
                ("candidate_samples", C.c_uint32 * BATCH_CANDIDATES),
                ("tune_rounds", C.c_uint32), ("tune_updates", C.c_uint64), ("tune_ms", C.c_double)]

    def as_dict(self):
        k = max(self.candidates, 1)
        return {"rounds": self.rounds, "host_rounds": self.host_rounds, "candidates": self.candidates,
                "chosen": self.chosen, "updates": self.updates, "ms": self.ms,
                "candidate_ns_per_update": [self.candidate_ns_per_update[i] for i in range(k)],
                "candidate": [{f: getattr(self.candidate[i], f) for f, _ in Tuning._fields_} for i in range(k)],
                "candidate_samples": [self.candidate_samples[i] for i in range(k)],
                "tune_rounds": self.tune_rounds, "tune_updates": self.tune_updates, "tune_ms": self.tune_ms}


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("num_hd", C.c_uint32), ("slots", C.c_uint64), ("bytes", C.c_uint64),
                ("epoch_size", C.c_uint32), ("epochs", C.c_uint32), ("overflow_events", C.c_uint32),
                ("published", C.c_int32), ("tier2_ms", C.c_double), ("tier2_runs", C.c_uint32),
                ("lost", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PoolConfig(C.Structure):
    _fields_ = [("max_graphs", C.c_uint32), ("batch_capacity", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32), ("hbm_budget", C.c_uint64)]


_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_recp = np.ctypeslib.ndpointer(dtype=EDGE_UPDATE, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# name -> (argtypes); every function returns int unless listed in _RESTYPE
SIGNATURES = {
    "sgc_default_config": [C.POINTER(Config)],
    "sgc_default_tuning": [C.POINTER(Tuning)],
    "sgc_set_tuning": [_vp, C.POINTER(Tuning)],
    "sgc_pool_set_tuning": [_vp, C.POINTER(Tuning)],
    "sgc_last_error": [_vp],
    "sgc_version": [],
    "sgc_device_count": [],
    "sgc_device_name": [C.c_int, C.c_char_p, C.c_uint32],
    "sgc_microbench_trc": [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                           C.POINTER(C.c_uint32)],
    "sgc_create": [C.POINTER(Config), C.c_uint32, _u32p, _u32p, _u32p, C.POINTER(_vp)],
    "sgc_destroy": [_vp],
    "sgc_submit": [_vp, _vp, C.c_uint64, C.c_int, C.POINTER(EpochStats), C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_apply": [_vp, _vp, C.c_uint64, C.c_int],
    "sgc_epoch": [_vp, C.POINTER(EpochStats)],
    "sgc_finish": [_vp, C.POINTER(EpochStats)],
    "sgc_query": [_vp, _u32p, C.c_uint32, _u16p],
    "sgc_set_flags": [_vp, C.c_uint32],
    "sgc_download_words": [_vp, _u32p],
    "sgc_download_colors": [_vp, _u16p],
    "sgc_download_snapshot": [_vp, _u16p],
    "sgc_num_slots": [_vp, C.POINTER(C.c_uint64)],
    "sgc_download_csr": [_vp, _u32p, _u32p],
    "sgc_download_bitmaps": [_vp, _u64p],
    "sgc_num_hd": [_vp, C.POINTER(C.c_uint32)],
    "sgc_download_hist": [_vp, _u32p],
    "sgc_upload_words": [_vp, _u32p],
    "sgc_upload_bitmaps": [_vp, _u64p],
    "sgc_upload_hist": [_vp, _u32p],
    "sgc_totals_get": [_vp, C.POINTER(Totals)],
    "sgc_overflow_count": [_vp, C.POINTER(C.c_uint32)],
    "sgc_find_color": [_vp, C.c_uint32, C.POINTER(C.c_uint16)],
    "sgc_validate": [_vp, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
    "sgc_history_reserve": [_vp, C.c_uint32],
    "sgc_history_count": [_vp, C.POINTER(C.c_uint32)],
    "sgc_validate_history": [_vp, _u32p, _u32p, _u64p, _u64p, C.c_uint64, _u64p, C.c_uint32, _u64p],
    "sgc_device_upload": [_vp, C.c_uint64, C.POINTER(_vp)],
    "sgc_device_free": [_vp],
    "sgc_last_timing": [_vp, C.POINTER(C.c_double)],
    "sgc_kernel_timing_enable": [_vp, C.c_int],
    "sgc_kernel_times": [_vp, C.c_char_p, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_uint32,
                         C.POINTER(C.c_uint32)],
    "sgc_pool_last_error": [_vp],
    "sgc_pool_create": [C.POINTER(PoolConfig), C.POINTER(_vp)],
    "sgc_pool_destroy": [_vp],
    "sgc_pool_graph_bytes": [C.POINTER(Config), C.c_uint32, _u32p, _u32p, C.POINTER(C.c_uint64)],
    "sgc_pool_add": [_vp, C.POINTER(Config), C.c_uint32, _u32p, _u32p, _u32p, C.POINTER(C.c_uint32)],
    "sgc_pool_remove": [_vp, C.c_uint32],
    "sgc_pool_usage": [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
    "sgc_pool_set_flags": [_vp, C.c_uint32],
    "sgc_pool_apply": [_vp, _vp, C.c_uint64, C.c_int],
    "sgc_pool_due": [_vp, _u32p, C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_pool_epoch": [_vp, _u32p, C.c_uint32, C.POINTER(EpochStats)],
    "sgc_pool_finish": [_vp, _u32p, C.c_uint32, C.POINTER(EpochStats)],
    "sgc_pool_submit": [_vp, _vp, C.c_uint64, C.c_int, C.POINTER(EpochStats), C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_pool_query": [_vp, C.c_uint32, _u32p, C.c_uint32, _u16p],
    "sgc_pool_query_noalloc": [_vp, C.c_uint32, _u32p, C.c_uint32, _u16p],
    "sgc_pool_download_words": [_vp, C.c_uint32, _u32p],
    "sgc_pool_upload_words": [_vp, C.c_uint32, _u32p],
    "sgc_pool_graph_info": [_vp, C.c_uint32, C.POINTER(GraphInfo)],
    "sgc_pool_download_csr": [_vp, C.c_uint32, _u32p, _u32p],
    "sgc_pool_download_bitmaps": [_vp, C.c_uint32, _u64p],
    "sgc_pool_download_hist": [_vp, C.c_uint32, _vp],
    "sgc_pool_download_snapshot": [_vp, C.c_uint32, _u16p],
    "sgc_pool_validate": [_vp, C.c_uint32, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                          C.POINTER(C.c_uint32)],
    "sgc_pool_history_reserve": [_vp, C.c_uint32, C.c_uint32],
    "sgc_pool_history_count": [_vp, C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_pool_validate_history": [_vp, C.c_uint32, _u32p, _u32p, _u64p, _u64p, C.c_uint64, _u64p, C.c_uint32,
                                  _u64p],
    "sgc_pool_totals": [_vp, C.POINTER(Totals)],
    "sgc_pool_kernel_timing_enable": [_vp, C.c_int],
    "sgc_pool_set_epoch_size": [_vp, C.c_uint32, C.c_uint32],
    "sgc_default_batch_config": [C.POINTER(BatchConfig)],
    "sgc_pool_batch_prepare": [_vp, _vp, C.c_uint64, C.POINTER(BatchConfig), C.POINTER(C.c_uint32)],
    "sgc_pool_batch_run": [_vp, C.POINTER(BatchStats)],
    "sgc_pool_batch_release": [_vp],
    "sgc_batch_prepare": [_vp, _vp, C.c_uint64, C.POINTER(BatchConfig), C.POINTER(C.c_uint32)],
    "sgc_batch_run": [_vp, C.POINTER(BatchStats)],
    "sgc_batch_release": [_vp],
    "sgc_pool_kernel_times": [_vp, C.c_char_p, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                              C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_default_server_config": [C.POINTER(ServerConfig)],
    "sgc_server_create": [_vp, C.POINTER(ServerConfig), C.POINTER(_vp)],
    "sgc_server_destroy": [_vp],
    "sgc_server_last_error": [_vp],
    "sgc_server_admit": [_vp, C.POINTER(Config), C.c_uint32, _u32p, _u32p, _u32p, C.POINTER(C.c_uint64)],
    "sgc_server_admitted": [_vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32)],
    "sgc_server_attach": [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)],
    "sgc_server_remove": [_vp, C.c_uint32],
    "sgc_server_submit": [_vp, C.c_uint32, _vp, C.c_uint32, C.POINTER(C.c_uint64)],
    "sgc_server_wait": [_vp, C.c_uint64, C.c_uint32],
    "sgc_server_flush": [_vp, C.c_uint32],
    "sgc_server_query": [_vp, C.c_uint32, _u32p, C.c_uint32, _u16p],
    "sgc_server_get_stats": [_vp, C.POINTER(ServerStats)],
    "sgc_server_latencies": [_vp, C.c_int, _vp, C.c_uint64, C.POINTER(C.c_uint64)],
    "sgc_server_reset_stats": [_vp],
    "sgc_server_watermarks": [_vp, C.c_uint32, _vp, C.c_uint64, C.POINTER(C.c_uint64)],
    "sgc_server_replay": [_vp, _vp, C.c_uint64, _vp, _vp, C.c_uint32, C.POINTER(ReplayStats)],
    "sgc_write_coloring": [C.c_char_p, _u16p, C.c_uint32],
    "sgc_write_meta": [C.c_char_p, C.POINTER(RunMeta)],
    "sgc_csr_load_mtx": [C.c_char_p, C.POINTER(_vp)],
    "sgc_csr_from_edges": [C.c_uint32, _u32p, _u32p, C.c_uint64, C.POINTER(_vp)],
    "sgc_csr_info": [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
    "sgc_csr_export": [_vp, _vp, _vp, _vp],
    "sgc_csr_free": [_vp],
    "sgc_gen_chung_lu": [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_uint64, C.c_int, C.POINTER(_vp)],
    "sgc_gen_rmat": [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.POINTER(_vp)],
    "sgc_core_numbers": [C.c_uint32, _u32p, _u32p, _u32p],
    "sgc_dense_growth_stream": [C.c_uint32, _u32p, _u32p, _u32p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64,
                                _vp],
    "sgc_uniform_stream": [C.c_uint32, _u32p, _u32p, C.c_uint64, C.c_uint64, _vp],
    "sgc_holdout_stream": [C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint64, C.POINTER(_vp), _vp,
                           C.POINTER(C.c_uint64)],
    "sgc_init_stream": [C.c_uint32, _u32p, _u32p, _u32p, _vp, C.POINTER(C.c_uint64)],
    "sgc_churn_stream": [C.c_uint32, _u32p, _u32p, C.c_uint64, C.c_uint32, C.c_uint64, _vp],
    "sgc_static_greedy": [C.c_uint32, _u32p, _u32p, C.POINTER(C.c_uint32)],
    "sgc_edge_intervals": [C.c_uint32, _u32p, _u32p, _vp, C.c_uint64, C.POINTER(C.c_uint64), _vp, _vp, _vp, _vp],
}
_RESTYPE = {"sgc_default_config": None, "sgc_default_tuning": None, "sgc_last_error": C.c_char_p, "sgc_version": C.c_char_p,
            "sgc_pool_last_error": C.c_char_p, "sgc_server_last_error": C.c_char_p,
            "sgc_default_server_config": None, "sgc_default_batch_config": None}

_lib = None


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m streamgc_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


class SgcError(RuntimeError):
    pass


SGC_ERR_NOMEM = -3


class SgcNoMemory(SgcError):
    """Admission refused: the pool is full or its HBM budget would be exceeded."""


def check(rc, engine=None, pool=None):
    if rc != SGC_OK:
        msg = lib().sgc_pool_last_error(pool) if pool is not None else lib().sgc_last_error(engine)
        text = f"streamgc error {rc}: {msg.decode() if msg else ''}"
        raise (SgcNoMemory if rc == SGC_ERR_NOMEM else SgcError)(text)
    return rc
