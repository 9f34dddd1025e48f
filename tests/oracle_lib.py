"""ctypes access to the oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests call it to produce expected results and
compare the HIP engine against them.  It is never part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SGC_ORACLE_LIB points the tests at another build of oracle.cpp: the mutation
# tests use it to run the oracle's own tests against a copy with a planted slip.
LIB_PATH = os.environ.get("SGC_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Config(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "epoch_size", "batch_size", "histogram_threshold", "slack_pct", "min_row_capacity",
        "overflow_capacity", "max_repair_iters", "plateau_iters", "tier2_passes", "flags")]


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "epoch_index", "overflow_count", "overflow_drained", "overflow_failed", "t1_iterations",
        "t1_conflicts_last", "t2_fixed", "t2_remaining", "max_color", "snapshot_published",
        "t2_ran", "reserved")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


FLAG_DETERMINISTIC = 1 << 0
FLAG_REPAIR_EVERY_EPOCH = 1 << 1
FLAG_EXACT_HISTOGRAM = 1 << 2
FLAG_ABL_FORCE_BITMAP = 1 << 5
FLAG_ABL_FORCE_HISTOGRAM = 1 << 6
FLAG_ABL_NO_OVERFLOW = 1 << 7
FLAG_ABL_LIVE_DEGREE = 1 << 8


def default_config(**kw):
    c = Config(epoch_size=32768, batch_size=32768, histogram_threshold=64, slack_pct=300,
               min_row_capacity=8, overflow_capacity=1 << 18, max_repair_iters=20,
               plateau_iters=3, tier2_passes=30, flags=0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = C.CDLL(LIB_PATH)
        L.orc_mt64_nth.restype = C.c_uint64
        L.orc_mt64_nth.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_uniform_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _u64p]
        L.orc_init_stream.restype = C.c_uint64
        L.orc_init_stream.argtypes = [C.c_uint32, _u32p, _u32p, _u32p, _u32p, _u32p]
        L.orc_bench_stream.restype = C.c_uint64
        L.orc_bench_stream.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, C.c_uint32, C.c_uint64,
                                       _u32p, _u32p, _u8p]
        L.orc_load_mtx.restype = C.c_void_p
        L.orc_load_mtx.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
        L.orc_csr_get.argtypes = [C.c_void_p, _u32p, _u32p, _u32p]
        L.orc_csr_free.argtypes = [C.c_void_p]
        L.orc_engine_create.restype = C.c_void_p
        L.orc_engine_create.argtypes = [C.c_uint32, _u32p, _u32p, _u32p, C.POINTER(Config)]
        L.orc_engine_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_flags.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_submit.restype = C.c_uint32
        L.orc_submit.argtypes = [C.c_void_p, _u32p, _u32p, _u8p, C.c_uint64, C.POINTER(Stats), C.c_uint32]
        L.orc_apply.argtypes = [C.c_void_p, _u32p, _u32p, _u8p, C.c_uint64]
        for f in ("orc_epoch", "orc_finish", "orc_tier1", "orc_tier2"):
            getattr(L, f).argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.orc_tier1_sweep.restype = C.c_uint32
        L.orc_tier1_sweep.argtypes = [C.c_void_p]
        L.orc_rebuild_bitmaps.argtypes = [C.c_void_p]
        L.orc_compact.argtypes = [C.c_void_p]
        L.orc_num_slots.restype = C.c_uint64
        L.orc_num_slots.argtypes = [C.c_void_p]
        L.orc_num_hd.restype = C.c_uint32
        L.orc_num_hd.argtypes = [C.c_void_p]
        L.orc_get_words.argtypes = [C.c_void_p, _u32p]
        L.orc_set_words.argtypes = [C.c_void_p, _u32p]
        L.orc_get_snapshot.argtypes = [C.c_void_p, _u16p]
        L.orc_get_bitmaps.argtypes = [C.c_void_p, _u64p]
        L.orc_set_bitmaps.argtypes = [C.c_void_p, _u64p]
        L.orc_get_hist.argtypes = [C.c_void_p, _u32p]
        L.orc_set_hist.argtypes = [C.c_void_p, _u32p]
        L.orc_get_csr.argtypes = [C.c_void_p, _u32p, _u32p]
        L.orc_set_col.argtypes = [C.c_void_p, _u32p]
        L.orc_overflow_count.restype = C.c_uint32
        L.orc_overflow_count.argtypes = [C.c_void_p]
        L.orc_find_color.restype = C.c_uint16
        L.orc_find_color.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_validate_edges.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, _u16p,
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_static_greedy.restype = C.c_uint32
        L.orc_static_greedy.argtypes = [C.c_uint32, _u32p, _u32p]
        _lib = L
    return _lib


class Csr:
    """Plain symmetric CSR with sorted rows: row_ptr[n+1], col[nnz], deg[n] (uint32)."""

    def __init__(self, n, row_ptr, col, deg=None):
        self.n = int(n)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        if deg is None:
            deg = np.diff(self.row_ptr.astype(np.int64))
        self.deg = np.ascontiguousarray(deg, dtype=np.uint32)

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def edges(self):
        """Undirected edges (u < v) in CSR order."""
        u = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.row_ptr.astype(np.int64)))
        m = u < self.col
        return u[m], self.col[m]

    @staticmethod
    def from_edges(n, edges):
        """Dedup, drop self-loops, symmetrise, sort rows (the loader's rules)."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        if len(e):
            e = e[e[:, 0] != e[:, 1]]
            e = np.sort(e, axis=1)
            e = np.unique(e, axis=0)
        a = np.concatenate([e[:, 0], e[:, 1]]) if len(e) else np.zeros(0, np.int64)
        b = np.concatenate([e[:, 1], e[:, 0]]) if len(e) else np.zeros(0, np.int64)
        order = np.lexsort((b, a))
        a, b = a[order], b[order]
        deg = np.bincount(a, minlength=n).astype(np.uint32)
        rp = np.zeros(n + 1, dtype=np.uint32)
        np.cumsum(deg, out=rp[1:])
        return Csr(n, rp, b.astype(np.uint32), deg)


def load_mtx(path):
    L = lib()
    n = C.c_uint32()
    nnz = C.c_uint64()
    h = L.orc_load_mtx(path.encode(), C.byref(n), C.byref(nnz))
    if not h:
        raise OSError(path)
    rp = np.zeros(n.value + 1, np.uint32)
    col = np.zeros(max(nnz.value, 1), np.uint32)
    deg = np.zeros(max(n.value, 1), np.uint32)
    L.orc_csr_get(h, rp, col, deg)
    L.orc_csr_free(h)
    return Csr(n.value, rp, col[:nnz.value], deg[:n.value])


def init_stream(g):
    m = g.nnz // 2
    u = np.zeros(max(m, 1), np.uint32)
    v = np.zeros(max(m, 1), np.uint32)
    k = lib().orc_init_stream(g.n, g.row_ptr, g.col, g.deg, u, v)
    return u[:k], v[:k]


def bench_stream(g, size, pct, seed):
    u = np.zeros(max(size, 1), np.uint32)
    v = np.zeros(max(size, 1), np.uint32)
    t = np.zeros(max(size, 1), np.uint8)
    lib().orc_bench_stream(g.n, g.row_ptr, g.col, size, pct, seed, u, v, t)
    return u[:size], v[:size], t[:size]


def records(u, v, t=None):
    """16-byte EdgeUpdate records (u32 u, u32 v, u8 type, 7 zero bytes)."""
    r = np.zeros((len(u), 4), dtype=np.uint32)
    r[:, 0] = u
    r[:, 1] = v
    if t is not None:
        r[:, 2] = t
    return r


class Engine:
    """The oracle engine on one graph (sequential, single-warp semantics)."""

    def __init__(self, g, cfg=None, threads=1):
        self.g = g
        self.cfg = cfg or default_config()
        self.h = lib().orc_engine_create(g.n, g.row_ptr, g.col, g.deg, C.byref(self.cfg))
        if not self.h:
            raise ValueError("oracle engine create failed")
        lib().orc_set_threads(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_engine_destroy(self.h)
            self.h = None

    def set_flags(self, flags):
        self.cfg.flags = flags
        lib().orc_set_flags(self.h, flags)

    def submit(self, u, v, t=None, max_epochs=4096):
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        t = np.ascontiguousarray(t if t is not None else np.zeros(len(u)), np.uint8)
        st = (Stats * max_epochs)()
        k = lib().orc_submit(self.h, u, v, t, len(u), st, max_epochs)
        return [st[i].as_dict() for i in range(min(k, max_epochs))]

    def apply(self, u, v, t=None):
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        t = np.ascontiguousarray(t if t is not None else np.zeros(len(u)), np.uint8)
        lib().orc_apply(self.h, u, v, t, len(u))

    def _stat_call(self, name):
        st = Stats()
        getattr(lib(), name)(self.h, C.byref(st))
        return st.as_dict()

    def epoch(self):
        return self._stat_call("orc_epoch")

    def finish(self):
        return self._stat_call("orc_finish")

    def tier1(self):
        return self._stat_call("orc_tier1")

    def tier2(self):
        return self._stat_call("orc_tier2")

    def tier1_sweep(self):
        return lib().orc_tier1_sweep(self.h)

    def rebuild_bitmaps(self):
        lib().orc_rebuild_bitmaps(self.h)

    def compact(self):
        lib().orc_compact(self.h)

    def words(self):
        w = np.zeros(self.g.n, np.uint32)
        lib().orc_get_words(self.h, w)
        return w

    def set_words(self, w):
        lib().orc_set_words(self.h, np.ascontiguousarray(w, np.uint32))

    def colors(self):
        return (self.words() & 0xFFFF).astype(np.uint16)

    def snapshot(self):
        s = np.zeros(self.g.n, np.uint16)
        lib().orc_get_snapshot(self.h, s)
        return s

    def bitmaps(self):
        b = np.zeros(self.g.n, np.uint64)
        lib().orc_get_bitmaps(self.h, b)
        return b

    def set_bitmaps(self, b):
        lib().orc_set_bitmaps(self.h, np.ascontiguousarray(b, np.uint64))

    def hist(self):
        k = lib().orc_num_hd(self.h)
        h = np.zeros(max(k * 1024, 1), np.uint32)
        lib().orc_get_hist(self.h, h)
        return h[:k * 1024].reshape(k, 1024)

    def set_hist(self, h):
        lib().orc_set_hist(self.h, np.ascontiguousarray(h, np.uint32).ravel())

    def csr(self):
        s = lib().orc_num_slots(self.h)
        rp = np.zeros(self.g.n + 1, np.uint32)
        col = np.zeros(max(s, 1), np.uint32)
        lib().orc_get_csr(self.h, rp, col)
        return rp, col[:s]

    def set_col(self, col):
        lib().orc_set_col(self.h, np.ascontiguousarray(col, np.uint32))

    def find_color(self, v):
        return lib().orc_find_color(self.h, v)

    def overflow_count(self):
        return lib().orc_overflow_count(self.h)


def validate_edges(n, eu, ev, colors):
    a = C.c_uint64()
    b = C.c_uint64()
    lib().orc_validate_edges(n, np.ascontiguousarray(eu, np.uint32), np.ascontiguousarray(ev, np.uint32),
                             len(eu), np.ascontiguousarray(colors, np.uint16), C.byref(a), C.byref(b))
    return a.value, b.value


def static_greedy(g):
    return lib().orc_static_greedy(g.n, g.row_ptr, g.col)


def live_edges_after(g, u, v, t):
    """Live undirected edge set after applying a stream (u<v keys) to base graph g."""
    bu, bv = g.edges()
    live = set((bu.astype(np.uint64) << np.uint64(32) | bv.astype(np.uint64)).tolist())
    for a, b, k in zip(u.tolist(), v.tolist(), t.tolist()):
        key = (min(a, b) << 32) | max(a, b)
        if k == 0:
            live.add(key)
        elif k == 1:
            live.discard(key)
    keys = np.array(sorted(live), dtype=np.uint64)
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
