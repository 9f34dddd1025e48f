// sgc_device.hpp -- device-side layout and wave64 building blocks of the
// StreamGC hot path for gfx950 (CDNA4).
//
// One 64-lane wavefront serves one edge update.  Row scans are lane-parallel
// (64 slots per load, coalesced); "first free or equal slot" and "first
// match" are one __ballot each; the packed vertex word is claimed with a
// 32-bit device-scope CAS; colour selection maps the 64-bit bitmap to a ffs
// and the 1024-bin histogram to 16 ballots (first zero) or a wave min
// reduction.  All scalar decisions are wave-uniform, so a single wave
// replaying a stream in order computes exactly what the reference's lane 0
// computes (A/src/kernels/local_update.cu:81-123).
//
// Shared mutable state (slots, vertex words, bitmaps, histograms) is read
// with relaxed device-scope atomic loads and written with device-scope
// atomics; lanes of one wave need no extra ordering (in-order completion
// per wave), and all concurrency between waves is resolved by CAS, exactly
// as optimistic as the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgc {

constexpr uint32_t INVALID = 0xFFFFFFFFu;   // empty slot / no histogram row (A/include/types.h:24-25)
constexpr uint32_t HIST_BINS = 1024;        // A/include/partition.cuh:6-9
constexpr uint32_t WAVE = 64;
// Row scans issue ROW_UNROLL 64-slot loads before consuming any of them, so a
// long row costs ROW_UNROLL times fewer memory round trips.
constexpr uint32_t ROW_UNROLL = 4;
constexpr int MAX_CAS_RETRIES = 64;          // recolour CAS retries before leaving it to repair
// Ablation flags carried by the descriptors (the values of SGC_ABL_* in streamgc.h).
constexpr uint32_t ABL_FORCE_BITMAP = 1u << 5, ABL_NO_OVERFLOW = 1u << 7, ABL_LIVE_DEGREE = 1u << 8;

// Per-graph u32 counters (device).
enum : uint32_t {
    C_OVF = 0,          // overflow reservations (may run past capacity, like the reference)
    C_OVF_RAW,          // C_OVF as read by the last drain
    C_OVF_SNAP,         // min(C_OVF_RAW, capacity): entries the drain retries
    C_DRAIN_FAILED,
    C_CONFLICTS,        // Tier-1 sweep count
    C_T2BAD,            // row slots conflicting / uncoloured (tier-2 check)
    C_VIOL,             // validator: conflicting slots
    C_UNCOL,            // validator: uncoloured vertices with a live slot
    C_MAXC,             // validator: max colour
    C_ASYM,             // overflow events ever (edges whose two rows may disagree); never reset
    C_T2TODO,           // device Tier-2: the last pass found a conflict
    C_T2FIXED,          //   vertices the last pass recoloured
    C_T2TOTAL,          //   vertices recoloured by this repair
    C_REVOVF,           //   reverse-only neighbours that did not fit the scratch list
    C_COUNT = 16
};

// Dirty marks.  Each maintenance sweep visits only the vertices marked for
// it since it last ran, and clears their marks: compaction (M_ROW: the row
// lost a slot), rebuild (M_BITS: the bitmap or histogram may be inexact), the
// Tier-1 check (M_CHECK) and the Tier-2 check (M_T2: a colour in the closed
// neighbourhood changed in a way that can change the vertex's verdict).  Each
// kind is its own byte array (stride `dstride`), so a mark is a plain byte
// store of 1 and concurrent marks never lose one another.  The endpoints of
// an overflowed edge, whose two rows may disagree, are sticky (M_STICKY) and
// visited by every sweep.
// Two more byte arrays serve the device Tier-2: M_TODO (vertices to recolour
// in the current pass) and M_REV (vertices with a neighbour found only in the
// neighbour's row).
enum : uint32_t { M_ROW = 0, M_BITS = 1, M_CHECK = 2, M_T2 = 3, M_STICKY = 4, M_TODO = 5, M_REV = 6, M_ARRAYS = 7 };
enum : uint32_t {
    D_ROW = 1u << M_ROW, D_BITS = 1u << M_BITS, D_CHECK = 1u << M_CHECK, D_T2 = 1u << M_T2,
    D_STATUS = D_CHECK | D_T2, D_ALL = D_ROW | D_BITS | D_CHECK | D_T2
};

// Launch-wide u64 totals (algorithmic counters).
enum : uint32_t {
    T_UPDATES = 0, T_SCAN, T_HD, T_RECOLOR, T_RC_CAP, T_RC_LIVE, T_OVF_EVENTS, T_LIVE_SLOTS, T_GROUPED,
    T_SWEEP_ROWS,                  // 4 counters: rows visited by compact, rebuild, tier-1, tier-2 check
    T_SWEEP_SLOTS = T_SWEEP_ROWS + 4,  // 4 counters: row slots those sweeps read
    T_COUNT = T_SWEEP_SLOTS + 4
};

struct GraphDev {
    uint32_t* word;           // n packed (version << 16) | colour
    const uint32_t* deg0;     // n loaded-graph degrees (priority)
    const uint32_t* row_ptr;  // n + 1 slot offsets
    uint32_t* col;            // S slots, INVALID = free
    uint32_t* fill;           // n: slots of the row that may hold a neighbour (the rest are INVALID)
    uint64_t* bitmap;         // n neighbour-colour bitmaps (bit c-1 = colour c)
    const uint32_t* hd_map;   // n histogram row or INVALID
    const uint32_t* hubs;     // nhd vertices in histogram mode (inverse of hd_map)
    uint32_t* hist;           // nhd * 1024 counters
    uint32_t* ovf;            // 2 * ovf_cap (vertex, neighbour) pairs
    uint32_t* ctr;            // C_COUNT counters
    uint16_t* snap_write;     // snapshot buffer the next publish writes
    uint8_t* dirty;           // M_ARRAYS mark arrays of dstride bytes (null = no marking, full sweeps)
    uint32_t n;
    uint32_t ovf_cap;
    uint32_t nhd;
    uint32_t dstride;         // n rounded up to a multiple of 4
    uint32_t flags;           // SGC_ABL_* ablation flags of the pool
};

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t bcast0(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ uint64_t wave_min64(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    return x;
}
__device__ __forceinline__ uint32_t wave_sum32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ uint32_t wave_max32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t y = __shfl_xor(x, o);
        x = y > x ? y : x;
    }
    return x;
}

// ---------------------------------------------------------------- device-scope memory access
// Descriptor pointers are generic; casting them to the global address space
// lets the compiler emit global_* instructions (SGPR base, no LDS check).
#define SGC_G __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ SGC_G T* gp(T* p) {
    return (SGC_G T*)p;
}
template <class T>
__device__ __forceinline__ SGC_G T* gp(const T* p) {
    return (SGC_G T*)const_cast<T*>(p);
}
template <class T>
__device__ __forceinline__ T ld(const T* p) {
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st(T* p, T v) {
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t cas32(uint32_t* p, uint32_t expect, uint32_t desired) {
    __hip_atomic_compare_exchange_strong(gp(p), &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expect;  // the value found
}
__device__ __forceinline__ void or64(uint64_t* p, uint64_t v) {
    __hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add64(uint64_t* p, uint64_t v) {
    __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t max32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Mark v for the sweeps in `bits`.
__device__ __forceinline__ void mark(const GraphDev& g, uint32_t v, uint32_t bits) {
    if (!g.dirty) return;
    SGC_G uint8_t* d = gp(g.dirty) + v;
    const size_t st = g.dstride;
    if (bits & D_ROW) d[M_ROW * st] = 1;
    if (bits & D_BITS) d[M_BITS * st] = 1;
    if (bits & D_CHECK) d[M_CHECK * st] = 1;
    if (bits & D_T2) d[M_T2 * st] = 1;
}
__device__ __forceinline__ void mark_sticky(const GraphDev& g, uint32_t v) {
    if (!g.dirty) return;
    mark(g, v, D_ALL);
    gp(g.dirty)[(size_t)M_STICKY * g.dstride + v] = 1;
}

__device__ __forceinline__ uint16_t color_of(uint32_t w) { return (uint16_t)(w & 0xFFFFu); }
__device__ __forceinline__ uint32_t bump(uint32_t w, uint16_t c) {     // pack(c, version + 1), u16 wrap
    return (((w >> 16) + 1u) << 16) | c;
}
// pi(u) < pi(v) with pi = (deg0 << 32) | id   (A/include/priority.cuh:8-23)
__device__ __forceinline__ bool pi_less(const GraphDev& g, uint32_t a, uint32_t b) {
    uint64_t pa = ((uint64_t)gp(g.deg0)[a] << 32) | a;
    uint64_t pb = ((uint64_t)gp(g.deg0)[b] << 32) | b;
    return pa < pb;
}

// Per-wave counters, flushed once per wave.
struct WaveAcc {
    uint64_t updates = 0, scan = 0, hd = 0, recolor = 0, rc_cap = 0, rc_live = 0, ovf = 0;
    __device__ __forceinline__ void flush(uint64_t* totals) {
        if (!totals || lane() != 0) return;
        if (updates) add64(&totals[T_UPDATES], updates);
        if (scan) add64(&totals[T_SCAN], scan);
        if (hd) add64(&totals[T_HD], hd);
        if (recolor) add64(&totals[T_RECOLOR], recolor);
        if (rc_cap) add64(&totals[T_RC_CAP], rc_cap);
        if (rc_live) add64(&totals[T_RC_LIVE], rc_live);
        if (ovf) add64(&totals[T_OVF_EVENTS], ovf);
    }
};

// ---------------------------------------------------------------- colour selection
// A/include/color_select.cuh:10-53.  Bitmap: lowest clear bit.  All 64 set:
// lane-parallel gather of row colours 1..255 into a 256-bit set, first free,
// else 256.
__device__ __forceinline__ uint16_t fas_bitmap(const GraphDev& g, uint32_t x) {
    uint64_t freeb = ~ld(&g.bitmap[x]);
    if (freeb) return (uint16_t)(__builtin_ctzll(freeb) + 1);
    const uint32_t s = gp(g.row_ptr)[x], e = s + ld(&g.fill[x]);
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            if (c[u] > 0 && c[u] < 256) {
                uint64_t b = 1ull << (c[u] & 63);
                uint32_t q = c[u] >> 6;
                m0 |= q == 0 ? b : 0;
                m1 |= q == 1 ? b : 0;
                m2 |= q == 2 ? b : 0;
                m3 |= q == 3 ? b : 0;
            }
        }
    }
    m0 = wave_or64(m0) | 1ull;  // colour 0 is not a colour
    m1 = wave_or64(m1);
    m2 = wave_or64(m2);
    m3 = wave_or64(m3);
    if (~m0) return (uint16_t)__builtin_ctzll(~m0);
    if (~m1) return (uint16_t)(64 + __builtin_ctzll(~m1));
    if (~m2) return (uint16_t)(128 + __builtin_ctzll(~m2));
    if (~m3) return (uint16_t)(192 + __builtin_ctzll(~m3));
    return 256;
}

// A/include/color_select.cuh:59-89.  First zero bin in ascending order (one
// ballot per 64 bins), else the minimum count with the lowest index.
__device__ __forceinline__ uint16_t fas_hist(const GraphDev& g, uint32_t hd) {
    const uint32_t* h = g.hist + (size_t)hd * HIST_BINS;
    uint64_t best = ~0ull;
    for (uint32_t k = 0; k < HIST_BINS; k += WAVE) {
        uint32_t cnt = ld(&h[k + lane()]);
        uint64_t z = ballot(cnt == 0);
        if (z) return (uint16_t)(k + __builtin_ctzll(z) + 1);
        uint64_t key = ((uint64_t)cnt << 32) | (k + lane());
        best = key < best ? key : best;
    }
    best = wave_min64(best);
    return (uint16_t)((uint32_t)best + 1);
}

// A/include/color_select.cuh:93-115
// Histogram selection for a hub unless every vertex is forced to its bitmap
// (ablation, A/include/color_select.cuh:98-100).
__device__ __forceinline__ bool select_by_hist(const GraphDev& g, uint32_t x) {
    return gp(g.hd_map)[x] != INVALID && !(g.flags & ABL_FORCE_BITMAP);
}
__device__ __forceinline__ uint16_t find_color(const GraphDev& g, uint32_t x) {
    return select_by_hist(g, x) ? fas_hist(g, gp(g.hd_map)[x]) : fas_bitmap(g, x);
}

// A/include/color_select.cuh:160-183 (lane 0 only: one neighbour)
__device__ __forceinline__ void on_new_neighbor(const GraphDev& g, uint32_t x, uint16_t c) {
    if (c == 0 || lane() != 0) return;
    if (c <= 64) or64(&g.bitmap[x], 1ull << (c - 1));
    uint32_t hd = gp(g.hd_map)[x];
    if (hd != INVALID && c <= HIST_BINS) add32(&g.hist[(size_t)hd * HIST_BINS + c - 1], 1u);
}
// A/include/color_select.cuh:188-209 (bitmap bits stay set until a rebuild)
__device__ __forceinline__ void on_removed_neighbor(const GraphDev& g, uint32_t x, uint16_t c) {
    if (c == 0 || lane() != 0) return;
    uint32_t hd = gp(g.hd_map)[x];
    if (hd != INVALID && c <= HIST_BINS) add32(&g.hist[(size_t)hd * HIST_BINS + c - 1], 0xFFFFFFFFu);
}
// A/include/color_select.cuh:120-157 (per lane: each lane owns one neighbour)
// hd = hd_map[nb].  A bitmap row keeps its old colour bit (inexact until a
// rebuild); a histogram stays exact.
__device__ __forceinline__ void on_neighbor_recolor(const GraphDev& g, uint32_t nb, uint32_t hd, uint16_t oldc,
                                                    uint16_t newc) {
    if (newc != 0 && newc <= 64) or64(&g.bitmap[nb], 1ull << (newc - 1));
    if (hd != INVALID) {
        uint32_t* h = g.hist + (size_t)hd * HIST_BINS;
        if (oldc != 0 && oldc <= HIST_BINS) add32(&h[oldc - 1], 0xFFFFFFFFu);
        if (newc != 0 && newc <= HIST_BINS) add32(&h[newc - 1], 1u);
    }
}

// ---------------------------------------------------------------- slack-CSR rows
// A/src/kernels/local_update.cu:12-39.  First slot that is free (claimed by
// CAS) or already holds nb.  Returns the slots scanned; a full row appends to
// the overflow buffer (entries past capacity are dropped, the counter keeps
// counting) or, with overflow disabled, reports failure.
// An insert that lands in a hole may duplicate a copy further along the row
// (the first-free-or-equal scan meets the hole first,
// A/src/kernels/local_update.cu:12-40); a later delete then removes one copy
// per row and the two rows disagree.  Such endpoints are maintained like an
// overflowed edge's (sticky, and counted in C_ASYM).
__device__ __forceinline__ void flag_duplicate(const GraphDev& g, uint32_t x, uint32_t nb, uint32_t from,
                                               uint32_t end) {
    bool dup = false;
    for (uint32_t base = from; base < end && !dup; base += WAVE) {
        const uint32_t j = base + lane();
        dup = ballot(j < end && ld(&g.col[j]) == nb) != 0;
    }
    if (dup && lane() == 0) {
        add32(&g.ctr[C_ASYM], 1u);
        mark_sticky(g, x);
        mark_sticky(g, nb);
    }
}

__device__ __forceinline__ uint32_t row_insert(const GraphDev& g, uint32_t x, uint32_t nb, bool use_overflow, bool* failed,
                               WaveAcc* acc) {
    const uint32_t s = gp(g.row_ptr)[x], e = gp(g.row_ptr)[x + 1];
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t val[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            val[u] = j < e ? ld(&g.col[j]) : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t b = base + u * WAVE;
            const bool in = b + lane() < e;
            uint64_t m = ballot(in && (val[u] == INVALID || val[u] == nb));
            while (m) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                const uint32_t v0 = __shfl(val[u], (int)k);
                if (v0 == nb) return b + k - s + 1;
                uint32_t old = 0, prev_fill = 0;
                if (lane() == 0) {
                    old = cas32(&g.col[b + k], INVALID, nb);
                    if (old == INVALID) prev_fill = max32(&g.fill[x], b + k - s + 1);
                }
                old = bcast0(old);
                prev_fill = bcast0(prev_fill);
                if (old == INVALID && prev_fill > b + k - s + 1 && nb < g.n)
                    flag_duplicate(g, x, nb, b + k + 1, s + prev_fill);
                if (old == INVALID || old == nb) return b + k - s + 1;
                m &= m - 1;
            }
        }
    }
    if (use_overflow) {
        if (lane() == 0) {
            add32(&g.ctr[C_ASYM], 1u);
            if (!(g.flags & ABL_NO_OVERFLOW)) {   // the ablation drops the edge unbuffered and uncounted
                uint32_t pos = add32(&g.ctr[C_OVF], 1u);
                if (pos < g.ovf_cap) {
                    st(&g.ovf[2 * (size_t)pos], x);
                    st(&g.ovf[2 * (size_t)pos + 1], nb);
                }
            }
            // the two rows of this edge may now disagree: always maintain both
            mark_sticky(g, x);
            mark_sticky(g, nb);
        }
        if (acc) acc->ovf++;
    } else if (failed) {
        *failed = true;
    }
    return e - s;
}

// v changed colour oldc -> newc without telling its neighbours (first
// colouring, Tier-1): every neighbour's bitmap or histogram may be stale, and
// a neighbour whose colour is oldc, newc or none may have a new verdict.
__device__ __forceinline__ void mark_neighbors(const GraphDev& g, uint32_t v, uint16_t oldc, uint16_t newc) {
    if (!g.dirty) return;
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            if (nb[u] != INVALID && nb[u] < g.n)
                mark(g, nb[u], D_BITS | ((c[u] == oldc || c[u] == newc || c[u] == 0) ? D_STATUS : 0u));
    }
}

// A/src/kernels/local_update.cu:42-56: the first slot holding nb becomes free.
__device__ __forceinline__ uint32_t row_remove(const GraphDev& g, uint32_t x, uint32_t nb) {
    const uint32_t s = gp(g.row_ptr)[x], e = s + ld(&g.fill[x]);
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t val[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            val[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t b = base + u * WAVE;
            uint64_t m = ballot(b + lane() < e && val[u] == nb);
            while (m) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                uint32_t old = 0;
                if (lane() == 0) old = cas32(&g.col[b + k], nb, INVALID);
                old = bcast0(old);
                if (old == nb) return b + k - s + 1;
                m &= m - 1;
            }
        }
    }
    return e - s;
}

// A/src/kernels/recolor.cu:50-82: every live neighbour learns the new colour.
__device__ __forceinline__ uint32_t propagate(const GraphDev& g, uint32_t x, uint16_t oldc, uint16_t newc) {
    if (oldc == newc) return 0;
    const uint32_t s = gp(g.row_ptr)[x], e = s + ld(&g.fill[x]);
    uint32_t live = 0;
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t nb[ROW_UNROLL], hd[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const bool ok = nb[u] != INVALID && nb[u] < g.n;
            hd[u] = ok ? gp(g.hd_map)[nb[u]] : INVALID;
            c[u] = (ok && g.dirty) ? color_of(ld(&g.word[nb[u]])) : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            if (nb[u] == INVALID || nb[u] >= g.n) continue;
            on_neighbor_recolor(g, nb[u], hd[u], oldc, newc);
            mark(g, nb[u], (hd[u] == INVALID ? D_BITS : 0u) |
                               ((c[u] == oldc || c[u] == newc || c[u] == 0) ? D_STATUS : 0u));
            live++;
        }
    }
    if (lane() == 0) mark(g, x, D_STATUS);
    return wave_sum32(live);
}

// A/src/kernels/recolor.cu:87-141.  Select, then publish pack(c', version+1)
// with one CAS on the packed word; on interference re-select and retry.
__device__ __forceinline__ bool recolor(const GraphDev& g, uint32_t x, WaveAcc& acc) {
    for (int attempt = 0; attempt < MAX_CAS_RETRIES; attempt++) {
        const uint32_t w = bcast0(ld(&g.word[x]));
        const uint16_t oldc = color_of(w);
        const uint16_t nc = find_color(g, x);
        if (nc == oldc) return false;
        const uint32_t nw = bump(w, nc);
        uint32_t got = 0;
        if (lane() == 0) got = cas32(&g.word[x], w, nw);
        got = bcast0(got);
        if (got == w) {
            uint32_t live = propagate(g, x, oldc, nc);
            acc.recolor++;
            acc.rc_cap += ld(&g.fill[x]);
            acc.rc_live += live;
            return true;
        }
    }
    return false;
}

// A/src/kernels/recolor.cu:145-163
__device__ __forceinline__ void try_reduce(const GraphDev& g, uint32_t x, uint16_t freed, WaveAcc& acc) {
    const uint16_t cur = color_of(bcast0(ld(&g.word[x])));
    if (cur == 0 || freed >= cur) return;
    if (find_color(g, x) < cur) recolor(g, x, acc);
}

// Colour an uncoloured endpoint: pack(c, 1) (A/src/kernels/local_update.cu:90-101).
// The reference tells only the update's other endpoint about the new colour.
__device__ __forceinline__ void color_fresh(const GraphDev& g, uint32_t x, uint32_t w) {
    const uint16_t c = find_color(g, x);
    if (lane() == 0) cas32(&g.word[x], w, (1u << 16) | c);
    mark_neighbors(g, x, 0, c);
}

// One update, A/src/kernels/local_update.cu:71-146.
__device__ __forceinline__ void apply_update(const GraphDev& g, uint32_t u, uint32_t v, uint32_t type, WaveAcc& acc) {
    if (u >= g.n || v >= g.n) return;  // hash miss: skipped (:77)
    if (type <= 1 && lane() == 0) {
        // an insert fills the first free slot (the row stays packed); a delete leaves a hole
        const uint32_t bits = type == 1 ? D_ALL : (D_BITS | D_STATUS);
        mark(g, u, bits);
        mark(g, v, bits);
    }
    if (type == 0) {
        uint32_t p = row_insert(g, u, v, true, nullptr, &acc);
        p += row_insert(g, v, u, true, nullptr, &acc);
        uint32_t wu = bcast0(ld(&g.word[u]));
        if (color_of(wu) == 0) color_fresh(g, u, wu);
        const uint32_t wv = bcast0(ld(&g.word[v]));
        if (color_of(wv) == 0) {
            on_new_neighbor(g, v, color_of(bcast0(ld(&g.word[u]))));
            color_fresh(g, v, wv);
        }
        const uint16_t cu = color_of(bcast0(ld(&g.word[u])));
        const uint16_t cv = color_of(bcast0(ld(&g.word[v])));
        on_new_neighbor(g, u, cv);
        on_new_neighbor(g, v, cu);
        const uint16_t cu2 = color_of(bcast0(ld(&g.word[u])));
        const uint16_t cv2 = color_of(bcast0(ld(&g.word[v])));
        if (ballot(cu2 == cv2 && cu2 != 0)) recolor(g, pi_less(g, u, v) ? u : v, acc);
        acc.updates++;
        acc.scan += p;
        acc.hd += (gp(g.hd_map)[u] != INVALID) + (gp(g.hd_map)[v] != INVALID);
    } else if (type == 1) {
        const uint16_t fu = color_of(bcast0(ld(&g.word[v])));
        const uint16_t fv = color_of(bcast0(ld(&g.word[u])));
        uint32_t p = row_remove(g, u, v);
        p += row_remove(g, v, u);
        on_removed_neighbor(g, u, fu);
        on_removed_neighbor(g, v, fv);
        try_reduce(g, u, fu, acc);
        try_reduce(g, v, fv, acc);
        acc.updates++;
        acc.scan += p;
        acc.hd += (gp(g.hd_map)[u] != INVALID) + (gp(g.hd_map)[v] != INVALID);
    }
}

}  // namespace sgc
