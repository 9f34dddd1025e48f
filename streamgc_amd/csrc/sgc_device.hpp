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

#include <type_traits>

namespace sgc {

constexpr uint32_t INVALID = 0xFFFFFFFFu;   // empty slot / no histogram row (A/include/types.h:24-25)
constexpr uint32_t HIST_BINS = 1024;        // A/include/partition.cuh:6-9
constexpr uint32_t WAVE = 64;
// Row scans issue ROW_UNROLL 64-slot loads before consuming any of them, so a
// long row costs ROW_UNROLL times fewer memory round trips.
constexpr uint32_t ROW_UNROLL = 4;
// An insert's or delete's scan of its two rows loads ROW_UNROLL 64-slot
// chunks of each per round trip, and SCAN_CHUNKS of the one row left once the
// other is done: a hub row's long search takes SCAN_CHUNKS x 64 slots per trip.
// 8 rather than 16: every trip runs the code of all SCAN_CHUNKS chunks, and
// the shorter trips made the per-update kernel 8% faster on a 4M-vertex
// power-law graph although hub rows take twice as many of them.
constexpr uint32_t SCAN_CHUNKS = 8;
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
    C_LOST,             // half-edges lost for good: overflow entries past capacity, failed drains, and full
                        // rows under the no-overflow ablation; never reset
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
    T_SWEEP_ROWS,                  // 5 counters: rows visited by compact, rebuild, tier-1 (parallel),
                                   // tier-2 check, tier-1 (vertex order)
    T_SWEEP_SLOTS = T_SWEEP_ROWS + 5,  // 5 counters: row slots those sweeps read
    T_COUNT = T_SWEEP_SLOTS + 5
};
// The totals are striped: T_STRIPES copies of the counters, T_STRIDE u64
// apart (two cache lines), a block adding into copy blockIdx % T_STRIPES;
// the host sums the copies.  Thousands of waves adding into one address
// serialise on it and hold up every other atomic of its memory channel.
constexpr uint32_t T_STRIPES = 256, T_STRIDE = 32;
static_assert(T_COUNT <= T_STRIDE, "totals stripe too small");

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
    uint32_t dstride;         // n rounded up to a multiple of 16
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
__device__ __forceinline__ void tot_add(uint64_t* totals, uint32_t idx, uint64_t v) {
    add64(&totals[(size_t)(blockIdx.x % T_STRIPES) * T_STRIDE + idx], v);
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
        if (updates) tot_add(totals, T_UPDATES, updates);
        if (scan) tot_add(totals, T_SCAN, scan);
        if (hd) tot_add(totals, T_HD, hd);
        if (recolor) tot_add(totals, T_RECOLOR, recolor);
        if (rc_cap) tot_add(totals, T_RC_CAP, rc_cap);
        if (rc_live) tot_add(totals, T_RC_LIVE, rc_live);
        if (ovf) tot_add(totals, T_OVF_EVENTS, ovf);
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

// A/include/color_select.cuh:160-183 (lane 0 only: one neighbour); hd = hd_map[x]
__device__ __forceinline__ void on_new_neighbor(const GraphDev& g, uint32_t x, uint32_t hd, uint16_t c) {
    if (c == 0 || lane() != 0) return;
    if (c <= 64) or64(&g.bitmap[x], 1ull << (c - 1));
    if (hd != INVALID && c <= HIST_BINS) add32(&g.hist[(size_t)hd * HIST_BINS + c - 1], 1u);
}
// A/include/color_select.cuh:188-209 (bitmap bits stay set until a rebuild)
__device__ __forceinline__ void on_removed_neighbor(const GraphDev& g, uint32_t x, uint32_t hd, uint16_t c) {
    if (c == 0 || lane() != 0) return;
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
// Whether slots [from, end) hold nb (wave-uniform).
__device__ __forceinline__ bool row_holds(const GraphDev& g, uint32_t nb, uint32_t from, uint32_t end) {
    constexpr uint32_t CH = SCAN_CHUNKS;   // 64-slot chunks per round trip
    bool dup = false;
    for (uint32_t base = from; base < end && !dup; base += CH * WAVE) {
        uint32_t val[CH];
#pragma unroll
        for (uint32_t q = 0; q < CH; q++) {
            const uint32_t j = base + q * WAVE + lane();
            val[q] = j < end ? ld(&g.col[j]) : INVALID;
        }
        bool hit = false;
#pragma unroll
        for (uint32_t q = 0; q < CH; q++) hit |= val[q] == nb;
        dup = ballot(hit) != 0;
    }
    return dup;
}
__device__ __forceinline__ void mark_duplicate(const GraphDev& g, uint32_t x, uint32_t nb) {
    if (lane() != 0) return;
    add32(&g.ctr[C_ASYM], 1u);
    mark_sticky(g, x);
    mark_sticky(g, nb);
}
__device__ __forceinline__ void flag_duplicate(const GraphDev& g, uint32_t x, uint32_t nb, uint32_t from,
                                               uint32_t end) {
    if (row_holds(g, nb, from, end)) mark_duplicate(g, x, nb);
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
            } else {
                add32(&g.ctr[C_LOST], 1u);
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

// Both rows of one update, scanned in the same memory round trips: an insert
// looks for the first slot that is free or already holds the neighbour
// (A/src/kernels/local_update.cu:12-39), a delete for the first slot holding
// it (A/src/kernels/local_update.cu:42-56), row 0 and row 1 each on its own.
// A round trip loads ROW_UNROLL x 64 slots of each row still searching
// (SCAN_CHUNKS x 64 when one row is left).  The rows' matches are then claimed together, lane 0
// for row 0 and lane 1 for row 1: one CAS each, and for an insert the raise
// of the row's fill to the claimed position (harmless when the CAS loses:
// the slot's winner raised it as far).  A row whose CAS lost resumes its scan
// after that slot.  The two rows must differ; a single wave applying a
// stream in order claims exactly the slots of the two scans one after the
// other.
// Out: pos = 1-based position of the row's slot (0: none before the end);
// pf = the fill before an insert's raise when this wave wrote the slot, else 0;
// had = an insert found the neighbour in the row already; free = the first
// free slot a delete's scan passed (absolute index, INVALID: none).
template <bool INSERT>
__device__ __forceinline__ void scan_pair(const GraphDev& g, bool run0, uint32_t x0, uint32_t nb0, uint32_t s0,
                                          uint32_t end0, bool run1, uint32_t x1, uint32_t nb1, uint32_t s1,
                                          uint32_t end1, uint32_t& pos0, uint32_t& pos1, uint32_t& pf0,
                                          uint32_t& pf1, bool& had0, bool& had1, uint32_t& free0,
                                          uint32_t& free1) {
    constexpr uint32_t CH = SCAN_CHUNKS;   // 64-slot chunks per round trip
    uint32_t b0 = s0, b1 = s1;
    pos0 = pos1 = pf0 = pf1 = 0;
    had0 = had1 = false;
    free0 = free1 = INVALID;
    run0 = run0 && b0 < end0;
    run1 = run1 && b1 < end1;
    uint32_t c0, c1;   // first match of each row in this trip
    bool eq0, eq1;     // it already holds the neighbour (insert)
    // One trip over N chunks: [0, per0) are row 0's, [per0, per0 + per1) row 1's.
    auto trip = [&](auto nc, uint32_t per0, uint32_t per1) {
        constexpr uint32_t N = decltype(nc)::value;
        uint32_t val[N];
#pragma unroll
        for (uint32_t q = 0; q < N; q++) {
            const bool r1 = q >= per0;
            const uint32_t j = (r1 ? b1 + (q - per0) * WAVE : b0 + q * WAVE) + lane();
            val[q] = q < per0 + per1 && j < (r1 ? end1 : end0) ? ld(&g.col[j]) : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < N; q++) {
            const bool r1 = q >= per0;
            const uint32_t b = r1 ? b1 + (q - per0) * WAVE : b0 + q * WAVE;
            const uint32_t want = r1 ? nb1 : nb0;
            const bool in = q < per0 + per1 && b + lane() < (r1 ? end1 : end0);
            const uint64_t m = ballot(in && (val[q] == want || (INSERT && val[q] == INVALID)));
            if (!INSERT) {
                const uint64_t mf = ballot(in && val[q] == INVALID);
                if (mf && (r1 ? free1 : free0) == INVALID) {
                    if (r1) free1 = b + (uint32_t)__builtin_ctzll(mf);
                    else free0 = b + (uint32_t)__builtin_ctzll(mf);
                }
            }
            if (m && (r1 ? c1 : c0) == INVALID) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                const bool eq = __builtin_amdgcn_readlane(val[q], k) == want;
                if (r1) {
                    c1 = b + k;
                    eq1 = eq;
                } else {
                    c0 = b + k;
                    eq0 = eq;
                }
            }
        }
    };
    while (run0 || run1) {
        c0 = c1 = INVALID;
        eq0 = eq1 = false;
        uint32_t per0, per1;
        if (run0 && run1 && end0 - b0 <= WAVE && end1 - b1 <= WAVE) {
            // Both rows end within one chunk: a trip of two chunks (light
            // rows, the common case, run far fewer instructions).
            per0 = per1 = 1;
            trip(std::integral_constant<uint32_t, 2>{}, per0, per1);
        } else {
            per0 = run0 ? (run1 ? ROW_UNROLL : CH) : 0;
            per1 = run1 ? (run0 ? ROW_UNROLL : CH) : 0;
            trip(std::integral_constant<uint32_t, CH>{}, per0, per1);
        }
        if (run0 && c0 == INVALID) {
            b0 += per0 * WAVE;
            run0 = b0 < end0;
        }
        if (run1 && c1 == INVALID) {
            b1 += per1 * WAVE;
            run1 = b1 < end1;
        }
        if (INSERT && c0 != INVALID && eq0) {   // the neighbour is there already
            pos0 = c0 - s0 + 1;
            had0 = true;
            run0 = false;
            c0 = INVALID;
        }
        if (INSERT && c1 != INVALID && eq1) {
            pos1 = c1 - s1 + 1;
            had1 = true;
            run1 = false;
            c1 = INVALID;
        }
        if (c0 == INVALID && c1 == INVALID) continue;
        uint32_t old = 0, pf = 0;
        const bool mine = lane() == 0 ? c0 != INVALID : (lane() == 1 && c1 != INVALID);
        if (mine) {
            const uint32_t at = lane() == 0 ? c0 : c1;
            const uint32_t want = lane() == 0 ? nb0 : nb1;
            if (INSERT) {
                old = cas32(&g.col[at], INVALID, want);
                pf = max32(&g.fill[lane() == 0 ? x0 : x1], at - (lane() == 0 ? s0 : s1) + 1);
            } else {
                old = cas32(&g.col[at], want, INVALID);
            }
        }
        const uint32_t old0 = __builtin_amdgcn_readlane(old, 0), old1 = __builtin_amdgcn_readlane(old, 1);
        const uint32_t f0 = __builtin_amdgcn_readlane(pf, 0), f1 = __builtin_amdgcn_readlane(pf, 1);
        if (c0 != INVALID) {
            if (old0 == nb0 || (INSERT && old0 == INVALID)) {
                pos0 = c0 - s0 + 1;
                pf0 = INSERT && old0 == INVALID ? f0 : 0u;
                had0 = INSERT && old0 == nb0;
                run0 = false;
            } else {   // another wave took the slot: go on after it
                b0 = c0 + 1;
                run0 = b0 < end0;
            }
        }
        if (c1 != INVALID) {
            if (old1 == nb1 || (INSERT && old1 == INVALID)) {
                pos1 = c1 - s1 + 1;
                pf1 = INSERT && old1 == INVALID ? f1 : 0u;
                had1 = INSERT && old1 == nb1;
                run1 = false;
            } else {
                b1 = c1 + 1;
                run1 = b1 < end1;
            }
        }
    }
}

// An insert into row x (slots [s, e)) after its scan: a slot written in a hole
// may duplicate a copy further along; no slot at all is an overflow event.
// Returns the slots scanned (the position found, or the whole row).
// sticky: both endpoints are sticky already (a duplicate would change nothing).
__device__ __forceinline__ uint32_t finish_insert(const GraphDev& g, uint32_t x, uint32_t nb, uint32_t s, uint32_t e,
                                                  uint32_t pos, uint32_t pf, bool sticky, WaveAcc& acc) {
    if (pos) {
        if (pf > pos && nb < g.n && !sticky) flag_duplicate(g, x, nb, s + pos, s + pf);
        return pos;
    }
    if (lane() == 0) {
        add32(&g.ctr[C_ASYM], 1u);
        if (!(g.flags & ABL_NO_OVERFLOW)) {   // the ablation drops the edge unbuffered and uncounted
            uint32_t at = add32(&g.ctr[C_OVF], 1u);
            if (at < g.ovf_cap) {
                st(&g.ovf[2 * (size_t)at], x);
                st(&g.ovf[2 * (size_t)at + 1], nb);
            }
        } else {
            add32(&g.ctr[C_LOST], 1u);
        }
        mark_sticky(g, x);
        mark_sticky(g, nb);
    }
    acc.ovf++;
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

// A/src/kernels/recolor.cu:145-163; cur = x's colour (read by the caller
// after the removal).  early = x's bitmap as read before the removal, hd =
// hd_map[x].  Bitmap bits are only ever added between rebuilds, so when that
// earlier read already leaves no free colour below cur (fewer than 64 bits
// set), a read now would not either, and the reduction ends without it.
__device__ __forceinline__ void try_reduce(const GraphDev& g, uint32_t x, uint16_t freed, uint16_t cur, uint64_t early,
                                           uint32_t hd, WaveAcc& acc) {
    if (cur == 0 || freed >= cur) return;
    const bool by_hist = hd != INVALID && !(g.flags & ABL_FORCE_BITMAP);
    if (!by_hist && ~early && (uint32_t)__builtin_ctzll(~early) + 1u >= cur) return;
    if (find_color(g, x) < cur) recolor(g, x, acc);
}

// Colour an uncoloured endpoint: pack(c, 1) (A/src/kernels/local_update.cu:90-101).
// The reference tells only the update's other endpoint about the new colour.
__device__ __forceinline__ void color_fresh(const GraphDev& g, uint32_t x, uint32_t w) {
    const uint16_t c = find_color(g, x);
    if (lane() == 0) cas32(&g.word[x], w, (1u << 16) | c);
    mark_neighbors(g, x, 0, c);
}

// One update, A/src/kernels/local_update.cu:71-146.  The reads each step
// needs are issued together (row bounds, histogram rows, words), and both rows
// are scanned and claimed in the same round trips (scan_pair).  A vertex word
// changes only by a recolour, so a colour read before the row writes equals
// the one read after them unless another wave recoloured in between, which
// is one more interleaving the parallel mode allows anyway.
//
// Marks.  A delete that removed the neighbour leaves a hole (ROW) and, in a
// bitmap row, a stale bit (BITS); a histogram row stays exact unless the edge
// was not in it (its count then drops for a colour nobody left, BITS).  A
// delete cannot give either endpoint a conflict.  An insert that stored the edge in
// both rows and then read the colours leaves both bitmaps and histograms
// exact (a recolour of an endpoint after those reads sees the new slot and
// tells the other endpoint itself), and once it has resolved a same-colour
// edge on the spot neither endpoint's Tier-1 verdict changed.  So an insert
// marks only what it leaves inexact or unresolved: BITS where the row held the
// edge already (its histogram counts it twice), everything where an endpoint
// was uncoloured, and CHECK/T2 where the endpoints still share a colour.
// With the live-degree ablation a row change moves the vertex's own
// priority, so both kinds mark CHECK/T2 too.
//
// A wavefront applying one edge's updates in order passes a RowHint from a
// delete to the next update: if that is the same edge's re-insert, the rows'
// bounds are known and the first free-or-equal slot of each row is the first
// free slot the delete's scan passed, or the one it freed (the delete found
// the first copy there, so no copy lies before it); the insert claims those
// slots directly and scans only when another wavefront took one.
// The delete's marks wait in the hint: a re-insert into the very slot the
// delete freed leaves nothing to compact, and one that reads the neighbour's
// colour as the delete did leaves the bitmap bit right.  Any other next
// update, or the end of the edge's updates, stores them as they were.
struct RowHint {
    uint32_t u, v, su, eu, sv, ev, hu, hv, fu, fv;
    uint32_t ou, ov;   // the slots the delete freed
    uint32_t mu, mv;   // the delete's marks, held back
    uint16_t cu, cv;   // the colours the delete read
    bool stu, stv, valid;
};
__device__ __forceinline__ void hint_release(const GraphDev& g, RowHint* h) {
    if (!h || !h->valid) return;
    if (lane() == 0) {
        mark(g, h->u, h->mu);
        mark(g, h->v, h->mv);
    }
    h->valid = false;
}
__device__ __forceinline__ void apply_update(const GraphDev& g, uint32_t u, uint32_t v, uint32_t type, WaveAcc& acc,
                                             RowHint* hint = nullptr) {
    const bool hinted = type == 0 && hint && hint->valid && hint->u == u && hint->v == v && u < g.n && v < g.n;
    if (!hinted) hint_release(g, hint);
    if (u >= g.n || v >= g.n || type > 1) return;  // hash miss: skipped (:77); other kinds do nothing
    uint32_t su, eu, sv, ev, hu, hv, su0 = 0, sv0 = 0, held_u = 0, held_v = 0;
    bool stu, stv;
    if (hinted) {
        su = hint->su, eu = hint->eu, sv = hint->sv, ev = hint->ev, hu = hint->hu, hv = hint->hv;
        stu = hint->stu, stv = hint->stv;
        su0 = hint->fu, sv0 = hint->fv;
        held_u = hint->mu, held_v = hint->mv;
        if (su0 == hint->ou) held_u &= ~D_ROW;   // back in the freed slot, if it can claim it
        if (sv0 == hint->ov) held_v &= ~D_ROW;
    } else {
        su = gp(g.row_ptr)[u], eu = gp(g.row_ptr)[u + 1];
        sv = gp(g.row_ptr)[v], ev = gp(g.row_ptr)[v + 1];
        hu = gp(g.hd_map)[u], hv = gp(g.hd_map)[v];
        const SGC_G uint8_t* sticky = g.dirty ? gp(g.dirty) + (size_t)M_STICKY * g.dstride : nullptr;
        stu = sticky && sticky[u], stv = sticky && sticky[v];
    }
    if (hint) hint->valid = false;
    if ((g.flags & ABL_LIVE_DEGREE) && lane() == 0) {
        mark(g, u, D_STATUS);
        mark(g, v, D_STATUS);
    }
    const bool two = u != v;   // a self-loop's two rows are one: scanned one after the other
    uint32_t pu = 0, pv = 0, fu = 0, fv = 0, d0, d1, xu, xv;
    bool hadu = false, hadv = false, h0;
    if (type == 0) {
        bool scan_u = true, scan_v = true;
        if (hinted) {   // claim the slots the delete left (u != v: a hint needs two rows)
            uint32_t old = 0, pf = 0;
            if (lane() < 2) {
                const uint32_t at = lane() == 0 ? su0 : sv0;
                old = cas32(&g.col[at], INVALID, lane() == 0 ? v : u);
                pf = max32(&g.fill[lane() == 0 ? u : v], at - (lane() == 0 ? su : sv) + 1);
            }
            if (__builtin_amdgcn_readlane(old, 0) == INVALID) {
                pu = su0 - su + 1;
                fu = __builtin_amdgcn_readlane(pf, 0);
                scan_u = false;
            }
            if (__builtin_amdgcn_readlane(old, 1) == INVALID) {
                pv = sv0 - sv + 1;
                fv = __builtin_amdgcn_readlane(pf, 1);
                scan_v = false;
            }
        }
        if (scan_u) held_u |= hinted ? D_ROW : 0u;   // another wavefront took the slot
        if (scan_v) held_v |= hinted ? D_ROW : 0u;
        if (scan_u || scan_v) {
            uint32_t qu, qv, gu, gv;
            bool au, av;
            scan_pair<true>(g, scan_u, u, v, su, eu, two && scan_v, v, u, sv, ev, qu, qv, gu, gv, au, av, xu, xv);
            if (scan_u) pu = qu, fu = gu, hadu = au;
            if (scan_v && two) pv = qv, fv = gv, hadv = av;
            if (!two) scan_pair<true>(g, false, u, v, su, eu, true, v, u, sv, ev, d0, pv, d1, fv, h0, hadv, xu, xv);
        }
        // No duplicate check when both rows took the slots the delete left:
        // each row lost one copy of the edge and gained one, so the rows hold
        // as many copies as before (and if they disagreed, both endpoints are
        // sticky already).
        bool skip_dup = (stu && stv) || (hinted && !scan_u && !scan_v);
        if (!skip_dup && two && pu && pv) {
            // The edge is in both rows now.  Unless both endpoints are sticky,
            // its two rows held it equally often before this insert (a row
            // that gains a second copy, or keeps one the other row lost, makes
            // both endpoints sticky).  So this insert made a second copy iff
            // one row had the edge already while the other took a hole, or
            // both took a hole and the row with fewer slots past its hole
            // holds a copy there (the other row then holds one too): one short
            // row is read instead of both, which spares a hub row's tail.
            bool dup = hadu != hadv;
            if (!hadu && !hadv) {
                const uint32_t tail_u = fu > pu ? fu - pu : 0u, tail_v = fv > pv ? fv - pv : 0u;
                dup = tail_u <= tail_v ? row_holds(g, v, su + pu, su + fu) : row_holds(g, u, sv + pv, sv + fv);
            }
            if (dup) mark_duplicate(g, u, v);
            skip_dup = true;
        }
        uint32_t p = finish_insert(g, u, v, su, eu, pu, fu, skip_dup, acc);
        p += finish_insert(g, v, u, sv, ev, pv, fv, skip_dup, acc);
        if (lane() == 0) {
            if (hadu) mark(g, u, D_BITS);
            if (hadv) mark(g, v, D_BITS);
        }
        const uint32_t ru = ld(&g.word[u]), rv = ld(&g.word[v]);
        uint16_t cu = color_of(bcast0(ru)), cv = color_of(bcast0(rv));
        if (hinted) {   // the delete's marks, less what the re-insert undid
            if (cv == hint->cv) held_u &= ~D_BITS;
            if (cu == hint->cu) held_v &= ~D_BITS;
            if (lane() == 0) {
                mark(g, u, held_u);
                mark(g, v, held_v);
            }
        }
        if (cu == 0 || cv == 0) {   // first colouring of an endpoint (:90-101)
            if (lane() == 0) {
                mark(g, u, D_BITS | D_STATUS);
                mark(g, v, D_BITS | D_STATUS);
            }
            const uint32_t wu = bcast0(ld(&g.word[u]));
            if (color_of(wu) == 0) color_fresh(g, u, wu);
            const uint32_t wv = bcast0(ld(&g.word[v]));
            if (color_of(wv) == 0) {
                on_new_neighbor(g, v, hv, color_of(bcast0(ld(&g.word[u]))));
                color_fresh(g, v, wv);
            }
            cu = color_of(bcast0(ld(&g.word[u])));
            cv = color_of(bcast0(ld(&g.word[v])));
        }
        on_new_neighbor(g, u, hu, cv);
        on_new_neighbor(g, v, hv, cu);
        // The colours were read after both slots were written: a recolour of
        // an endpoint after that read sees the new slot and marks the other
        // endpoint itself, so no second read is needed.
        if (cu == cv) {   // same colour (or both uncoloured): the sweeps look again whatever happens here
            if (lane() == 0) {
                mark(g, u, D_STATUS);
                mark(g, v, D_STATUS);
            }
            if (cu != 0) recolor(g, pi_less(g, u, v) ? u : v, acc);
        }
        acc.scan += p;
    } else {
        const uint32_t ru = ld(&g.word[u]), rv = ld(&g.word[v]);
        const uint32_t lu = ld(&g.fill[u]), lv = ld(&g.fill[v]);
        const uint64_t bu = ld(&g.bitmap[u]), bv = ld(&g.bitmap[v]);   // for the reductions below
        const uint16_t cu = color_of(bcast0(ru)), cv = color_of(bcast0(rv));   // freed at v, freed at u
        const uint32_t tu = su + bcast0(lu), tv = sv + bcast0(lv);
        scan_pair<false>(g, true, u, v, su, tu, two, v, u, sv, tv, pu, pv, fu, fv, hadu, hadv, xu, xv);
        if (!two) scan_pair<false>(g, false, u, v, su, tu, true, v, u, sv, tv, d0, pv, d1, fv, h0, hadv, d0, xv);
        acc.scan += (pu ? pu : tu - su) + (pv ? pv : tv - sv);
        const uint32_t mu = (pu ? D_ROW : 0u) | ((hu == INVALID) == (pu != 0) ? D_BITS : 0u);
        const uint32_t mv = (pv ? D_ROW : 0u) | ((hv == INVALID) == (pv != 0) ? D_BITS : 0u);
        const bool hold = hint && two && pu && pv;   // the same edge's re-insert may follow
        if (hold) {   // both copies removed: the re-insert's slots; the marks wait
            const uint32_t ou = su + pu - 1, ov = sv + pv - 1;
            *hint = RowHint{u, v, su, eu, sv, ev, hu, hv, xu < ou ? xu : ou, xv < ov ? xv : ov, ou, ov, mu, mv,
                            cu, cv, stu, stv, true};
        } else if (lane() == 0) {
            mark(g, u, mu);
            mark(g, v, mv);
        }
        on_removed_neighbor(g, u, hu, cv);
        on_removed_neighbor(g, v, hv, cu);
        try_reduce(g, u, cv, cu, bu, hu, acc);
        try_reduce(g, v, cu, cv, bv, hv, acc);
    }
    acc.updates++;
    acc.hd += (hu != INVALID) + (hv != INVALID);
}

}  // namespace sgc
