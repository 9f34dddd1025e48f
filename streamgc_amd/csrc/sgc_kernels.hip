// sgc_kernels.hip -- HIP kernels of the StreamGC hot path for gfx950.
//
// Work decomposition: updates are one wavefront each (grid-stride), or one
// wavefront per segment replayed in order (deterministic mode).  Epoch sweeps
// run one wavefront per vertex row over the vertex space of the graphs taking
// part (their vertex counts laid end to end); rows are streamed 64
// slots per load.  The deterministic repair sweep runs one wavefront per
// graph in vertex order.
#include <hip/hip_runtime.h>

#include "sgc_device.hpp"
#include "sgc_launch.hpp"

namespace sgc {

__device__ __forceinline__ uint32_t wave_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / WAVE; }
__device__ __forceinline__ uint32_t wave_count() { return (gridDim.x * blockDim.x) / WAVE; }

// The vertex space of a launch: the vertices of the graphs taking part, one
// graph after another.  A launch over a few graphs of a large pool touches
// only their vertices.
struct Space {
    const uint32_t* base;  // ns + 1 prefix of vertex counts (null: a single graph)
    const uint32_t* gid;   // graph id of each part (null: part k is graph k)
    uint32_t ns;
    uint32_t total;
    __device__ __forceinline__ void locate(uint32_t i, uint32_t& g, uint32_t& v) const {
        uint32_t lo = 0;
        if (base && ns > 1) {
            uint32_t hi = ns;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (base[mid] <= i) lo = mid;
                else hi = mid;
            }
        }
        g = gid ? gid[lo] : lo;
        v = base ? i - base[lo] : i;
    }
};

// A vertex space written on the device (a captured round's control kernel
// lays out the graphs still taking part; see dspace_words).
__device__ __forceinline__ Space device_space(const uint32_t* dspace, uint32_t ng) {
    return Space{dspace + 4, dspace + 4 + ng + 1, dspace[0], dspace[1]};
}

// ---------------------------------------------------------------- sweep sets
// A maintenance sweep visits either every virtual vertex (list == null, one
// wavefront per row) or the worklist launch_build_list left on the device.
// The worklist keeps rows of at most `long_row` filled slots at its front
// (count[0] of them, one wavefront each) and longer rows at its back
// (count[1], list[cap - 1 - k], one 256-thread block each), so a hub row
// costs a few block-wide round trips instead of hundreds of wavefront ones.
// Lengths are read on the device: the host never waits for them.
constexpr uint32_t BLOCK = 256;

struct SweepList {
    const uint32_t* list;   // null: every vertex, wavefront per row
    const uint32_t* count;  // [small, big]
    uint32_t cap;
};
__device__ __forceinline__ uint32_t sweep_small(const SweepList& sl, uint32_t V) { return sl.list ? sl.count[0] : V; }
__device__ __forceinline__ uint32_t sweep_big(const SweepList& sl) { return sl.list ? sl.count[1] : 0; }
__device__ __forceinline__ uint32_t small_item(const SweepList& sl, uint32_t k) { return sl.list ? sl.list[k] : k; }
__device__ __forceinline__ uint32_t big_item(const SweepList& sl, uint32_t k) { return sl.list[sl.cap - 1 - k]; }

// Block-wide sum of per-wave values into one u64 total (one atomic per block).
// Every thread of the block must call it.
__device__ __forceinline__ void block_add64(uint64_t* totals, uint32_t idx, uint64_t mine) {
    __shared__ uint64_t part[BLOCK / WAVE];
    const uint32_t w = threadIdx.x / WAVE;
    if (lane() == 0) part[w] = mine;
    __syncthreads();
    if (threadIdx.x == 0 && totals) {
        uint64_t t = 0;
        for (uint32_t k = 0; k < BLOCK / WAVE; k++) t += part[k];
        if (t) tot_add(totals, idx, t);
    }
    __syncthreads();
}

// Block-wide sum (every thread gets it).  Every thread must call it.
__device__ __forceinline__ uint32_t block_sum32(uint32_t x) {
    __shared__ uint32_t part[BLOCK / WAVE];
    x = wave_sum32(x);
    if (lane() == 0) part[threadIdx.x / WAVE] = x;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t k = 0; k < BLOCK / WAVE; k++) t += part[k];
    __syncthreads();
    return t;
}

// Exclusive prefix of x over the block (thread order); *total = block sum.
__device__ __forceinline__ uint32_t block_scan32(uint32_t x, uint32_t* total) {
    __shared__ uint32_t part[BLOCK / WAVE];
    uint32_t inc = x;
    for (uint32_t d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d);
        if (lane() >= d) inc += y;
    }
    const uint32_t w = threadIdx.x / WAVE;
    if (lane() == WAVE - 1) part[w] = inc;
    __syncthreads();
    uint32_t before = 0, t = 0;
    for (uint32_t k = 0; k < BLOCK / WAVE; k++) {
        before += k < w ? part[k] : 0;
        t += part[k];
    }
    __syncthreads();
    *total = t;
    return before + inc - x;
}

// The worklist of a sweep (the vertices marked `kind`, or sticky; their marks
// are cleared), in two passes with no shared counter.  The vertices are cut
// into tiles of LIST_TILE_V, each inside one graph (with several graphs, each
// graph's tiles follow the previous graph's).  Each thread reads the marks of
// 4 vertices as one word (mark arrays are padded to a multiple of 4) with no
// per-vertex branches.  Pass 1: each block takes a tile, lists its hits in
// the tile's own region of a scratch array (short rows up from the start,
// long rows down from the end), clears their marks and stores its two
// counts.  Pass 2: each block adds up the counts of the tiles before it and
// copies its hits to their places in the worklist; the last block stores the
// totals.
constexpr uint32_t LIST_WORDS = 4;                          // mark words per thread
constexpr uint32_t LIST_TILE_V = BLOCK * LIST_WORDS * 4;    // vertices per tile
static_assert(LIST_TILE_V == 4096, "list_words() sizes the scratch for 4096-vertex tiles");

struct ListTile {
    uint32_t small[LIST_WORDS], big[LIST_WORDS];   // 4-bit vertex masks per word
};
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t x) {
    return ((x & 0xFFu) != 0) | (((x >> 8) & 0xFFu) != 0) << 1 | (((x >> 16) & 0xFFu) != 0) << 2 |
           (((x >> 24) & 0xFFu) != 0) << 3;
}
__device__ __forceinline__ void list_tile(const GraphDev& g, uint32_t kind, uint32_t long_row, uint32_t tile,
                                          uint32_t words, ListTile& out) {
    SGC_G uint32_t* mk = (SGC_G uint32_t*)(gp(g.dirty) + (size_t)kind * g.dstride);
    const SGC_G uint32_t* sk = (const SGC_G uint32_t*)(gp(g.dirty) + (size_t)M_STICKY * g.dstride);
    // Sticky marks come only with an event counted in C_ASYM (overflow,
    // duplicate): a graph that never had one has none, and its sticky array
    // is not read.
    const bool any_sticky = ld(&g.ctr[C_ASYM]) != 0;
    uint32_t m[LIST_WORDS], h[LIST_WORDS];
#pragma unroll
    for (uint32_t k = 0; k < LIST_WORDS; k++) {
        const uint32_t w = tile * (BLOCK * LIST_WORDS) + k * BLOCK + threadIdx.x;
        m[k] = w < words ? mk[w] : 0u;
        h[k] = any_sticky && w < words ? sk[w] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < LIST_WORDS; k++) {
        const uint32_t w = tile * (BLOCK * LIST_WORDS) + k * BLOCK + threadIdx.x;
        uint32_t hit = nonzero_bytes(m[k] | h[k]);
        if (4 * w + 4 > g.n) hit &= (1u << (g.n > 4 * w ? g.n - 4 * w : 0u)) - 1u;   // padding bytes
        uint32_t big = 0;
        if (hit) {   // fill is allocated to a multiple of 4 vertices
            const uint4 f = *reinterpret_cast<const uint4*>(g.fill + 4 * (size_t)w);
            big = ((f.x > long_row) | (f.y > long_row) << 1 | (f.z > long_row) << 2 | (f.w > long_row) << 3) & hit;
        }
        out.small[k] = hit & ~big;
        out.big[k] = big;
        if (m[k]) mk[w] = 0u;
    }
}

__global__ __launch_bounds__(256) void k_list_tiles(const GraphDev* __restrict__ graphs, Space sp,
                                                    const uint32_t* __restrict__ stile, uint32_t kind,
                                                    uint32_t long_row, uint32_t* __restrict__ tmp,
                                                    uint32_t* __restrict__ tcount,
                                                    const uint32_t* __restrict__ active,
                                                    const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const uint32_t t = blockIdx.x;
    // Several graphs: their tiles laid end to end (stile = prefix of each
    // part's tile count); the part holding tile t, found by the whole block.
    uint32_t part = 0, first = 0;
    if (stile && sp.ns > 1) {
        uint32_t lo = 0, hi = sp.ns;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (stile[mid] <= t) lo = mid;
            else hi = mid;
        }
        part = lo;
        first = stile[lo];
    }
    const uint32_t gid = sp.gid ? sp.gid[part] : part;
    const uint32_t off = sp.base ? sp.base[part] : 0u;   // virtual index of the graph's vertex 0
    if (active && active[gid] == 0) {
        if (stile && sp.ns > 1 && threadIdx.x == 0) {   // nothing listed from this tile
            gp(tcount)[2 * t] = 0;
            gp(tcount)[2 * t + 1] = 0;
        }
        return;
    }
    const GraphDev g = graphs[gid];
    const uint32_t words = (g.n + 3) / 4;
    const uint32_t lt_index = t - first;   // the tile within its graph
    ListTile lt;
    list_tile(g, kind, long_row, lt_index, words, lt);
    uint32_t cs = 0, cb = 0;
#pragma unroll
    for (uint32_t k = 0; k < LIST_WORDS; k++) {
        cs += __builtin_popcount(lt.small[k]);
        cb += __builtin_popcount(lt.big[k]);
    }
    uint32_t ts, tb;
    uint32_t ps = block_scan32(cs, &ts);
    uint32_t pb = block_scan32(cb, &tb);
    SGC_G uint32_t* region = gp(tmp) + (size_t)t * LIST_TILE_V;
#pragma unroll
    for (uint32_t k = 0; k < LIST_WORDS; k++) {
        const uint32_t v0 = off + 4 * (lt_index * (BLOCK * LIST_WORDS) + k * BLOCK + threadIdx.x);
        for (uint32_t q = 0; q < 4; q++) {
            if ((lt.small[k] >> q) & 1u) region[ps++] = v0 + q;
            if ((lt.big[k] >> q) & 1u) region[LIST_TILE_V - 1 - pb++] = v0 + q;
        }
    }
    if (threadIdx.x == 0) {
        gp(tcount)[2 * t] = ts;
        gp(tcount)[2 * t + 1] = tb;
    }
}

__global__ __launch_bounds__(256) void k_list_gather(const GraphDev* __restrict__ graphs, Space sp, int multi,
                                                     const uint32_t* __restrict__ tmp,
                                                     const uint32_t* __restrict__ tcount,
                                                     uint32_t* __restrict__ list, uint32_t cap,
                                                     uint32_t* __restrict__ count,
                                                     const uint32_t* __restrict__ active,
                                                     const uint32_t* __restrict__ gate) {
    // one graph not taking part: nothing listed (several graphs: the tiles of
    // those not taking part counted nothing)
    const uint32_t gid = sp.gid ? sp.gid[0] : 0u;
    if ((gate && *gate == 0) || (!multi && active && active[gid] == 0)) {
        if (blockIdx.x == 0 && threadIdx.x < 2) count[threadIdx.x] = 0;
        return;
    }
    const uint32_t t = blockIdx.x;
    uint32_t bs = 0, bb = 0;
    for (uint32_t k = threadIdx.x; k < t; k += BLOCK) {
        bs += gp(tcount)[2 * k];
        bb += gp(tcount)[2 * k + 1];
    }
    bs = block_sum32(bs);
    bb = block_sum32(bb);
    const uint32_t ns = gp(tcount)[2 * t], nb = gp(tcount)[2 * t + 1];
    const SGC_G uint32_t* region = gp(tmp) + (size_t)t * LIST_TILE_V;
    for (uint32_t k = threadIdx.x; k < ns; k += BLOCK) gp(list)[bs + k] = region[k];
    for (uint32_t k = threadIdx.x; k < nb; k += BLOCK) gp(list)[cap - 1 - (bb + k)] = region[LIST_TILE_V - 1 - k];
    if (t == gridDim.x - 1 && threadIdx.x == 0) {
        count[0] = bs + ns;
        count[1] = bb + nb;
    }
}

__global__ void k_zero_list_count(uint32_t* p) {
    p[0] = 0;
    p[1] = 0;
}

// ---------------------------------------------------------------- per-edge ordering
// Parallel waves apply a batch in no particular order, so two updates of one
// edge in one batch (a delete and its re-insert, frequent in churn streams)
// could land in either order and leave the rows disagreeing with the stream.
// Before a parallel batch, the updates are grouped by edge (global vertex ids
// of the pool, lo << 32 | hi) in a hash table: each slot heads a linked list
// of the batch positions of that edge.  The wave that owns the list head
// applies the whole group in stream order; different edges stay parallel.
constexpr uint64_t KEY_EMPTY = ~0ull;
constexpr uint32_t NONE = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t edge_key(const GraphDev* __restrict__ graphs, const uint32_t* __restrict__ gbase,
                                             const uint4& r, int multi) {
    const uint32_t gid = multi ? r.w : 0u;
    const uint32_t n = graphs[gid].n;
    if (r.x >= n || r.y >= n || (r.z & 0xFFu) > 1u) return KEY_EMPTY;
    const uint32_t base = multi ? gbase[gid] : 0u;
    const uint32_t a = base + (r.x < r.y ? r.x : r.y), b = base + (r.x < r.y ? r.y : r.x);
    return ((uint64_t)a << 32) | b;
}
__device__ __forceinline__ uint32_t key_hash(uint64_t k, uint32_t mask) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}

__global__ __launch_bounds__(256) void k_group_clear(uint64_t* __restrict__ keys, uint32_t* __restrict__ head,
                                                     uint32_t size, uint32_t* __restrict__ nheads) {
    if (blockIdx.x == 0 && threadIdx.x < 2) nheads[threadIdx.x] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
        gp(keys)[i] = KEY_EMPTY;
        gp(head)[i] = NONE;
    }
}

// Slot of `key` (inserted when absent; *created tells whether this call
// inserted it).  The table holds at least twice the batch, so a free slot
// always exists and the probe ends.
__device__ __forceinline__ uint32_t group_slot(uint64_t* __restrict__ keys, uint64_t key, uint32_t mask,
                                               bool* created) {
    uint32_t s = key_hash(key, mask);
    for (uint32_t probes = 0; probes <= mask; s = (s + 1) & mask, probes++) {
        uint64_t found = KEY_EMPTY;
        __hip_atomic_compare_exchange_strong(gp(&keys[s]), &found, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        *created = found == KEY_EMPTY;
        if (found == KEY_EMPTY || found == key) return s;
    }
    return NONE;
}

// Links every update of the batch into its edge's list, and lists each edge
// once (the update that inserted the key adds its slot; an update whose key
// found no slot is listed alone, position | GROUP_SINGLE): the per-update
// kernel then runs one wavefront per edge.  Records that change nothing
// (out-of-range vertex, other kinds) are not listed.  Edges with an endpoint
// row filled past HEAVY_FILL slots (the long scans) go to the front of the
// list, the others to the back, so that the long ones start first and the
// short ones fill in behind them.  Each block appends its entries with one
// atomic per list end; nheads = [front, back] counts.
constexpr uint32_t GROUP_SINGLE = 0x80000000u;
constexpr uint32_t HEAVY_FILL = 1024;

__global__ __launch_bounds__(256) void k_group_link(const GraphDev* __restrict__ graphs,
                                                    const uint32_t* __restrict__ gbase,
                                                    const uint4* __restrict__ upd, uint32_t count, int multi,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ head,
                                                    uint32_t* __restrict__ next, uint32_t mask,
                                                    uint32_t* __restrict__ heads, uint32_t cap,
                                                    uint32_t* __restrict__ nheads, const uint32_t* __restrict__ dyn) {
    __shared__ uint32_t block_at[2];
    if (dyn) {   // the batch is a slice of a device-resident stream: [offset, count] read here
        upd += dyn[0];
        count = dyn[1];
    }
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < count; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        uint32_t entry = NONE;
        bool heavy = false;
        if (i < count) {
            const uint4 r = upd[i];
            const uint64_t key = edge_key(graphs, gbase, r, multi);
            uint32_t prev = NONE;
            if (key != KEY_EMPTY) {
                bool created = false;
                const uint32_t s = group_slot(keys, key, mask, &created);
                if (s != NONE) {
                    prev = __hip_atomic_exchange(gp(&head[s]), i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (created) entry = s;
                } else {
                    entry = i | GROUP_SINGLE;
                }
                if (entry != NONE) {
                    const GraphDev& g = graphs[multi ? r.w : 0u];
                    heavy = gp(g.fill)[r.x] > HEAVY_FILL || gp(g.fill)[r.y] > HEAVY_FILL;
                }
            }
            gp(next)[i] = prev;
        }
        uint32_t nf, nb;
        const uint32_t af = block_scan32(entry != NONE && heavy ? 1u : 0u, &nf);
        const uint32_t ab = block_scan32(entry != NONE && !heavy ? 1u : 0u, &nb);
        if (threadIdx.x == 0) {
            block_at[0] = nf ? add32(&nheads[0], nf) : 0u;
            block_at[1] = nb ? add32(&nheads[1], nb) : 0u;
        }
        __syncthreads();
        if (entry != NONE) gp(heads)[heavy ? block_at[0] + af : cap - 1 - (block_at[1] + ab)] = entry;
        __syncthreads();
    }
}

// One thread per listed edge: the records of its updates, in stream order,
// into the edge's own GROUP_PACK x 16-byte line of `pack`, so that the
// per-update kernel has a whole group after one load instead of walking
// head, list links and records one dependent load after another.  Unused
// entries are PACK_NONE records; an edge with more than GROUP_PACK updates is
// left to the walk (its first entry says so and carries the list entry).
constexpr uint32_t GROUP_PACK = 4;
constexpr uint32_t PACK_NONE = 0xFFFFFFFFu;   // x of an entry with no record (a listed record has x < n)
constexpr uint32_t PACK_END = 0, PACK_WALK = 1;   // y of such an entry

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

__global__ __launch_bounds__(256) void k_group_pack(const uint4* __restrict__ upd, const uint32_t* __restrict__ ghead,
                                                    const uint32_t* __restrict__ gnext,
                                                    const uint32_t* __restrict__ heads, uint32_t cap,
                                                    const uint32_t* __restrict__ nheads, uint4* __restrict__ pack,
                                                    const uint32_t* __restrict__ dyn) {
    if (dyn) upd += dyn[0];
    const uint32_t nf = nheads[0], n = nf + nheads[1];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t e = j < nf ? heads[j] : heads[cap - 1 - (j - nf)];
        uint32_t p0 = NONE, p1 = NONE, p2 = NONE, p3 = NONE;
        bool more = false;
        if (e & GROUP_SINGLE) {
            p0 = e & ~GROUP_SINGLE;
        } else {
            p0 = gp(ghead)[e];
            if (p0 != NONE) p1 = gp(gnext)[p0];
            if (p1 != NONE) p2 = gp(gnext)[p1];
            if (p2 != NONE) p3 = gp(gnext)[p2];
            more = p3 != NONE && gp(gnext)[p3] != NONE;
        }
        uint4* out = pack + (size_t)j * GROUP_PACK;
        if (more) {
            out[0] = make_uint4(PACK_NONE, PACK_WALK, 0u, e);
            continue;
        }
        cswap(p0, p1);   // sorting network for 4 (NONE sorts last)
        cswap(p2, p3);
        cswap(p0, p2);
        cswap(p1, p3);
        cswap(p1, p2);
        // (each record loaded under its own branch: selects between whole
        // vector records put them in scratch memory)
        const uint4 none = make_uint4(PACK_NONE, PACK_END, 0u, 0u);
        uint4 r0 = none, r1 = none, r2 = none, r3 = none;
        if (p0 != NONE) r0 = upd[p0];
        if (p1 != NONE) r1 = upd[p1];
        if (p2 != NONE) r2 = upd[p2];
        if (p3 != NONE) r3 = upd[p3];
        out[0] = r0;
        out[1] = r1;
        out[2] = r2;
        out[3] = r3;
    }
}

// ---------------------------------------------------------------- updates
__device__ __forceinline__ void apply_rec(const GraphDev* __restrict__ graphs, const uint4& r, int multi, WaveAcc& acc,
                                          RowHint* hint = nullptr) {
    const uint32_t gid = multi ? r.w : 0u;
    const GraphDev g = graphs[gid];
    apply_update(g, r.x, r.y, r.z & 0xFFu, acc, hint);
}
__device__ __forceinline__ void apply_record(const GraphDev* __restrict__ graphs, const uint4* __restrict__ upd,
                                             uint32_t i, int multi, WaveAcc& acc, RowHint* hint = nullptr) {
    apply_rec(graphs, upd[i], multi, acc, hint);
}

__global__ __launch_bounds__(64) void k_apply_elem(const GraphDev* __restrict__ graphs,
                                                    const uint4* __restrict__ upd, uint32_t count, int multi,
                                                    const uint32_t* __restrict__ ghead,
                                                    const uint32_t* __restrict__ gnext,
                                                    const uint32_t* __restrict__ heads, uint32_t cap,
                                                    const uint32_t* __restrict__ nheads,
                                                    const uint4* __restrict__ pack, uint64_t* totals,
                                                    const uint32_t* __restrict__ dyn) {
    WaveAcc acc;
    uint64_t grouped = 0;
    if (dyn) {
        upd += dyn[0];
        count = dyn[1];
    }
    if (!heads) {   // no grouping: one wavefront per position
        for (uint32_t i = wave_id(); i < count; i += wave_count()) apply_record(graphs, upd, i, multi, acc);
        acc.flush(totals);
        return;
    }
    // One wavefront per edge of the batch: its updates in stream order.  The
    // records come from the edge's packed 64-byte line (lanes 0..3 load one
    // each), or, for a group longer than the line, from its list (head, then
    // next links): repeatedly the record at the smallest position above the
    // last one applied.
    const uint32_t nf = nheads[0], n = nf + nheads[1];
    for (uint32_t j = wave_id(); j < n; j += wave_count()) {
        uint4 mine = make_uint4(PACK_NONE, PACK_WALK, 0u, 0u);
        if (pack && lane() < GROUP_PACK) mine = pack[(size_t)j * GROUP_PACK + lane()];
        const bool walk = __builtin_amdgcn_readlane(mine.x, 0) == PACK_NONE &&
                          __builtin_amdgcn_readlane(mine.y, 0) == PACK_WALK;
        uint32_t h = NONE;
        bool single = false;
        if (walk) {
            const uint32_t e = j < nf ? heads[j] : heads[cap - 1 - (j - nf)];   // long scans first
            single = (e & GROUP_SINGLE) != 0;
            h = single ? e & ~GROUP_SINGLE : gp(ghead)[e];
        }
        RowHint hint;   // a delete's slots for the same edge's re-insert (a last delete's marks wait in it)
        hint.valid = false;
        uint32_t applied = 0, last = 0, last_g = 0;
        for (;;) {
            uint4 r;
            if (!walk) {
                if (applied == GROUP_PACK) break;
                r = make_uint4(__builtin_amdgcn_readlane(mine.x, applied), __builtin_amdgcn_readlane(mine.y, applied),
                               __builtin_amdgcn_readlane(mine.z, applied), __builtin_amdgcn_readlane(mine.w, applied));
                if (r.x == PACK_NONE) break;
            } else if (single) {
                if (applied) break;
                r = upd[h];
            } else {
                uint32_t best = NONE;
                for (uint32_t k2 = h; k2 != NONE; k2 = gp(gnext)[k2])
                    if ((applied == 0 || k2 > last) && (best == NONE || k2 < best)) best = k2;
                if (best == NONE) break;
                r = upd[best];
                last = best;
            }
            apply_rec(graphs, r, multi, acc, &hint);
            last_g = r.w;
            applied++;
        }
        if (hint.valid) hint_release(graphs[multi ? last_g : 0u], &hint);   // marks of a last delete
        grouped += applied ? applied - 1 : 0;
    }
    acc.flush(totals);
    if (grouped && totals && lane() == 0) tot_add(totals, T_GROUPED, grouped);
}

// One wavefront per segment, the segment's updates in order; positions go
// through `perm` when given (per-graph order inside a mixed batch).
__global__ __launch_bounds__(256) void k_apply_seg(const GraphDev* __restrict__ graphs,
                                                   const uint4* __restrict__ upd, const uint32_t* __restrict__ seg,
                                                   uint32_t nseg, const uint32_t* __restrict__ perm, int multi,
                                                   uint64_t* totals, const uint32_t* __restrict__ dyn) {
    WaveAcc acc;
    if (dyn) {   // segments of the current round (absolute stream positions); none when its count is 0
        if (dyn[1] == 0) return;
        seg += dyn[2];
    }
    for (uint32_t k = wave_id(); k < nseg; k += wave_count())
        for (uint32_t i = seg[2 * k]; i < seg[2 * k + 1]; i++)
            apply_record(graphs, upd, perm ? perm[i] : i, multi, acc);
    acc.flush(totals);
}

// ---------------------------------------------------------------- per-graph control
__global__ void k_zero_ctr(const GraphDev* __restrict__ graphs, uint32_t ng, const uint32_t* __restrict__ active,
                           uint32_t idx) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < ng && (!active || active[gid])) gp(graphs[gid].ctr)[idx] = 0;
}

// Maximum colour per graph into ctr[C_MAXC] (A/src/streamgc.cu:549-553).
__global__ __launch_bounds__(256) void k_max_color(const GraphDev* __restrict__ graphs, Space sp) {
    uint32_t cur = NONE, mx = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sp.total; i += gridDim.x * blockDim.x) {
        uint32_t gid, v;
        sp.locate(i, gid, v);
        if (gid != cur) {
            if (cur != NONE && mx) max32(&graphs[cur].ctr[C_MAXC], mx);
            cur = gid;
            mx = 0;
        }
        const uint32_t c = color_of(gp(graphs[gid].word)[v]);
        mx = c > mx ? c : mx;
    }
    const uint32_t c0 = bcast0(cur);
    if (ballot(cur != c0) == 0) {              // the whole wave on one graph: one atomic
        mx = wave_max32(mx);
        if (lane() == 0 && cur != NONE && mx) max32(&graphs[cur].ctr[C_MAXC], mx);
    } else if (cur != NONE && mx) {
        max32(&graphs[cur].ctr[C_MAXC], mx);
    }
}

// One graph: 4 words per 16-byte read, a block-wide maximum, one atomic per
// block (thousands of waves folding into one counter serialise on it).
__global__ __launch_bounds__(256) void k_max_color1(const GraphDev* __restrict__ graphs, Space sp) {
    __shared__ uint32_t part[BLOCK / WAVE];
    const uint32_t gid = sp.gid ? sp.gid[0] : 0u;
    const GraphDev g = graphs[gid];
    const uint32_t quads = g.n / 4;
    uint32_t mx = 0;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
        const uint4 w = *reinterpret_cast<const uint4*>(g.word + 4 * (size_t)q);
        const uint32_t a = color_of(w.x) > color_of(w.y) ? color_of(w.x) : color_of(w.y);
        const uint32_t b = color_of(w.z) > color_of(w.w) ? color_of(w.z) : color_of(w.w);
        mx = a > mx ? a : mx;
        mx = b > mx ? b : mx;
    }
    if (blockIdx.x == 0 && threadIdx.x < g.n % 4) {
        const uint32_t c = color_of(gp(g.word)[4 * quads + threadIdx.x]);
        mx = c > mx ? c : mx;
    }
    mx = wave_max32(mx);
    if (lane() == 0) part[threadIdx.x / WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < BLOCK / WAVE; k++) t = part[k] > t ? part[k] : t;
        if (t) max32(&g.ctr[C_MAXC], t);
    }
}

// ---------------------------------------------------------------- snapshot
// A/src/kernels/epoch_maintenance.cu:5-16, into the back buffer.  In a
// captured round (fl_round and hist set) the same colours also go to the
// graph's next history slot, if it keeps a history and has one free, so the
// snapshot is not read back for it.
__device__ __forceinline__ SGC_G uint16_t* history_slot(const uint32_t* fl_round, const HistDev* hist, uint32_t gid,
                                                       uint32_t n) {
    if (!hist || !fl_round[gid]) return nullptr;
    const HistDev h = hist[gid];
    return h.base && h.count < h.cap ? gp(h.base + (size_t)h.count * n) : nullptr;
}

// One graph: each thread copies the colours of 4 consecutive vertices (one
// 16-byte read of their words; the snapshot half may start at any even
// address, so four 2-byte stores).
__global__ __launch_bounds__(256) void k_snapshot1(const GraphDev* __restrict__ graphs, Space sp,
                                                   const uint32_t* __restrict__ active,
                                                   const uint32_t* __restrict__ fl_round,
                                                   const HistDev* __restrict__ hist) {
    const uint32_t gid = sp.gid ? sp.gid[0] : 0u;
    if (active && !active[gid]) return;
    const GraphDev g = graphs[gid];
    SGC_G uint16_t* out = gp(g.snap_write);
    SGC_G uint16_t* keep = history_slot(fl_round, hist, gid, g.n);
    const uint32_t quads = g.n / 4;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
        const uint4 w = *reinterpret_cast<const uint4*>(g.word + 4 * (size_t)q);
        const uint16_t c0 = color_of(w.x), c1 = color_of(w.y), c2 = color_of(w.z), c3 = color_of(w.w);
        out[4 * q] = c0;
        out[4 * q + 1] = c1;
        out[4 * q + 2] = c2;
        out[4 * q + 3] = c3;
        if (keep) {
            keep[4 * q] = c0;
            keep[4 * q + 1] = c1;
            keep[4 * q + 2] = c2;
            keep[4 * q + 3] = c3;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < g.n % 4) {
        const uint32_t v = 4 * quads + threadIdx.x;
        const uint16_t c = color_of(gp(g.word)[v]);
        out[v] = c;
        if (keep) keep[v] = c;
    }
}

__global__ __launch_bounds__(256) void k_snapshot(const GraphDev* __restrict__ graphs, Space sp,
                                                  const uint32_t* __restrict__ active,
                                                  const uint32_t* __restrict__ fl_round,
                                                  const HistDev* __restrict__ hist) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sp.total; i += gridDim.x * blockDim.x) {
        uint32_t gid, v;
        sp.locate(i, gid, v);
        if (active && !active[gid]) continue;
        const GraphDev& g = graphs[gid];
        const uint16_t c = color_of(gp(g.word)[v]);
        gp(g.snap_write)[v] = c;
        SGC_G uint16_t* keep = history_slot(fl_round, hist, gid, g.n);
        if (keep) keep[v] = c;
    }
}

// ---------------------------------------------------------------- compaction
// A/src/kernels/epoch_maintenance.cu:20-61: stable left-pack of the live
// slots, free slots after them.  One pass, writing only slots whose content
// changes: a live value moves to w + (live lanes before it); a slot whose
// value moved and that lies past the chunk's new write front is freed (a
// later chunk may refill it, in program order).
// One long row, all threads of the block: windows of 1024 slots (4
// consecutive slots per thread), a block prefix of the live counts, then the
// moves.  Writes of a window land below its end, reads of the next window
// start at its end.  Returns the live count; the row's fill becomes it.
__device__ __forceinline__ uint32_t compact_row_block(const GraphDev& g, uint32_t v) {
    const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
    uint32_t w = s;
    for (uint32_t base = s; base < e; base += 4 * BLOCK) {
        const uint32_t j0 = base + 4 * threadIdx.x;
        uint32_t val[4], cnt = 0;
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            val[q] = j0 + q < e ? gp(g.col)[j0 + q] : INVALID;
            cnt += val[q] != INVALID;
        }
        uint32_t total = 0;
        uint32_t pos = w + block_scan32(cnt, &total);
        const uint32_t w_after = w + total;
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) {
            const uint32_t j = j0 + q;
            if (val[q] == INVALID) continue;
            if (pos != j) {
                gp(g.col)[pos] = val[q];
                if (j >= w_after) gp(g.col)[j] = INVALID;
            }
            pos++;
        }
        w = w_after;
        __syncthreads();
    }
    if (threadIdx.x == 0) gp(g.fill)[v] = w - s;
    return w - s;
}

// One row, one wavefront: ROW_UNROLL chunks loaded, then packed in order.
__device__ __forceinline__ uint32_t compact_row_wave(const GraphDev& g, uint32_t v) {
    const uint64_t lt_mask = (1ull << lane()) - 1ull;
    const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
    uint32_t w = s;
    for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
        uint32_t val[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            val[u] = j < e ? gp(g.col)[j] : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            const bool live = val[u] != INVALID;
            const uint64_t m = ballot(live);
            const uint32_t pos = w + (uint32_t)__builtin_popcountll(m & lt_mask);
            const uint32_t w_after = w + (uint32_t)__builtin_popcountll(m);
            if (live && pos != j) gp(g.col)[pos] = val[u];
            if (live && pos != j && j >= w_after) gp(g.col)[j] = INVALID;
            w = w_after;
        }
    }
    if (lane() == 0) gp(g.fill)[v] = w - s;
    return w - s;
}

// The drain's counter snapshot of one graph (A/src/streamgc.cu:972-983): the
// overflow entries it retries (capped at the buffer), those lost past the
// buffer, counters reset.
__device__ __forceinline__ void drain_prep_graph(const GraphDev* __restrict__ graphs, uint32_t gid) {
    SGC_G uint32_t* c = gp(graphs[gid].ctr);
    uint32_t raw = c[C_OVF];
    c[C_OVF_RAW] = raw;
    c[C_OVF_SNAP] = raw < graphs[gid].ovf_cap ? raw : graphs[gid].ovf_cap;
    if (raw > graphs[gid].ovf_cap) c[C_LOST] += raw - graphs[gid].ovf_cap;   // never buffered
    c[C_OVF] = 0;
    c[C_DRAIN_FAILED] = 0;
}

// prep_ng > 0: block 0 also takes the drain's counter snapshot of the
// active graphs (k_drain_prep), which the drain launched next reads;
// compaction touches none of those counters.
__global__ __launch_bounds__(256) void k_compact(const GraphDev* __restrict__ graphs, Space sp, SweepList sl,
                                                 uint64_t* totals, uint32_t prep_ng,
                                                 const uint32_t* __restrict__ prep_active) {
    if (blockIdx.x == 0)
        for (uint32_t gid = threadIdx.x; gid < prep_ng; gid += blockDim.x)
            if (!prep_active || prep_active[gid]) drain_prep_graph(graphs, gid);
    uint64_t live_total = 0, rows = 0, slots = 0;
    const uint32_t nbig = sweep_big(sl);
    for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
        uint32_t gid, v;
        sp.locate(big_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t f = gp(g.fill)[v];
        const uint32_t live = compact_row_block(g, v);
        if (threadIdx.x == 0) {
            live_total += live;
            rows++;
            slots += f;
        }
    }
    const uint32_t total = sweep_small(sl, sp.total);
    for (uint32_t k = wave_id(); k < total; k += wave_count()) {
        uint32_t gid, v;
        sp.locate(small_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t f = gp(g.fill)[v];
        live_total += compact_row_wave(g, v);
        rows++;
        slots += f;
    }
    if (lane() != 0) live_total = rows = slots = 0;
    block_add64(totals, T_LIVE_SLOTS, live_total);
    block_add64(totals, T_SWEEP_ROWS + 0, rows);
    block_add64(totals, T_SWEEP_SLOTS + 0, slots);
}

// ---------------------------------------------------------------- overflow drain
// A/src/streamgc.cu:972-983: read the counter, reset it, clamp to capacity.
__global__ void k_drain_prep(const GraphDev* __restrict__ graphs, uint32_t ng, const uint32_t* __restrict__ active) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ng || (active && !active[gid])) return;
    drain_prep_graph(graphs, gid);
}

// A/src/kernels/epoch_maintenance.cu:66-87: one wavefront per entry, graph = blockIdx.y.
__global__ __launch_bounds__(256) void k_drain_elem(const GraphDev* __restrict__ graphs,
                                                    const uint32_t* __restrict__ active) {
    const uint32_t gid = blockIdx.y;
    if (active && !active[gid]) return;
    const GraphDev g = graphs[gid];
    const uint32_t count = gp(g.ctr)[C_OVF_SNAP];
    uint32_t failed = 0;
    for (uint32_t k = wave_id(); k < count; k += wave_count()) {
        bool f = false;
        const uint32_t x = gp(g.ovf)[2 * (size_t)k];
        row_insert(g, x, gp(g.ovf)[2 * (size_t)k + 1], false, &f, nullptr);
        if (lane() == 0 && x < g.n) mark(g, x, D_BITS | D_STATUS);
        failed += f;
    }
    if (failed && lane() == 0) {
        add32(&g.ctr[C_DRAIN_FAILED], failed);
        add32(&g.ctr[C_LOST], failed);
    }
}

// The same in buffer order, one wavefront per graph (deterministic mode).
__global__ __launch_bounds__(64) void k_drain_serial(const GraphDev* __restrict__ graphs, uint32_t ng,
                                                     const uint32_t* __restrict__ active) {
    for (uint32_t gid = blockIdx.x; gid < ng; gid += gridDim.x) {
        if (active && !active[gid]) continue;
        const GraphDev g = graphs[gid];
        const uint32_t count = gp(g.ctr)[C_OVF_SNAP];
        uint32_t failed = 0;
        for (uint32_t k = 0; k < count; k++) {
            bool f = false;
            const uint32_t x = gp(g.ovf)[2 * (size_t)k];
            row_insert(g, x, gp(g.ovf)[2 * (size_t)k + 1], false, &f, nullptr);
            if (lane() == 0 && x < g.n) mark(g, x, D_BITS | D_STATUS);
            failed += f;
        }
        if (lane() == 0) {
            gp(g.ctr)[C_DRAIN_FAILED] = failed;
            gp(g.ctr)[C_LOST] += failed;
        }
    }
}

// ---------------------------------------------------------------- bitmap / histogram rebuild
// A/include/color_select.cuh:213-251, A/src/kernels/epoch_maintenance.cu:113-128:
// bitmap-mode rows get the exact OR of their live neighbours' colours <= 64.
// Exact-histogram mode also recounts every hub's 1024 counters from its live
// row (colours 1..1024): one wavefront per hub, the bins in that wavefront's
// own 4 KiB of LDS, written out as 16 coalesced dwords per lane.  The
// reference never rebuilds histograms (A/src/kernels/epoch_maintenance.cu:120-123),
// so its counts drift; this mode keeps them exact at every rebuild.
// A long hub row: the whole block counts into one shared 1024-bin tile.
__device__ __forceinline__ void rebuild_hub_block(const GraphDev& g, uint32_t v, uint32_t hd, uint32_t* bins) {
    for (uint32_t t = threadIdx.x; t < HIST_BINS; t += BLOCK) bins[t] = 0;
    __syncthreads();
    const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
    for (uint32_t base = s; base < e; base += ROW_UNROLL * BLOCK) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * BLOCK + threadIdx.x;
            nb[u] = j < e ? gp(g.col)[j] : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(gp(g.word)[nb[u]]) : 0;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            if (c[u] != 0 && c[u] <= HIST_BINS) atomicAdd(&bins[c[u] - 1], 1u);
    }
    __syncthreads();
    SGC_G uint32_t* dst = gp(g.hist) + (size_t)hd * HIST_BINS;
    for (uint32_t t = threadIdx.x; t < HIST_BINS; t += BLOCK) dst[t] = bins[t];
    __syncthreads();
}

// A bitmap row: the OR of its live neighbours' colours <= 64 (one wavefront,
// or the whole block for a long row).
template <bool kBlock>
__device__ __forceinline__ void rebuild_bitmap(const GraphDev& g, uint32_t v) {
    const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
    const uint32_t width = kBlock ? BLOCK : WAVE, me = kBlock ? threadIdx.x : lane();
    uint64_t bm = 0;
    for (uint32_t base = s; base < e; base += ROW_UNROLL * width) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * width + me;
            nb[u] = j < e ? gp(g.col)[j] : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(gp(g.word)[nb[u]]) : 0;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            if (c[u] != 0 && c[u] <= 64) bm |= 1ull << (c[u] - 1);
    }
    bm = wave_or64(bm);
    if (kBlock) {
        __shared__ uint64_t part[BLOCK / WAVE];
        if (lane() == 0) part[threadIdx.x / WAVE] = bm;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0;
            for (uint32_t k = 0; k < BLOCK / WAVE; k++) t |= part[k];
            gp(g.bitmap)[v] = t;
        }
        __syncthreads();
    } else if (lane() == 0) {
        gp(g.bitmap)[v] = bm;
    }
}

// A/include/color_select.cuh:213-251, A/src/kernels/epoch_maintenance.cu:113-128:
// bitmap-mode rows get the exact OR of their live neighbours' colours <= 64.
// Exact-histogram mode also recounts every listed hub's 1024 counters from
// its live row (colours 1..1024) in LDS, then writes them out coalesced.  The
// reference never rebuilds histograms (A/src/kernels/epoch_maintenance.cu:120-123),
// so its counts drift; this mode keeps them exact at every rebuild.
__global__ __launch_bounds__(256) void k_rebuild(const GraphDev* __restrict__ graphs, Space sp, SweepList sl,
                                                 int exact_hist, uint64_t* totals) {
    __shared__ uint32_t bins[BLOCK / WAVE][HIST_BINS];
    uint64_t rows = 0, slots = 0;
    const uint32_t nbig = sweep_big(sl);
    for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
        uint32_t gid, v;
        sp.locate(big_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t hd = gp(g.hd_map)[v];
        if (hd != INVALID && !exact_hist) continue;
        if (threadIdx.x == 0) {
            rows++;
            slots += gp(g.fill)[v];
        }
        if (hd == INVALID) rebuild_bitmap<true>(g, v);
        else rebuild_hub_block(g, v, hd, &bins[0][0]);
    }
    uint32_t* b = bins[threadIdx.x / WAVE];
    const uint32_t total = sweep_small(sl, sp.total);
    for (uint32_t k = wave_id(); k < total; k += wave_count()) {
        uint32_t gid, v;
        sp.locate(small_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t hd = gp(g.hd_map)[v];
        if (hd != INVALID && !exact_hist) continue;
        const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
        rows++;
        slots += e - s;
        if (hd == INVALID) {
            rebuild_bitmap<false>(g, v);
            continue;
        }
        for (uint32_t t = lane(); t < HIST_BINS; t += WAVE) b[t] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t base = s; base < e; base += ROW_UNROLL * WAVE) {
            uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) {
                const uint32_t j = base + u * WAVE + lane();
                nb[u] = j < e ? gp(g.col)[j] : INVALID;
            }
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++)
                c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(gp(g.word)[nb[u]]) : 0;
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++)
                if (c[u] != 0 && c[u] <= HIST_BINS) atomicAdd(&b[c[u] - 1], 1u);
        }
        __builtin_amdgcn_wave_barrier();
        SGC_G uint32_t* dst = gp(g.hist) + (size_t)hd * HIST_BINS;
        for (uint32_t t = lane(); t < HIST_BINS; t += WAVE) dst[t] = b[t];
        __builtin_amdgcn_wave_barrier();
    }
    if (lane() != 0) rows = slots = 0;
    block_add64(totals, T_SWEEP_ROWS + 1, rows);
    block_add64(totals, T_SWEEP_SLOTS + 1, slots);
}

// ---------------------------------------------------------------- tier-1 repair
// Colours 1..255 present among a row's live neighbours, bit (c & 63) of
// m[c >> 6].
struct ColourSet {
    uint64_t m[4] = {0, 0, 0, 0};
    __device__ __forceinline__ void add(uint32_t c) {
        if (c == 0 || c >= 256) return;
        const uint64_t b = 1ull << (c & 63);
        const uint32_t q = c >> 6;
        m[0] |= q == 0 ? b : 0;
        m[1] |= q == 1 ? b : 0;
        m[2] |= q == 2 ? b : 0;
        m[3] |= q == 3 ? b : 0;
    }
    __device__ __forceinline__ void wave_reduce() {
        for (int k = 0; k < 4; k++) m[k] = wave_or64(m[k]);
    }
    // find_available_color_bitmap on an exact bitmap: the lowest colour in
    // 1..255 not present, else 256 (A/include/color_select.cuh:20-52).
    __device__ __forceinline__ uint16_t lowest_free() const {
        const uint64_t m0 = m[0] | 1ull;   // colour 0 is not a colour
        if (~m0) return (uint16_t)__builtin_ctzll(~m0);
        for (int k = 1; k < 4; k++)
            if (~m[k]) return (uint16_t)(64 * k + __builtin_ctzll(~m[k]));
        return 256;
    }
};

// Colours 1..1024 present among a hub's live neighbours: 32 words of LDS
// per wavefront (or per block).  The lowest absent colour is what
// find_available_color_histogram picks from an exact histogram
// (A/include/color_select.cuh:59-89); with all 1024 present it falls back to
// the histogram's minimum.
constexpr uint32_t PRESENCE_WORDS = HIST_BINS / 32;
__device__ __forceinline__ void presence_add(uint32_t* p, uint32_t c) {
    if (c != 0 && c <= HIST_BINS) atomicOr(&p[(c - 1) >> 5], 1u << ((c - 1) & 31));
}
// One wavefront reads the 1024 bits: the lowest clear bit, or 0 if none.
__device__ __forceinline__ uint16_t presence_lowest_free(const uint32_t* p) {
    const uint32_t word = lane() < PRESENCE_WORDS ? p[lane()] : 0xFFFFFFFFu;
    const uint64_t open = ballot(word != 0xFFFFFFFFu);
    if (!open) return 0;
    const uint32_t k = (uint32_t)__builtin_ctzll(open);
    const uint32_t wk = (uint32_t)__shfl((int)word, (int)k);
    return (uint16_t)(32 * k + __builtin_ctz(~wk) + 1);
}

__device__ __forceinline__ uint16_t fresh_pick(const GraphDev& g, uint32_t v, bool hub, ColourSet& cs,
                                               const uint32_t* pres) {
    if (!hub) {
        cs.wave_reduce();
        return cs.lowest_free();
    }
    __builtin_amdgcn_wave_barrier();
    const uint16_t c = presence_lowest_free(pres);
    return c ? c : fas_hist(g, gp(g.hd_map)[v]);
}

// A/src/kernels/conflict_repair.cu:10-114 for one vertex; returns 1 if counted.
// Fresh selection (parallel mode): a loser selects from the exact colours of its
// neighbours, gathered by the same row scan that found the conflict -- the
// bitmap or histogram a rebuild at this instant would leave -- instead of
// the ones rebuilt before the sweep, which miss the recolours made earlier in
// the same sweep (the reference does not propagate them,
// A/src/kernels/conflict_repair.cu).  Deterministic mode keeps the
// reference's selection.  pres: this wavefront's PRESENCE_WORDS of LDS.
// fresh: 0 = the reference's selection, 1 = exact colours for bitmap-mode
// rows, 2 = for histogram-mode rows too.
// Live slots of v's row (the ablation's priority degree,
// A/src/kernels/conflict_repair.cu:46-58): one wavefront, or the whole block.
template <bool kBlock>
__device__ __forceinline__ uint32_t live_row_count(const GraphDev& g, uint32_t v) {
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    const uint32_t width = kBlock ? BLOCK : WAVE, me = kBlock ? threadIdx.x : lane();
    uint32_t k = 0;
    for (uint32_t j = s + me; j < e; j += width) k += ld(&g.col[j]) != INVALID;
    return kBlock ? block_sum32(k) : wave_sum32(k);
}

__device__ __forceinline__ uint32_t tier1_vertex(const GraphDev& g, uint32_t v, uint32_t* pres, int fresh_level) {
    const uint32_t w = bcast0(ld(&g.word[v]));
    const uint16_t mine = color_of(w);
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    const bool hub = select_by_hist(g, v);
    const bool fresh = fresh_level >= (hub ? 2 : 1);
    ColourSet cs;
    if (fresh && hub) {
        if (lane() < PRESENCE_WORDS) pres[lane()] = 0;
        __builtin_amdgcn_wave_barrier();
    }
    if (mine == 0) {
        bool has_nb = false;
        for (uint32_t base = s; base < e && (fresh || !has_nb); base += ROW_UNROLL * WAVE) {
            uint32_t nb[ROW_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) {
                const uint32_t j = base + u * WAVE + lane();
                nb[u] = j < e ? ld(&g.col[j]) : INVALID;
            }
            bool any = false;
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) {
                any |= nb[u] != INVALID;
                if (fresh && nb[u] != INVALID && nb[u] < g.n) {
                    const uint32_t cn = color_of(ld(&g.word[nb[u]]));
                    if (hub) presence_add(pres, cn);
                    else cs.add(cn);
                }
            }
            has_nb |= ballot(any) != 0;
        }
        if (!has_nb) return 0;
        const uint16_t c = fresh ? fresh_pick(g, v, hub, cs, pres) : find_color(g, v);
        if (lane() == 0) {
            st(&g.word[v], bump(w, c));
            mark(g, v, D_BITS | D_STATUS);
        }
        mark_neighbors(g, v, 0, c);
        return 1;
    }
    const uint32_t my_deg = (g.flags & ABL_LIVE_DEGREE) ? live_row_count<false>(g, v) : gp(g.deg0)[v];
    bool need = false;
    for (uint32_t base = s; base < e && (fresh || !need); base += ROW_UNROLL * WAVE) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
        bool hit = false;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            if (fresh) {
                if (hub) presence_add(pres, c[u]);
                else cs.add(c[u]);
            }
            if (c[u] == mine) {
                const uint32_t d = gp(g.deg0)[nb[u]];
                hit |= (my_deg < d) || (my_deg == d && v < nb[u]);  // pi(v) < pi(nb)
            }
        }
        need |= ballot(hit) != 0;
    }
    if (!need) return 0;
    if (mine <= 64 && lane() == 0) or64(&g.bitmap[v], 1ull << (mine - 1));
    if (fresh && !hub) {   // own colour OR-ed in (A/src/kernels/conflict_repair.cu:105-107)
        if (mine < 64) cs.m[0] |= 1ull << mine;
        if (mine == 64) cs.m[1] |= 1ull;
    }
    const uint16_t c = fresh ? fresh_pick(g, v, hub, cs, pres) : find_color(g, v);
    if (lane() == 0) {
        st(&g.word[v], bump(w, c));
        mark(g, v, D_BITS | D_STATUS);   // its own colour bit was OR-ed in
    }
    mark_neighbors(g, v, mine, c);
    return 1;
}

// The same for a long row, all threads of the block: the verdict is a block
// OR; the first wavefront selects and stores; the block marks the neighbours.
__device__ __forceinline__ uint32_t tier1_vertex_block(const GraphDev& g, uint32_t v, int fresh_level) {
    __shared__ uint32_t pick;
    __shared__ uint64_t sets[BLOCK / WAVE][4];
    __shared__ uint32_t pres[PRESENCE_WORDS];
    const uint32_t w = ld(&g.word[v]);
    const uint16_t mine = color_of(w);
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    const uint32_t my_deg = (g.flags & ABL_LIVE_DEGREE) ? live_row_count<true>(g, v) : gp(g.deg0)[v];
    const bool hub = select_by_hist(g, v);
    const bool fresh = fresh_level >= (hub ? 2 : 1);   // see tier1_vertex
    ColourSet cs;
    uint32_t hit = 0;
    if (fresh && hub) {
        if (threadIdx.x < PRESENCE_WORDS) pres[threadIdx.x] = 0;
        __syncthreads();
    }
    for (uint32_t base = s; base < e; base += ROW_UNROLL * BLOCK) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * BLOCK + threadIdx.x;
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++)
            c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            if (fresh && hub) presence_add(pres, c[u]);
            else if (fresh) cs.add(c[u]);
            if (mine == 0) {
                hit |= nb[u] != INVALID;
            } else if (c[u] == mine) {
                const uint32_t d = gp(g.deg0)[nb[u]];
                hit |= (my_deg < d) || (my_deg == d && v < nb[u]);
            }
        }
    }
    if (block_sum32(hit) == 0) return 0;   // (its barrier also completes the presence bits)
    if (fresh && !hub) {
        cs.wave_reduce();
        if (lane() == 0)
            for (int k = 0; k < 4; k++) sets[threadIdx.x / WAVE][k] = cs.m[k];
        __syncthreads();
    }
    if (threadIdx.x < WAVE) {
        if (mine != 0 && mine <= 64 && lane() == 0) or64(&g.bitmap[v], 1ull << (mine - 1));
        uint16_t c;
        if (!fresh) {
            c = find_color(g, v);
        } else if (!hub) {
            ColourSet all;
            for (uint32_t wv = 0; wv < BLOCK / WAVE; wv++)
                for (int k = 0; k < 4; k++) all.m[k] |= sets[wv][k];
            if (mine != 0 && mine < 64) all.m[0] |= 1ull << mine;
            if (mine == 64) all.m[1] |= 1ull;
            c = all.lowest_free();
        } else {
            c = presence_lowest_free(pres);
            if (!c) c = fas_hist(g, gp(g.hd_map)[v]);
        }
        if (lane() == 0) {
            st(&g.word[v], bump(w, c));
            mark(g, v, D_BITS | D_STATUS);
            pick = c;
        }
    }
    __syncthreads();
    const uint16_t c = (uint16_t)pick;
    __syncthreads();
    if (g.dirty) {
        for (uint32_t base = s; base < e; base += ROW_UNROLL * BLOCK) {
            uint32_t nb[ROW_UNROLL], cn[ROW_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) {
                const uint32_t j = base + u * BLOCK + threadIdx.x;
                nb[u] = j < e ? ld(&g.col[j]) : INVALID;
            }
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++)
                cn[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++)
                if (nb[u] != INVALID && nb[u] < g.n)
                    mark(g, nb[u], D_BITS | ((cn[u] == mine || cn[u] == c || cn[u] == 0) ? D_STATUS : 0u));
        }
    }
    return 1;
}

__global__ __launch_bounds__(256) void k_tier1_elem(const GraphDev* __restrict__ graphs, Space sp, SweepList sl,
                                                    int fresh, uint64_t* totals) {
    __shared__ uint32_t presence[BLOCK / WAVE][PRESENCE_WORDS];
    uint64_t rows = 0, slots = 0;
    const uint32_t nbig = sweep_big(sl);
    for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
        uint32_t gid, v;
        sp.locate(big_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t f = gp(g.fill)[v];
        const uint32_t r = tier1_vertex_block(g, v, fresh);
        if (threadIdx.x == 0) {
            if (r) add32(&g.ctr[C_CONFLICTS], 1u);
            rows++;
            slots += f;
        }
    }
    uint32_t cnt = 0, cur_gid = 0xFFFFFFFFu;
    const uint32_t total = sweep_small(sl, sp.total);
    for (uint32_t k = wave_id(); k < total; k += wave_count()) {
        uint32_t gid, v;
        sp.locate(small_item(sl, k), gid, v);
        if (gid != cur_gid) {
            if (cnt && lane() == 0) add32(&graphs[cur_gid].ctr[C_CONFLICTS], cnt);
            cnt = 0;
            cur_gid = gid;
        }
        const GraphDev& g = graphs[gid];
        rows++;
        slots += gp(g.fill)[v];
        cnt += tier1_vertex(g, v, presence[threadIdx.x / WAVE], fresh);
    }
    if (cnt && lane() == 0) add32(&graphs[cur_gid].ctr[C_CONFLICTS], cnt);
    if (lane() != 0) rows = slots = 0;
    block_add64(totals, T_SWEEP_ROWS + 2, rows);
    block_add64(totals, T_SWEEP_SLOTS + 2, slots);
}

// Whether tier1_vertex(g, v) would find v conflict-free (return 0, write
// nothing): uncoloured with no neighbour in its row, or coloured with no
// neighbour of its colour that has the higher priority.  Reads only.
__device__ __forceinline__ bool tier1_clean(const GraphDev& g, uint32_t v) {
    const uint16_t mine = color_of(bcast0(ld(&g.word[v])));
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    uint32_t my_deg = 0;
    if (mine != 0) my_deg = (g.flags & ABL_LIVE_DEGREE) ? live_row_count<false>(g, v) : gp(g.deg0)[v];
    bool bad = false;
    for (uint32_t base = s; base < e && !bad; base += ROW_UNROLL * WAVE) {
        uint32_t nb[ROW_UNROLL], c[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * WAVE + lane();
            nb[u] = j < e ? ld(&g.col[j]) : INVALID;
        }
        bool hit = false;
        if (mine == 0) {
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) hit |= nb[u] != INVALID;
        } else {
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++)
                c[u] = (nb[u] != INVALID && nb[u] < g.n) ? color_of(ld(&g.word[nb[u]])) : 0;
#pragma unroll
            for (uint32_t u = 0; u < ROW_UNROLL; u++) {
                if (c[u] == mine) {
                    const uint32_t d = gp(g.deg0)[nb[u]];
                    hit |= (my_deg < d) || (my_deg == d && v < nb[u]);  // pi(v) < pi(nb)
                }
            }
        }
        bad = ballot(hit) != 0;
    }
    return !bad;
}

// Vertex order per graph (Gauss-Seidel, deterministic), one block per graph.
// With dirty tracking, a vertex is checked only if its check bit is set (or
// it is sticky) when the sweep reaches it; the bits are read again after
// every vertex that changed (a recolour marks neighbours).  A vertex whose
// bit is clear would be found conflict-free, so the sweep changes exactly
// what the full sweep changes.
// The block's SERIAL_WAVES wavefronts take the next SERIAL_WAVES marked
// vertices of the current 1024-vertex chunk and find, at once, which of them
// are conflict-free (tier1_clean, reads only).  Wavefront 0 then goes through
// them in order: a conflict-free vertex changes nothing, so the next one's
// verdict still holds; the first one that is not is checked and repaired
// (tier1_vertex) in the state the sweep would have reached it in, and the
// marks are read again from the vertex after it.  The result is the
// one-vertex-at-a-time sweep's, in fewer round trips when most marked
// vertices turn out clean.
constexpr uint32_t SERIAL_WAVES = 4;
__global__ __launch_bounds__(SERIAL_WAVES * WAVE) void k_tier1_serial(const GraphDev* __restrict__ graphs,
                                                                     uint32_t ng,
                                                                     const uint32_t* __restrict__ active,
                                                                     int fresh, uint64_t* totals) {
    __shared__ uint32_t presence[PRESENCE_WORDS];
    __shared__ uint32_t cand[SERIAL_WAVES], clean[SERIAL_WAVES], next_from, marked[SERIAL_WAVES];
    const uint32_t wv = threadIdx.x / WAVE;
    uint64_t rows = 0, slots = 0;
    for (uint32_t gid = blockIdx.x; gid < ng; gid += gridDim.x) {
        if (active && !active[gid]) continue;
        const GraphDev g = graphs[gid];
        uint32_t cnt = 0;
        if (!g.dirty) {   // every vertex, on wavefront 0 alone
            if (wv == 0) {
                for (uint32_t v = 0; v < g.n; v++) {
                    slots += gp(g.fill)[v];
                    cnt += tier1_vertex(g, v, presence, fresh);
                }
                rows += g.n;
            }
        } else {
            // each lane of wavefront 0 reads the marks of 16 vertices of the
            // chunk as two 8-byte words (mark arrays are padded to a multiple of 16)
            const size_t cst = (size_t)M_CHECK * g.dstride, sst = (size_t)M_STICKY * g.dstride;
            bool look = true;   // the last chunk had nothing marked (block-uniform)
            for (uint32_t base = 0; base < g.n; base += 16 * WAVE) {
                // After a chunk with nothing marked, skip unmarked chunks
                // SERIAL_WAVES at a time: each wavefront reads the marks of one
                // of the next chunks in the same round trip, and the sweep goes
                // on at the first chunk with any.  The chunks skipped stay
                // unmarked: only this block's repairs mark this graph's
                // vertices, and none ran since the read.
                if (look) {
                    const uint32_t c0 = base + 16 * WAVE * wv, a = c0 + 16 * lane();
                    bool any = false;
                    if (a < g.n) {
                        const uint64_t* cw = (const uint64_t*)(g.dirty + cst + a);
                        const uint64_t* sw = (const uint64_t*)(g.dirty + sst + a);
                        const uint32_t k = g.n - a >= 16 ? 16u : g.n - a;   // live vertices of the 16
                        const uint64_t lo = k >= 8 ? ~0ull : (1ull << (8 * k)) - 1ull;
                        const uint64_t hi = k >= 16 ? ~0ull : k <= 8 ? 0ull : (1ull << (8 * (k - 8))) - 1ull;
                        any = (((ld(&cw[0]) | gp(sw)[0]) & lo) | ((ld(&cw[1]) | gp(sw)[1]) & hi)) != 0;
                    }
                    const uint64_t m = ballot(any);
                    if (lane() == 0) marked[wv] = m != 0;
                    __syncthreads();
                    uint32_t first = SERIAL_WAVES;
                    for (uint32_t f = SERIAL_WAVES; f-- > 0;)
                        if (marked[f]) first = f;
                    __syncthreads();
                    if (first == SERIAL_WAVES) {   // nothing marked in the next SERIAL_WAVES chunks
                        base += (SERIAL_WAVES - 1) * 16 * WAVE;
                        continue;
                    }
                    base += first * 16 * WAVE;
                }
                const uint32_t v0 = base + 16 * lane();
                const uint64_t* cw = (const uint64_t*)(g.dirty + cst + v0);
                const uint64_t* sw = (const uint64_t*)(g.dirty + sst + v0);
                uint64_t s0 = 0, s1 = 0;
                if (wv == 0 && v0 < g.n) {
                    s0 = gp(sw)[0];
                    s1 = gp(sw)[1];
                }
                const uint32_t live = v0 >= g.n ? 0u : g.n - v0 >= 16 ? 0xFFFFu : (1u << (g.n - v0)) - 1u;
                uint32_t from = 0;   // chunk offset of the next vertex the sweep may visit (block-uniform)
                look = true;
                for (;;) {
                    if (wv == 0) {   // the next marked vertices
                        const uint64_t w0 = (v0 < g.n ? ld(&cw[0]) : 0ull) | s0;
                        const uint64_t w1 = (v0 < g.n ? ld(&cw[1]) : 0ull) | s1;
                        uint32_t m16 = 0;
#pragma unroll
                        for (uint32_t q = 0; q < 8; q++) {
                            m16 |= (((w0 >> (8 * q)) & 0xFFull) != 0) << q;
                            m16 |= (((w1 >> (8 * q)) & 0xFFull) != 0) << (q + 8);
                        }
                        m16 &= live;
                        const uint32_t at = 16 * lane();
                        m16 &= from <= at ? 0xFFFFu : from >= at + 16 ? 0u : ~((1u << (from - at)) - 1u) & 0xFFFFu;
                        for (uint32_t f = 0; f < SERIAL_WAVES; f++) {
                            const uint64_t any = ballot(m16 != 0);
                            uint32_t u = NONE;
                            if (any) {
                                const uint32_t k = (uint32_t)__builtin_ctzll(any);
                                const uint32_t q = (uint32_t)__builtin_ctz(__builtin_amdgcn_readlane(m16, k));
                                u = base + 16 * k + q;
                                if (lane() == k) m16 &= m16 - 1u;   // q was its lowest bit
                            }
                            if (lane() == 0) cand[f] = u;
                        }
                    }
                    __syncthreads();
                    const bool stop = cand[0] == NONE;
                    const uint32_t mine = cand[wv];
                    const bool ok = !stop && mine != NONE && tier1_clean(g, mine);
                    if (lane() == 0) clean[wv] = ok ? 1u : 0u;
                    __syncthreads();
                    if (stop) break;
                    look = false;
                    if (wv == 0) {   // in vertex order, up to the first vertex that is not clean
                        uint32_t last = base + from;
                        for (uint32_t f = 0; f < SERIAL_WAVES; f++) {
                            const uint32_t u = cand[f];
                            if (u == NONE) break;
                            if (lane() == 0) gp(g.dirty)[cst + u] = 0;
                            slots += gp(g.fill)[u];
                            rows++;
                            last = u + 1;
                            if (clean[f]) continue;
                            cnt += tier1_vertex(g, u, presence, fresh);
                            break;
                        }
                        if (lane() == 0) next_from = last - base;
                    }
                    __syncthreads();
                    from = next_from;
                }
            }
        }
        if (wv == 0 && lane() == 0 && cnt) add32(&g.ctr[C_CONFLICTS], cnt);
    }
    if (wv == 0 && lane() == 0 && totals && rows) {
        tot_add(totals, T_SWEEP_ROWS + 4, rows);
        tot_add(totals, T_SWEEP_SLOTS + 4, slots);
    }
}

// Vertices the host Tier-2 recoloured: their own status and their
// neighbourhoods are marked as a recolour on the device marks them.  One
// wavefront per (graph, vertex, old colour) item.
__global__ __launch_bounds__(256) void k_mark_recolored(const GraphDev* __restrict__ graphs,
                                                        const uint32_t* __restrict__ items, uint32_t count) {
    for (uint32_t k = wave_id(); k < count; k += wave_count()) {
        const GraphDev& g = graphs[items[3 * (size_t)k]];
        const uint32_t v = items[3 * (size_t)k + 1];
        if (v >= g.n) continue;
        const uint16_t oldc = (uint16_t)items[3 * (size_t)k + 2];
        const uint16_t newc = color_of(gp(g.word)[v]);
        if (lane() == 0) mark(g, v, D_STATUS);
        mark_neighbors(g, v, oldc, newc);
    }
}

// ---------------------------------------------------------------- Tier-2 on the device
// cpu_serial_repair (A/src/streamgc.cu:1057-1240) without leaving the GPU,
// with the same result as the host restatement: the adjacency of a vertex is
// its row plus every vertex whose row holds it ("reverse-only" neighbours
// exist only where the two rows of an edge disagree, i.e. between sticky
// vertices); uncoloured vertices with an edge are coloured first-fit in id
// order; then each pass takes the higher-id endpoint of every same-coloured
// edge and recolours them in ascending order with the lowest colour in
// [1, 1024) no neighbour holds.  The sequential parts run on one 1024-thread
// block per graph; the scans over all rows are grid-wide.
struct RevPair {
    uint32_t gid, v, u;   // u holds v in its row; v's row lacks u
};
constexpr uint32_t T2_BLOCK = 1024;

__device__ __forceinline__ bool mark_byte(const GraphDev& g, uint32_t kind, uint32_t v) {
    return gp(g.dirty)[(size_t)kind * g.dstride + v] != 0;
}

// Clear the pass flags and the reverse-only flags of the graphs taking part.
__global__ __launch_bounds__(256) void k_t2_clear(const GraphDev* __restrict__ graphs, Space sp,
                                                  const uint32_t* __restrict__ active, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sp.total; i += gridDim.x * blockDim.x) {
        uint32_t gid, v;
        sp.locate(i, gid, v);
        if (active && !active[gid]) continue;
        const GraphDev& g = graphs[gid];
        gp(g.dirty)[(size_t)M_TODO * g.dstride + v] = 0;
        gp(g.dirty)[(size_t)M_REV * g.dstride + v] = 0;
    }
}

// Reverse-only neighbours: for a sticky u and a sticky v in u's row, (v, u)
// when v's row lacks u.  One wavefront per vertex.
__global__ __launch_bounds__(256) void k_t2_rev(const GraphDev* __restrict__ graphs, Space sp,
                                                const uint32_t* __restrict__ active, RevPair* __restrict__ rev,
                                                uint32_t* __restrict__ rev_count, uint32_t cap,
                                                const uint32_t* __restrict__ dspace, uint32_t ng,
                                                const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    if (dspace) sp = device_space(dspace, ng);
    for (uint32_t i = wave_id(); i < sp.total; i += wave_count()) {
        uint32_t gid, u;
        sp.locate(i, gid, u);
        if (active && !active[gid]) continue;
        const GraphDev& g = graphs[gid];
        if (!mark_byte(g, M_STICKY, u)) continue;
        const uint32_t s = gp(g.row_ptr)[u], e = s + ld(&g.fill[u]);
        for (uint32_t base = s; base < e; base += WAVE) {
            const uint32_t j = base + lane();
            const uint32_t v = j < e ? ld(&g.col[j]) : INVALID;
            uint64_t m = ballot(v != INVALID && v < g.n && v != u && mark_byte(g, M_STICKY, v));
            while (m) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t w = (uint32_t)__shfl((int)v, (int)k);
                const uint32_t s2 = gp(g.row_ptr)[w], e2 = s2 + ld(&g.fill[w]);
                bool found = false;
                for (uint32_t b2 = s2; b2 < e2 && !found; b2 += WAVE) {
                    const uint32_t j2 = b2 + lane();
                    found = ballot(j2 < e2 && ld(&g.col[j2]) == u) != 0;
                }
                if (!found && lane() == 0) {
                    const uint32_t at = add32(rev_count, 1u);
                    if (at < cap) {
                        st(&rev[at].gid, gid);
                        st(&rev[at].v, w);
                        st(&rev[at].u, u);
                    } else {
                        add32(&g.ctr[C_REVOVF], 1u);
                    }
                    gp(g.dirty)[(size_t)M_REV * g.dstride + w] = 1;
                }
            }
        }
    }
}

// first_fit of the host restatement for vertex v, by the whole block.
// Returns true when v changed colour (it is then marked like any recolour).
__device__ __forceinline__ bool t2_first_fit(const GraphDev& g, uint32_t gid, uint32_t v,
                                             const RevPair* __restrict__ rev, uint32_t nrev) {
    __shared__ uint32_t seen[HIST_BINS / 32];
    __shared__ uint32_t pick;
    if (threadIdx.x < HIST_BINS / 32) seen[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t s = gp(g.row_ptr)[v], e = s + ld(&g.fill[v]);
    for (uint32_t j = s + threadIdx.x; j < e; j += T2_BLOCK) {
        const uint32_t nb = ld(&g.col[j]);
        if (nb == INVALID || nb >= g.n || nb == v) continue;
        const uint32_t c = color_of(ld(&g.word[nb]));
        if (c != 0 && c < HIST_BINS) atomicOr(&seen[c >> 5], 1u << (c & 31));
    }
    if (mark_byte(g, M_REV, v))
        for (uint32_t k = threadIdx.x; k < nrev; k += T2_BLOCK) {
            if (rev[k].gid != gid || rev[k].v != v) continue;
            const uint32_t c = color_of(ld(&g.word[rev[k].u]));
            if (c != 0 && c < HIST_BINS) atomicOr(&seen[c >> 5], 1u << (c & 31));
        }
    __syncthreads();
    if (threadIdx.x < WAVE) {
        uint32_t word = threadIdx.x < HIST_BINS / 32 ? seen[threadIdx.x] : 0xFFFFFFFFu;
        if (threadIdx.x == 0) word |= 1u;   // colour 0 is not a colour
        const uint64_t open = ballot(word != 0xFFFFFFFFu);
        uint32_t c = 0;
        if (open) {
            const uint32_t k = (uint32_t)__builtin_ctzll(open);
            const uint32_t wk = (uint32_t)__shfl((int)word, (int)k);
            c = 32 * k + (uint32_t)__builtin_ctz(~wk);
        }
        if (threadIdx.x == 0) {
            const uint32_t w = ld(&g.word[v]);
            const uint32_t oc = color_of(w);
            if (c == 0 || c == oc) {
                pick = 0;
            } else {
                st(&g.word[v], bump(w, (uint16_t)c));
                pick = 0x80000000u | (oc << 16) | c;
            }
        }
    }
    __syncthreads();
    const uint32_t p = pick;
    __syncthreads();
    if (!p) return false;
    if (threadIdx.x < WAVE) {
        if (lane() == 0) mark(g, v, D_STATUS);
        mark_neighbors(g, v, (uint16_t)((p >> 16) & 0x7FFFu), (uint16_t)(p & 0xFFFFu));
    }
    return true;
}

// Vertices of one graph flagged in mark array `kind` (or, for kind ==
// M_ARRAYS, uncoloured vertices with an edge), visited in ascending order by
// the block: flags are cleared and first_fit applied.  Returns the recolours.
__device__ __forceinline__ uint32_t t2_ordered_pass(const GraphDev& g, uint32_t gid, uint32_t kind,
                                                    const RevPair* __restrict__ rev, uint32_t nrev) {
    __shared__ uint64_t cand[T2_BLOCK / WAVE];
    __shared__ uint32_t has;
    uint32_t fixed = 0;
    for (uint32_t base = 0; base < g.n; base += T2_BLOCK) {
        const uint32_t v = base + threadIdx.x;
        bool c = false;
        if (v < g.n) {
            if (kind == M_ARRAYS) {
                c = color_of(ld(&g.word[v])) == 0;
            } else if (mark_byte(g, kind, v)) {
                c = true;
                gp(g.dirty)[(size_t)kind * g.dstride + v] = 0;
            }
        }
        const uint64_t m = ballot(c);
        if (lane() == 0) cand[threadIdx.x / WAVE] = m;
        __syncthreads();
        for (uint32_t w = 0; w < T2_BLOCK / WAVE; w++) {
            uint64_t mw = cand[w];
            while (mw) {
                const uint32_t b = (uint32_t)__builtin_ctzll(mw);
                mw &= mw - 1;
                const uint32_t u = base + w * WAVE + b;
                if (kind == M_ARRAYS) {   // only a vertex with an edge is coloured
                    if (threadIdx.x == 0) has = mark_byte(g, M_REV, u) ? 1u : 0u;
                    __syncthreads();
                    const uint32_t s = gp(g.row_ptr)[u], e = s + ld(&g.fill[u]);
                    bool any = false;
                    for (uint32_t j = s + threadIdx.x; j < e && !any; j += T2_BLOCK) {
                        const uint32_t nb = ld(&g.col[j]);
                        any = nb != INVALID && nb < g.n && nb != u;
                    }
                    if (any) has = 1;
                    __syncthreads();
                    const bool go = has != 0;
                    __syncthreads();
                    if (!go) continue;
                }
                fixed += t2_first_fit(g, gid, u, rev, nrev) ? 1u : 0u;
            }
        }
        __syncthreads();
    }
    return fixed;
}

// The uncoloured pass (pass == 0) or one conflict pass over the flagged
// vertices, one block per graph taking part.
__global__ __launch_bounds__(T2_BLOCK) void k_t2_serial(const GraphDev* __restrict__ graphs, uint32_t ng,
                                                        const uint32_t* __restrict__ active,
                                                        const RevPair* __restrict__ rev,
                                                        const uint32_t* __restrict__ rev_count, uint32_t cap,
                                                        int uncoloured, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const uint32_t nrev = *rev_count < cap ? *rev_count : cap;
    for (uint32_t gid = blockIdx.x; gid < ng; gid += gridDim.x) {
        if (active && !active[gid]) continue;
        const GraphDev g = graphs[gid];
        const uint32_t f = t2_ordered_pass(g, gid, uncoloured ? M_ARRAYS : M_TODO, rev, nrev);
        if (threadIdx.x == 0) {
            if (!uncoloured) gp(g.ctr)[C_T2FIXED] = f;
            gp(g.ctr)[C_T2TOTAL] += f;
        }
        __syncthreads();
    }
}

// The next pass's list: the higher-id endpoint of every same-coloured slot
// pair of every row.  One wavefront per row; C_T2TODO becomes 1 if any.
__global__ __launch_bounds__(256) void k_t2_flags(const GraphDev* __restrict__ graphs, Space sp,
                                                  const uint32_t* __restrict__ active,
                                                  const uint32_t* __restrict__ dspace, uint32_t ng,
                                                  const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    if (dspace) sp = device_space(dspace, ng);
    for (uint32_t i = wave_id(); i < sp.total; i += wave_count()) {
        uint32_t gid, x;
        sp.locate(i, gid, x);
        if (active && !active[gid]) continue;
        const GraphDev& g = graphs[gid];
        const uint32_t cx = color_of(ld(&g.word[x]));
        if (cx == 0) continue;
        const uint32_t s = gp(g.row_ptr)[x], e = s + ld(&g.fill[x]);
        bool any = false;
        for (uint32_t base = s; base < e; base += WAVE) {
            const uint32_t j = base + lane();
            const uint32_t y = j < e ? ld(&g.col[j]) : INVALID;
            if (y != INVALID && y < g.n && y != x && color_of(ld(&g.word[y])) == cx) {
                gp(g.dirty)[(size_t)M_TODO * g.dstride + (y > x ? y : x)] = 1;
                any = true;
            }
        }
        if (ballot(any) && lane() == 0) gp(g.ctr)[C_T2TODO] = 1;
    }
}

// ---------------------------------------------------------------- checks
// Tier-2 has work iff some row slot joins an uncoloured endpoint or two
// endpoints of one colour (A/src/streamgc.cu:1190-1218 then change nothing).
// Bad slots of row v: an uncoloured endpoint or two equal colours.
template <bool kBlock>
__device__ __forceinline__ uint32_t t2_row(const GraphDev& g, uint32_t v) {
    const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
    const uint32_t width = kBlock ? BLOCK : WAVE, me = kBlock ? threadIdx.x : lane();
    const uint16_t cv = color_of(gp(g.word)[v]);
    uint32_t bad = 0;
    for (uint32_t base = s; base < e; base += ROW_UNROLL * width) {
        uint32_t nb[ROW_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            const uint32_t j = base + u * width + me;
            nb[u] = j < e ? gp(g.col)[j] : INVALID;
        }
#pragma unroll
        for (uint32_t u = 0; u < ROW_UNROLL; u++) {
            if (nb[u] != INVALID && nb[u] < g.n && nb[u] != v) {
                const uint16_t cn = color_of(gp(g.word)[nb[u]]);
                bad += (cv == 0 || cn == 0 || cv == cn);
            }
        }
    }
    return kBlock ? block_sum32(bad) : wave_sum32(bad);
}

// Tier-2 has work iff some row slot joins an uncoloured endpoint or two
// endpoints of one colour (A/src/streamgc.cu:1190-1218 then change nothing).
// A vertex found with a bad slot keeps its mark, so the next check counts it
// again until a repair changes it.
__global__ __launch_bounds__(256) void k_t2_check(const GraphDev* __restrict__ graphs, Space sp, SweepList sl,
                                                  uint64_t* totals) {
    uint64_t rows = 0, slots = 0;
    const uint32_t nbig = sweep_big(sl);
    for (uint32_t k = blockIdx.x; k < nbig; k += gridDim.x) {
        uint32_t gid, v;
        sp.locate(big_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t bad = t2_row<true>(g, v);
        if (threadIdx.x == 0) {
            rows++;
            slots += gp(g.fill)[v];
            if (bad) {
                add32(&g.ctr[C_T2BAD], bad);
                if (g.dirty) gp(g.dirty)[(size_t)M_T2 * g.dstride + v] = 1;
            }
        }
    }
    const uint32_t total = sweep_small(sl, sp.total);
    for (uint32_t k = wave_id(); k < total; k += wave_count()) {
        uint32_t gid, v;
        sp.locate(small_item(sl, k), gid, v);
        const GraphDev& g = graphs[gid];
        const uint32_t bad = t2_row<false>(g, v);
        rows++;
        slots += gp(g.fill)[v];
        if (bad && lane() == 0) {
            add32(&g.ctr[C_T2BAD], bad);
            if (g.dirty) gp(g.dirty)[(size_t)M_T2 * g.dstride + v] = 1;
        }
    }
    if (lane() != 0) rows = slots = 0;
    block_add64(totals, T_SWEEP_ROWS + 3, rows);
    block_add64(totals, T_SWEEP_SLOTS + 3, slots);
}

// check_coloring_valid on the engine's rows (A/src/streamgc.cu:448-556);
// colours from `colors` when given (a snapshot), else the live words.
__global__ __launch_bounds__(256) void k_validate(const GraphDev* __restrict__ graphs, uint32_t gid,
                                                  const uint16_t* __restrict__ colors) {
    const GraphDev& g = graphs[gid];
    uint32_t viol = 0, unc = 0, mx = 0;
    for (uint32_t v = wave_id(); v < g.n; v += wave_count()) {
        const uint32_t s = gp(g.row_ptr)[v], e = s + gp(g.fill)[v];
        const uint16_t cv = colors ? colors[v] : color_of(gp(g.word)[v]);
        uint32_t bad = 0, has = 0;
        for (uint32_t base = s; base < e; base += WAVE) {
            const uint32_t j = base + lane();
            const uint32_t nb = j < e ? gp(g.col)[j] : INVALID;
            if (nb != INVALID) {
                has = 1;
                if (nb < g.n) {
                    const uint16_t cn = colors ? colors[nb] : color_of(gp(g.word)[nb]);
                    bad += (cv != 0 && cv == cn);
                }
            }
        }
        viol += wave_sum32(bad);
        unc += (cv == 0 && ballot(has != 0)) ? 1 : 0;
        mx = cv > mx ? cv : mx;
    }
    if (lane() == 0) {
        if (viol) add32(&g.ctr[C_VIOL], viol);
        if (unc) add32(&g.ctr[C_UNCOL], unc);
        max32(&g.ctr[C_MAXC], mx);
    }
}

__global__ void k_find_color(const GraphDev* __restrict__ graphs, uint32_t gid, uint32_t v, uint16_t* out) {
    const GraphDev g = graphs[gid];
    const uint16_t c = find_color(g, v);
    if (lane() == 0) *out = c;
}

// Independent validator: edges with life intervals checked against snapshot
// k at stream position pos[k] (blockIdx.y = k).
__global__ __launch_bounds__(256) void k_validate_history(const uint16_t* __restrict__ snaps, uint32_t n,
                                                          const uint32_t* __restrict__ eu,
                                                          const uint32_t* __restrict__ ev,
                                                          const uint64_t* __restrict__ born,
                                                          const uint64_t* __restrict__ died, uint64_t m,
                                                          const uint64_t* __restrict__ pos,
                                                          unsigned long long* bad_out) {
    const uint32_t k = blockIdx.y;
    const uint16_t* snap = snaps + (size_t)k * n;
    const uint64_t p = pos[k];
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        if (born[i] <= p && p < died[i]) {
            const uint16_t a = snap[eu[i]], b = snap[ev[i]];
            bad += (a == 0) + (b == 0) + (a != 0 && a == b);
        }
    }
    bad = wave_sum32(bad);
    if (bad && lane() == 0) atomicAdd(&bad_out[k], (unsigned long long)bad);
}

// Slack-CSR construction (A/src/streamgc.cu:350-433): copy each loaded row to
// the front of its slack row, free slots after it.
__global__ __launch_bounds__(256) void k_fill_rows(uint32_t n, const uint32_t* __restrict__ rp0,
                                                   const uint32_t* __restrict__ col0,
                                                   const uint32_t* __restrict__ rp, uint32_t* __restrict__ col,
                                                   uint32_t* __restrict__ fill) {
    for (uint32_t v = wave_id(); v < n; v += wave_count()) {
        const uint32_t s0 = rp0[v], d = rp0[v + 1] - s0;
        const uint32_t s = rp[v], e = rp[v + 1];
        if (lane() == 0) fill[v] = d < e - s ? d : e - s;
        for (uint32_t j = s + lane(); j < e; j += WAVE) {
            const uint32_t k = j - s;
            col[j] = k < d ? col0[s0 + k] : INVALID;
        }
    }
}

}  // namespace sgc

// ---------------------------------------------------------------- launchers
namespace sgc {

static inline Space space(const LaunchCtx& L) { return Space{L.sbase, L.sgid, L.ns, L.V}; }

static inline uint32_t grid_for_waves(uint64_t waves, uint32_t waves_per_block, uint32_t cap) {
    uint64_t b = (waves + waves_per_block - 1) / waves_per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (uint32_t)b;
}

hipError_t launch_apply(const LaunchCtx& L, const void* upd, uint32_t count, int multi, const EdgeGroups* eg) {
    if (count == 0) return hipSuccess;
    const uint32_t *head = nullptr, *next = nullptr, *heads = nullptr, *nheads = nullptr;
    if (eg && eg->keys && count > 1) {
        if ((uint64_t)eg->size < 2ull * count || eg->next_cap < count || !eg->heads || !eg->nheads)
            return hipErrorInvalidValue;
        head = eg->head;
        next = eg->next;
        heads = eg->heads;
        nheads = eg->nheads;
    }
    // one-wavefront blocks: a wavefront's slot is refilled as soon as it ends
    uint32_t grid = grid_for_waves(count, 1, 4 * L.max_blocks);
    k_apply_elem<<<grid, WAVE, 0, L.stream>>>(L.graphs, (const uint4*)upd, count, multi, head, next, heads,
                                               eg ? eg->next_cap : 0u, nheads, heads ? eg->pack : nullptr,
                                               L.totals, L.dyn);
    return hipGetLastError();
}
hipError_t launch_group(const LaunchCtx& L, const void* upd, uint32_t count, int multi, const EdgeGroups* eg) {
    if (count <= 1 || !eg || !eg->keys) return hipSuccess;
    if ((uint64_t)eg->size < 2ull * count || eg->next_cap < count) return hipErrorInvalidValue;
    const uint32_t mask = eg->size - 1;
    uint32_t cg = grid_for_waves(eg->size, 4 * WAVE, 1024);
    k_group_clear<<<cg, 256, 0, L.stream>>>(eg->keys, eg->head, eg->size, eg->nheads);
    uint32_t mg = grid_for_waves(count, 4 * WAVE, 1024);
    k_group_link<<<mg, 256, 0, L.stream>>>(L.graphs, L.gbase, (const uint4*)upd, count, multi, eg->keys, eg->head,
                                           eg->next, mask, eg->heads, eg->next_cap, eg->nheads, L.dyn);
    if (eg->pack)
        k_group_pack<<<mg, 256, 0, L.stream>>>((const uint4*)upd, eg->head, eg->next, eg->heads, eg->next_cap,
                                               eg->nheads, eg->pack, L.dyn);
    return hipGetLastError();
}
hipError_t launch_apply_seg(const LaunchCtx& L, const void* upd, const uint32_t* seg, uint32_t nseg,
                            const uint32_t* perm, int multi) {
    if (nseg == 0) return hipSuccess;
    uint32_t grid = grid_for_waves(nseg, 1, L.max_blocks);
    k_apply_seg<<<grid, 64, 0, L.stream>>>(L.graphs, (const uint4*)upd, seg, nseg, perm, multi, L.totals, L.dyn);
    return hipGetLastError();
}
hipError_t launch_zero_ctr(const LaunchCtx& L, uint32_t idx) {
    k_zero_ctr<<<(L.ng + 63) / 64, 64, 0, L.stream>>>(L.graphs, L.ng, L.active, idx);
    return hipGetLastError();
}
hipError_t launch_max_color(const LaunchCtx& L) {
    if (L.ns == 1) {
        k_max_color1<<<256, 256, 0, L.stream>>>(L.graphs, space(L));
        return hipGetLastError();
    }
    uint32_t grid = grid_for_waves(L.V, 4 * WAVE, 1024);
    k_max_color<<<grid, 256, 0, L.stream>>>(L.graphs, space(L));
    return hipGetLastError();
}
hipError_t launch_snapshot(const LaunchCtx& L, const uint32_t* fl_round, const HistDev* hist) {
    if (!fl_round) hist = nullptr;
    if (L.ns == 1) {
        const uint32_t grid = grid_for_waves(L.V / 4 + 1, 4 * WAVE, L.max_blocks);
        k_snapshot1<<<grid, 256, 0, L.stream>>>(L.graphs, space(L), L.active, fl_round, hist);
        return hipGetLastError();
    }
    uint32_t grid = grid_for_waves(L.V, 4 * WAVE, L.max_blocks);
    k_snapshot<<<grid, 256, 0, L.stream>>>(L.graphs, space(L), L.active, fl_round, hist);
    return hipGetLastError();
}
// Sweep grids: one wavefront per row (blocks first take the long rows of a
// worklist), grid-stride; a worklist sweep takes the grid of a full sweep
// capped at sweep_blocks (its length is on the device).
static inline uint32_t sweep_grid(const LaunchCtx& L) {
    return grid_for_waves(L.V, 4, L.list ? L.sweep_blocks : L.max_blocks);
}
static inline SweepList sweep_list(const LaunchCtx& L) { return SweepList{L.list, L.list_count, L.list_cap}; }
hipError_t launch_build_list(const LaunchCtx& L, uint32_t kind, uint32_t* list, uint32_t cap, uint32_t* count) {
    if (kind >= M_STICKY || cap < L.V) return hipErrorInvalidValue;
    // Two passes over 4096-vertex tiles, each tile inside one graph (scratch
    // behind the list): one graph's own tiles, or the tiles of several graphs
    // laid end to end (L.stile, L.tiles).
    const bool multi = L.ns > 1;
    const uint32_t tiles = multi ? L.tiles : (L.V + LIST_TILE_V - 1) / LIST_TILE_V;
    if (multi && (!L.stile || tiles > list_tiles(cap, L.ng))) return hipErrorInvalidValue;
    if (tiles == 0) {
        k_zero_list_count<<<1, 1, 0, L.stream>>>(count);
        return hipGetLastError();
    }
    uint32_t* tmp = list + cap;
    uint32_t* tcount = tmp + (size_t)tiles * LIST_TILE_V;
    k_list_tiles<<<tiles, 256, 0, L.stream>>>(L.graphs, space(L), multi ? L.stile : nullptr, kind, L.long_row, tmp,
                                              tcount, L.active, L.gate);
    k_list_gather<<<tiles, 256, 0, L.stream>>>(L.graphs, space(L), multi ? 1 : 0, tmp, tcount, list, cap, count,
                                               L.active, L.gate);
    return hipGetLastError();
}
hipError_t launch_compact(const LaunchCtx& L, bool drain_prep) {
    k_compact<<<sweep_grid(L), 256, 0, L.stream>>>(L.graphs, space(L), sweep_list(L), L.totals,
                                                   drain_prep ? L.ng : 0u, L.active);
    return hipGetLastError();
}
hipError_t launch_drain(const LaunchCtx& L, bool serial, uint32_t max_count, bool prepped) {
    if (!prepped) k_drain_prep<<<(L.ng + 63) / 64, 64, 0, L.stream>>>(L.graphs, L.ng, L.active);
    if (serial) {
        k_drain_serial<<<L.ng < 1024 ? L.ng : 1024, 64, 0, L.stream>>>(L.graphs, L.ng, L.active);
    } else {
        // entries of one graph share grid.x blocks; many graphs take fewer each
        const uint32_t per_graph = L.ng > 4 ? (16384 / L.ng > 0 ? 16384 / L.ng : 1) : 4096;
        dim3 grid(grid_for_waves(max_count, 4, per_graph), L.ng);
        k_drain_elem<<<grid, 256, 0, L.stream>>>(L.graphs, L.active);
    }
    return hipGetLastError();
}
hipError_t launch_rebuild(const LaunchCtx& L, bool exact_hist) {
    k_rebuild<<<sweep_grid(L), 256, 0, L.stream>>>(L.graphs, space(L), sweep_list(L),
                                                   exact_hist ? 1 : 0, L.totals);
    return hipGetLastError();
}
hipError_t launch_tier1(const LaunchCtx& L, bool serial, int fresh) {
    if (serial) {
        const uint32_t grid = L.ng < 1024 ? L.ng : 1024;
        k_tier1_serial<<<grid, SERIAL_WAVES * WAVE, 0, L.stream>>>(L.graphs, L.ng, L.active, fresh, L.totals);
    } else {
        k_tier1_elem<<<sweep_grid(L), 256, 0, L.stream>>>(L.graphs, space(L), sweep_list(L), fresh,
                                                          L.totals);
    }
    return hipGetLastError();
}
hipError_t launch_t2_check(const LaunchCtx& L) {
    k_t2_check<<<sweep_grid(L), 256, 0, L.stream>>>(L.graphs, space(L), sweep_list(L),
                                                    L.totals);
    return hipGetLastError();
}
hipError_t launch_mark_recolored(hipStream_t s, const GraphDev* graphs, const uint32_t* items, uint32_t count) {
    if (count == 0) return hipSuccess;
    k_mark_recolored<<<grid_for_waves(count, 4, 4096), 256, 0, s>>>(graphs, items, count);
    return hipGetLastError();
}
hipError_t launch_t2_begin(const LaunchCtx& L, void* rev, uint32_t* rev_count, uint32_t cap) {
    k_zero_list_count<<<1, 1, 0, L.stream>>>(rev_count);   // [0] is the pair count
    k_t2_clear<<<grid_for_waves(L.V, 4 * WAVE, 4096), 256, 0, L.stream>>>(L.graphs, space(L), L.active, L.gate);
    k_t2_rev<<<grid_for_waves(L.V, 4, L.max_blocks), 256, 0, L.stream>>>(L.graphs, space(L), L.active,
                                                                         (RevPair*)rev, rev_count, cap, L.dspace,
                                                                         L.ng, L.gate);
    return hipGetLastError();
}
hipError_t launch_t2_pass(const LaunchCtx& L, const void* rev, const uint32_t* rev_count, uint32_t cap,
                          bool uncoloured) {
    if (!uncoloured)
        k_t2_flags<<<grid_for_waves(L.V, 4, L.max_blocks), 256, 0, L.stream>>>(L.graphs, space(L), L.active,
                                                                               L.dspace, L.ng, L.gate);
    k_t2_serial<<<L.ng < 1024 ? L.ng : 1024, T2_BLOCK, 0, L.stream>>>(L.graphs, L.ng, L.active, (const RevPair*)rev,
                                                                      rev_count, cap, uncoloured ? 1 : 0, L.gate);
    return hipGetLastError();
}
hipError_t launch_validate(const LaunchCtx& L, uint32_t gid, uint32_t n, const uint16_t* colors) {
    uint32_t grid = grid_for_waves(n, 4, L.max_blocks);
    k_validate<<<grid, 256, 0, L.stream>>>(L.graphs, gid, colors);
    return hipGetLastError();
}
hipError_t launch_find_color(const LaunchCtx& L, uint32_t gid, uint32_t v, uint16_t* out) {
    k_find_color<<<1, 64, 0, L.stream>>>(L.graphs, gid, v, out);
    return hipGetLastError();
}
hipError_t launch_validate_history(hipStream_t s, const uint16_t* snaps, uint32_t n, const uint32_t* eu,
                                   const uint32_t* ev, const uint64_t* born, const uint64_t* died, uint64_t m,
                                   const uint64_t* pos, uint32_t count, unsigned long long* bad_out) {
    if (count == 0 || m == 0) return hipSuccess;
    uint64_t blocks = (m + 255) / 256;
    dim3 grid((uint32_t)(blocks < 2048 ? blocks : 2048), count);
    k_validate_history<<<grid, 256, 0, s>>>(snaps, n, eu, ev, born, died, m, pos, bad_out);
    return hipGetLastError();
}
hipError_t launch_fill_rows(hipStream_t s, uint32_t n, const uint32_t* rp0, const uint32_t* col0, const uint32_t* rp,
                            uint32_t* col, uint32_t* fill) {
    uint32_t grid = grid_for_waves(n, 4, 16384);
    k_fill_rows<<<grid, 256, 0, s>>>(n, rp0, col0, rp, col, fill);
    return hipGetLastError();
}

// ---------------------------------------------------------------- T_rc microbenchmark
// A/benchmark/microbench/t_rc_atomic.cu:18-28: one lane runs `iters`
// dependent CAS operations on a vertex word, each bumping the version and
// toggling the colour's low bit (a failed CAS takes the value it found).
// Here every wavefront's lane 0 does so, on word (wave x stride): stride 32
// gives each wavefront its own 128-byte line, stride 0 makes them all share
// one word.
__global__ __launch_bounds__(64) void k_trc(uint32_t* __restrict__ words, uint32_t stride, uint32_t iters) {
    if (threadIdx.x != 0) return;
    uint32_t* w = words + (size_t)blockIdx.x * stride;
    uint32_t cur = ld(w);
    for (uint32_t i = 0; i < iters; i++) {
        const uint32_t nxt = bump(cur, (uint16_t)((cur & 0xFFFFu) ^ 1u));
        const uint32_t got = cas32(w, cur, nxt);
        cur = got == cur ? nxt : got;
    }
}
hipError_t launch_trc(hipStream_t s, uint32_t* words, uint32_t waves, uint32_t stride, uint32_t iters) {
    k_trc<<<waves, WAVE, 0, s>>>(words, stride, iters);
    return hipGetLastError();
}

}  // namespace sgc
