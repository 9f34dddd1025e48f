// sgc_host.cpp -- host pipeline of the StreamGC engine (C++17): .mtx loader,
// seeded update streams, Tier-2 host repair, static greedy, edge intervals.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

#include "sgc_internal.hpp"
#include "streamgc.h"

namespace sgc {

// ---------------------------------------------------------------- random numbers
// mt19937_64 exactly as [rand.eng.mt] specifies it (the reference draws from
// std::mt19937_64, A/src/host/stream_gen.cpp:73); implemented here so the
// stream bytes do not depend on the C++ library the host is built with.
class Mt19937_64 {
  public:
    explicit Mt19937_64(uint64_t seed) {
        s_[0] = seed;
        for (int i = 1; i < kN; i++) s_[i] = 6364136223846793005ULL * (s_[i - 1] ^ (s_[i - 1] >> 62)) + (uint64_t)i;
        pos_ = kN;
    }
    uint64_t operator()() {
        if (pos_ == kN) refill();
        uint64_t z = s_[pos_++];
        z ^= (z >> 29) & 0x5555555555555555ULL;
        z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
        z ^= (z << 37) & 0xFFF7EEE000000000ULL;
        return z ^ (z >> 43);
    }

  private:
    static constexpr int kN = 312, kM = 156;
    void refill() {
        constexpr uint64_t hi = 0xFFFFFFFF80000000ULL, lo = 0x7FFFFFFFULL, a = 0xB5026F5AA96619E9ULL;
        for (int i = 0; i < kN; i++) {
            const uint64_t y = (s_[i] & hi) | (s_[(i + 1) % kN] & lo);
            s_[i] = s_[(i + kM) % kN] ^ (y >> 1) ^ ((y & 1) ? a : 0);
        }
        pos_ = 0;
    }
    uint64_t s_[kN];
    int pos_;
};

// Unbiased integer in [0, range) by Lemire's multiply-and-reject method
// (ACM TOMACS 29(1), 2019), with a 128-bit product for a 64-bit engine -- the
// draw std::uniform_int_distribution makes in GCC 11's libstdc++.
static inline uint64_t below(Mt19937_64& g, uint64_t range) {
    unsigned __int128 m = (unsigned __int128)g() * range;
    if ((uint64_t)m < range) {
        const uint64_t reject_under = (0 - range) % range;
        while ((uint64_t)m < reject_under) m = (unsigned __int128)g() * range;
    }
    return (uint64_t)(m >> 64);
}

// ---------------------------------------------------------------- u64 key set
// Open addressing with linear probing and backward-shift deletion (no
// tombstones); keys are (lo << 32) | hi with lo < hi, so ~0 is free.
class EdgeSet {
  public:
    explicit EdgeSet(uint64_t expect) {
        uint64_t cap = 16;
        while (cap < expect * 2) cap <<= 1;
        slot_.assign(cap, kFree);
        mask_ = cap - 1;
    }
    bool contains(uint64_t k) const {
        for (uint64_t i = hash(k);; i = (i + 1) & mask_) {
            if (slot_[i] == k) return true;
            if (slot_[i] == kFree) return false;
        }
    }
    bool insert(uint64_t k) {
        if ((size_ + 1) * 2 > slot_.size()) grow();
        for (uint64_t i = hash(k);; i = (i + 1) & mask_) {
            if (slot_[i] == k) return false;
            if (slot_[i] == kFree) {
                slot_[i] = k;
                size_++;
                return true;
            }
        }
    }
    bool erase(uint64_t k) {
        uint64_t i = hash(k);
        for (;; i = (i + 1) & mask_) {
            if (slot_[i] == kFree) return false;
            if (slot_[i] == k) break;
        }
        uint64_t hole = i;
        for (uint64_t j = (hole + 1) & mask_; slot_[j] != kFree; j = (j + 1) & mask_) {
            const uint64_t home = hash(slot_[j]);
            // move slot j back into the hole unless its home lies cyclically in (hole, j]
            const bool stays = hole <= j ? (hole < home && home <= j) : (hole < home || home <= j);
            if (!stays) {
                slot_[hole] = slot_[j];
                hole = j;
            }
        }
        slot_[hole] = kFree;
        size_--;
        return true;
    }
    uint64_t size() const { return size_; }

  private:
    static constexpr uint64_t kFree = ~0ULL;
    uint64_t hash(uint64_t k) const {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdULL;
        k ^= k >> 33;
        return k & mask_;
    }
    void grow() {
        std::vector<uint64_t> old;
        old.swap(slot_);
        slot_.assign(old.size() * 2, kFree);
        mask_ = slot_.size() - 1;
        size_ = 0;
        for (uint64_t k : old)
            if (k != kFree) insert(k);
    }
    std::vector<uint64_t> slot_;
    uint64_t mask_ = 0, size_ = 0;
};

static inline uint64_t ekey(uint32_t a, uint32_t b) {
    return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
}
static inline void put(sgc_edge_update& r, uint32_t u, uint32_t v, uint8_t t) {
    std::memset(&r, 0, sizeof r);
    r.u = u;
    r.v = v;
    r.type = t;
}

}  // namespace sgc

using namespace sgc;

extern "C" {

// generate_init_stream (A/src/host/stream_gen.cpp:30-61): every u < v edge in
// CSR order, sorted by (min degree desc, max degree desc) with std::sort --
// the same library sort on the same sequence, so ties land where the
// reference's land.
int sgc_init_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const uint32_t* deg,
                    sgc_edge_update* out, uint64_t* count) {
    if (!row_ptr || !col_idx || !deg || !count) return SGC_ERR_INVALID;
    std::vector<std::pair<uint32_t, uint32_t>> e;
    e.reserve(row_ptr[n] / 2);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++)
            if (u < col_idx[j]) e.emplace_back(u, col_idx[j]);
    std::sort(e.begin(), e.end(), [deg](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
        const uint32_t da = deg[a.first], db = deg[a.second], dc = deg[b.first], dd = deg[b.second];
        const uint32_t lo_a = da < db ? da : db, lo_b = dc < dd ? dc : dd;
        if (lo_a != lo_b) return lo_a > lo_b;
        return (da < db ? db : da) > (dc < dd ? dd : dc);
    });
    if (out)
        for (size_t i = 0; i < e.size(); i++) put(out[i], e[i].first, e[i].second, 0);
    *count = e.size();
    return SGC_OK;
}

// generate_bench_stream_edges (A/src/host/stream_gen.cpp:65-177), same draws in
// the same order: per event U{0..99} < pct decides insert; an insert re-adds a
// uniform pick from the deleted pool (swap-remove) or, with the pool empty,
// draws two uniform vertices (v+1 mod n on a tie) and retries on a live edge;
// a delete removes a uniform pick from the live list (swap-remove).  A retry
// re-draws the percentage too.
int sgc_churn_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint64_t size, uint32_t insert_pct,
                     uint64_t seed, sgc_edge_update* out) {
    if (!row_ptr || !col_idx || (!out && size) || n == 0) return SGC_ERR_INVALID;
    Mt19937_64 g(seed);
    const uint64_t m0 = row_ptr[n] / 2;
    EdgeSet live(m0 + size);
    std::vector<uint64_t> live_list, pool;
    live_list.reserve(m0 + size);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++)
            if (u < col_idx[j]) {
                const uint64_t k = ((uint64_t)u << 32) | col_idx[j];
                live.insert(k);
                live_list.push_back(k);
            }
    uint64_t k = 0;
    while (k < size) {
        const bool ins = below(g, 100) < insert_pct;
        if (ins) {
            if (!pool.empty()) {
                const uint64_t i = below(g, pool.size());
                const uint64_t e = pool[i];
                pool[i] = pool.back();
                pool.pop_back();
                live.insert(e);
                live_list.push_back(e);
                put(out[k++], (uint32_t)(e >> 32), (uint32_t)e, 0);
            } else {
                const uint32_t a = (uint32_t)below(g, n);
                uint32_t b = (uint32_t)below(g, n);
                if (a == b) b = (b + 1) % n;
                const uint64_t e = ekey(a, b);
                if (!live.contains(e)) {
                    live.insert(e);
                    live_list.push_back(e);
                    put(out[k++], (uint32_t)(e >> 32), (uint32_t)e, 0);
                }
            }
        } else if (!live_list.empty()) {
            const uint64_t i = below(g, live_list.size());
            const uint64_t e = live_list[i];
            live.erase(e);
            live_list[i] = live_list.back();
            live_list.pop_back();
            pool.push_back(e);
            put(out[k++], (uint32_t)(e >> 32), (uint32_t)e, 1);
        }
    }
    return SGC_OK;
}

// static_greedy_coloring (A/src/streamgc.cu:559-592)
int sgc_static_greedy(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint32_t* max_color) {
    if (!row_ptr || !col_idx || !max_color) return SGC_ERR_INVALID;
    std::vector<uint32_t> color(n, 0), mark(n + 2, 0xFFFFFFFFu);
    uint32_t mx = 0;
    for (uint32_t v = 0; v < n; v++) {
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) mark[color[col_idx[j]]] = v;
        uint32_t c = 1;
        while (mark[c] == v) c++;
        color[v] = c;
        mx = c > mx ? c : mx;
    }
    *max_color = mx;
    return SGC_OK;
}

// Life intervals of every edge over a stream: base edges are born at 0, an
// ADD at position i opens [i + 1, .), a DEL at position i closes it at i + 1.
int sgc_edge_intervals(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const sgc_edge_update* stream,
                       uint64_t count, uint64_t* m_out, uint32_t* eu, uint32_t* ev, uint64_t* born, uint64_t* died) {
    if (!row_ptr || !col_idx || (!stream && count) || !m_out) return SGC_ERR_INVALID;
    struct Iv {
        uint64_t key, born, died;
    };
    std::vector<Iv> iv;
    iv.reserve(row_ptr[n] / 2 + count);
    // open interval index per key: a set of keys plus a parallel index map
    std::vector<uint64_t> keys;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++)
            if (u < col_idx[j]) iv.push_back({((uint64_t)u << 32) | col_idx[j], 0, ~0ULL});
    // index: sorted unique keys -> open interval id (updated as intervals open/close)
    // Implemented with a hash map over keys.
    struct Map {
        std::vector<uint64_t> k;
        std::vector<uint64_t> v;
        uint64_t mask;
        explicit Map(uint64_t expect) {
            uint64_t cap = 16;
            while (cap < expect * 2) cap <<= 1;
            k.assign(cap, ~0ULL);
            v.assign(cap, ~0ULL);
            mask = cap - 1;
        }
        uint64_t slot(uint64_t key) const {
            uint64_t h = key * 0x9E3779B97F4A7C15ULL;
            uint64_t i = (h ^ (h >> 29)) & mask;
            while (k[i] != ~0ULL && k[i] != key) i = (i + 1) & mask;
            return i;
        }
    } open(iv.size() + count);
    for (uint64_t i = 0; i < iv.size(); i++) {
        uint64_t s = open.slot(iv[i].key);
        open.k[s] = iv[i].key;
        open.v[s] = i;
    }
    for (uint64_t i = 0; i < count; i++) {
        const sgc_edge_update& r = stream[i];
        if (r.u >= n || r.v >= n || r.u == r.v || r.type > 1) continue;
        const uint64_t key = ekey(r.u, r.v);
        const uint64_t s = open.slot(key);
        const bool is_open = open.k[s] == key && open.v[s] != ~0ULL;
        if (r.type == 0 && !is_open) {
            open.k[s] = key;
            open.v[s] = iv.size();
            iv.push_back({key, i + 1, ~0ULL});
        } else if (r.type == 1 && is_open) {
            iv[open.v[s]].died = i + 1;
            open.v[s] = ~0ULL;
        }
    }
    *m_out = iv.size();
    if (eu && ev && born && died)
        for (uint64_t i = 0; i < iv.size(); i++) {
            eu[i] = (uint32_t)(iv[i].key >> 32);
            ev[i] = (uint32_t)iv[i].key;
            born[i] = iv[i].born;
            died[i] = iv[i].died;
        }
    return SGC_OK;
}

}  // extern "C"

namespace sgc {

// Tier-2 (A/src/streamgc.cu:1057-1240): adjacency from both row directions
// (duplicates kept), uncoloured vertices with an edge coloured first-fit in id
// order, then up to `passes` passes that collect the higher-id endpoint of
// every conflicting edge and recolour them in ascending order with the lowest
// colour in [1, 1024) unused by a neighbour.  Returns vertices recoloured.
uint32_t tier2_host(uint32_t n, const uint32_t* rp, const uint32_t* col, uint32_t* word, uint32_t passes) {
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = rp[u]; j < rp[u + 1]; j++) {
            const uint32_t v = col[j];
            if (v == 0xFFFFFFFFu || v >= n || v == u) continue;
            off[u + 1]++;
            off[v + 1]++;
        }
    for (uint32_t u = 0; u < n; u++) off[u + 1] += off[u];
    std::vector<uint32_t> adj(off[n]);
    {
        std::vector<uint64_t> at(off.begin(), off.end() - 1);
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t j = rp[u]; j < rp[u + 1]; j++) {
                const uint32_t v = col[j];
                if (v == 0xFFFFFFFFu || v >= n || v == u) continue;
                adj[at[u]++] = v;
                adj[at[v]++] = u;
            }
    }
    auto colour = [&](uint32_t v) { return (uint16_t)(word[v] & 0xFFFF); };
    std::vector<uint32_t> seen(1024, 0xFFFFFFFFu);
    uint32_t stamp = 0;
    auto first_fit = [&](uint32_t v) -> bool {
        stamp++;
        for (uint64_t k = off[v]; k < off[v + 1]; k++) {
            const uint16_t c = colour(adj[k]);
            if (c != 0 && c < 1024) seen[c] = stamp;
        }
        for (uint32_t c = 1; c < 1024; c++)
            if (seen[c] != stamp) {
                if (colour(v) == c) return false;
                word[v] = ((((word[v] >> 16) + 1) & 0xFFFF) << 16) | c;
                return true;
            }
        return false;
    };
    uint32_t fixed = 0;
    for (uint32_t u = 0; u < n; u++)
        if (colour(u) == 0 && off[u] != off[u + 1] && first_fit(u)) fixed++;
    std::vector<uint32_t> todo;
    for (uint32_t pass = 0; pass < passes; pass++) {
        todo.clear();
        for (uint32_t u = 0; u < n; u++) {
            const uint16_t cu = colour(u);
            if (cu == 0) continue;
            for (uint64_t k = off[u]; k < off[u + 1]; k++)
                if (adj[k] > u && colour(adj[k]) == cu) todo.push_back(adj[k]);
        }
        if (todo.empty()) break;
        std::sort(todo.begin(), todo.end());
        todo.erase(std::unique(todo.begin(), todo.end()), todo.end());
        uint32_t f = 0;
        for (uint32_t v : todo) f += first_fit(v);
        fixed += f;
        if (f == 0) break;
    }
    return fixed;
}

}  // namespace sgc
