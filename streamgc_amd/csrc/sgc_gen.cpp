// sgc_gen.cpp -- graph containers and synthetic inputs (C++17 + OpenMP).
//
// All generators are deterministic for a given seed whatever the thread
// count: random numbers come from a counter-based stream (splitmix64 of
// seed and draw index), parallel work is split into fixed chunks, and
// anything order-dependent (deduplication to an exact edge count) runs
// sequentially in chunk order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

#include "sgc_internal.hpp"
#include "streamgc.h"

struct sgc_csr {
    uint32_t n = 0;
    std::vector<uint32_t> rp, col, deg;
};

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// the k-th 64-bit draw of stream `s` under `seed`
inline uint64_t draw(uint64_t seed, uint64_t s, uint64_t k) { return mix64(mix64(seed ^ mix64(s)) + k); }
inline double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

template <class It>
void psort(It a, It b) {
#ifdef _OPENMP
    __gnu_parallel::sort(a, b);
#else
    std::sort(a, b);
#endif
}

// CSR from unique undirected keys (lo << 32 | hi, lo < hi): rows sorted.
sgc_csr* csr_from_keys(uint32_t n, std::vector<uint64_t>& keys) {
    psort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    sgc_csr* g = new sgc_csr;
    g->n = n;
    g->deg.assign(n, 0);
    for (uint64_t k : keys) {
        g->deg[k >> 32]++;
        g->deg[(uint32_t)k]++;
    }
    g->rp.assign((size_t)n + 1, 0);
    for (uint32_t v = 0; v < n; v++) g->rp[v + 1] = g->rp[v] + g->deg[v];
    g->col.resize(keys.size() * 2);
    std::vector<uint32_t> at(g->rp.begin(), g->rp.end() - 1);
    // keys ascend by (lo, hi): row lo gets its hi neighbours in order; row hi
    // gets lo neighbours in ascending lo order, but interleaves with its own
    // hi list -- place lo-neighbours first (all < hi) then hi-neighbours.
    // Count lo-neighbours per row first so both groups land sorted.
    std::vector<uint32_t> below(n, 0);
    for (uint64_t k : keys) below[(uint32_t)k]++;
    std::vector<uint32_t> at_lo(at), at_hi(n);
    for (uint32_t v = 0; v < n; v++) at_hi[v] = at[v] + below[v];
    for (uint64_t k : keys) {
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        g->col[at_hi[a]++] = b;   // b > a: goes to a's upper group
        g->col[at_lo[b]++] = a;   // a < b: goes to b's lower group
    }
    return g;
}

}  // namespace

extern "C" {

int sgc_csr_from_edges(uint32_t n, const uint32_t* u, const uint32_t* v, uint64_t m, sgc_csr** out) {
    if (!out || (m && (!u || !v))) return SGC_ERR_INVALID;
    std::vector<uint64_t> keys;
    keys.reserve(m);
    for (uint64_t i = 0; i < m; i++) {
        if (u[i] >= n || v[i] >= n) return SGC_ERR_INVALID;
        if (u[i] == v[i]) continue;
        keys.push_back(u[i] < v[i] ? ((uint64_t)u[i] << 32) | v[i] : ((uint64_t)v[i] << 32) | u[i]);
    }
    *out = csr_from_keys(n, keys);
    return SGC_OK;
}

// load_graph_csr (A/src/host/graph_loader.cpp:29-169): the banner decides
// pattern or weighted; '%' lines skipped; header "rows cols nnz" with
// n = max(rows, cols); nnz entries "u v [rest of line]", 1-indexed;
// self-loops dropped; (min, max) normalised, sorted, deduplicated; symmetric
// CSR with sorted rows.  The file is read whole and parsed in memory.
int sgc_csr_load_mtx(const char* path, sgc_csr** out) {
    if (!path || !out) return SGC_ERR_INVALID;
    FILE* f = fopen(path, "rb");
    if (!f) return SGC_ERR_IO;
    std::string buf;
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, f) : 0;
    fclose(f);
    buf.resize(got);
    const char* p = buf.data();
    const char* end = p + buf.size();
    auto line_end = [&](const char* q) {
        const char* e = (const char*)memchr(q, '\n', (size_t)(end - q));
        return e ? e : end;
    };
    const char* e0 = line_end(p);
    const bool pattern = std::string(p, e0).find("pattern") != std::string::npos;
    p = e0 < end ? e0 + 1 : end;
    while (p < end && *p == '%') {
        const char* e = line_end(p);
        p = e < end ? e + 1 : end;
    }
    const char* eh = line_end(p);
    unsigned rows = 0, cols = 0;
    unsigned long long nnz = 0;
    if (sscanf(std::string(p, eh).c_str(), "%u %u %llu", &rows, &cols, &nnz) != 3) return SGC_ERR_IO;
    p = eh < end ? eh + 1 : end;
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; };
    auto read_u32 = [&](uint32_t& x) {
        while (p < end && ws(*p)) p++;
        if (p >= end) return false;
        bool neg = false;
        if (*p == '+' || *p == '-') neg = *p++ == '-';
        if (p >= end || *p < '0' || *p > '9') return false;
        uint64_t val = 0;
        while (p < end && *p >= '0' && *p <= '9') val = val * 10 + (uint64_t)(*p++ - '0');
        x = (uint32_t)(neg ? 0 - val : val);
        return true;
    };
    const uint32_t n = rows > cols ? rows : cols;
    std::vector<uint64_t> keys;
    keys.reserve(nnz);
    for (unsigned long long i = 0; i < nnz; i++) {
        uint32_t a, b;
        if (!read_u32(a) || !read_u32(b)) break;
        if (!pattern) {
            const char* e = line_end(p);
            p = e < end ? e + 1 : end;
        }
        a--;
        b--;
        if (a == b) continue;
        if (a >= n || b >= n) return SGC_ERR_RANGE;
        keys.push_back(a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a);
    }
    *out = csr_from_keys(n, keys);
    return SGC_OK;
}

int sgc_csr_info(const sgc_csr* g, uint32_t* n, uint64_t* nnz, uint32_t* max_deg) {
    if (!g) return SGC_ERR_INVALID;
    if (n) *n = g->n;
    if (nnz) *nnz = g->col.size();
    if (max_deg) *max_deg = g->deg.empty() ? 0 : *std::max_element(g->deg.begin(), g->deg.end());
    return SGC_OK;
}

int sgc_csr_export(const sgc_csr* g, uint32_t* row_ptr, uint32_t* col_idx, uint32_t* deg) {
    if (!g) return SGC_ERR_INVALID;
    if (row_ptr) memcpy(row_ptr, g->rp.data(), g->rp.size() * 4);
    if (col_idx && !g->col.empty()) memcpy(col_idx, g->col.data(), g->col.size() * 4);
    if (deg && g->n) memcpy(deg, g->deg.data(), (size_t)g->n * 4);
    return SGC_OK;
}

int sgc_csr_free(sgc_csr* g) {
    delete g;
    return SGC_OK;
}

// Chung-Lu graph with a power-law expected-degree sequence
// w_i ∝ (i + i0)^(-1/(gamma-1)); i0 is set so that max/mean weight equals
// max_degree / (2m/n).  Endpoint pairs are drawn independently ∝ w (inverse
// CDF), self-loops skipped, and distinct pairs accepted in draw order until
// exactly m edges exist.  Vertex ids are then randomly permuted.
int sgc_gen_chung_lu(uint32_t n, uint64_t m, double gamma, double max_degree, uint64_t seed, int permute,
                     sgc_csr** out) {
    if (!out || n < 2 || gamma <= 1.0 || m == 0 || m > (uint64_t)n * (n - 1) / 2) return SGC_ERR_INVALID;
    const double alpha = 1.0 / (gamma - 1.0);
    const double want_ratio = max_degree / (2.0 * (double)m / n);
    auto ratio = [&](double i0) {
        double s = 0;
        for (uint32_t i = 0; i < n; i++) s += std::pow((double)i + i0, -alpha);
        return std::pow(i0, -alpha) / (s / n);
    };
    double lo = 1e-3, hi = (double)n;   // ratio falls as i0 grows
    for (int it = 0; it < 60; it++) {
        const double mid = std::sqrt(lo * hi);
        if (ratio(mid) > want_ratio) lo = mid;
        else hi = mid;
    }
    const double i0 = std::sqrt(lo * hi);
    std::vector<double> cdf(n);
    double acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        acc += std::pow((double)i + i0, -alpha);
        cdf[i] = acc;
    }
    for (uint32_t i = 0; i < n; i++) cdf[i] /= acc;
    cdf[n - 1] = 1.0;
    auto pick = [&](double r) { return (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin()); };
    const uint64_t chunk = 1u << 20;
    std::vector<uint64_t> keys;
    keys.reserve(m);
    struct Set {
        std::vector<uint64_t> s;
        uint64_t mask;
        explicit Set(uint64_t expect) {
            uint64_t cap = 16;
            while (cap < expect * 2) cap <<= 1;
            s.assign(cap, ~0ULL);
            mask = cap - 1;
        }
        bool add(uint64_t k) {
            uint64_t i = mix64(k) & mask;
            while (s[i] != ~0ULL) {
                if (s[i] == k) return false;
                i = (i + 1) & mask;
            }
            s[i] = k;
            return true;
        }
    } set(m);
    std::vector<uint64_t> cand(chunk);
    for (uint64_t c = 0; keys.size() < m; c++) {
        if (c > 4096 + 64 * (m / chunk)) return SGC_ERR_RANGE;   // cannot reach m distinct pairs
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)chunk; k++) {
            const uint64_t idx = c * chunk + (uint64_t)k;
            const uint32_t a = pick(unit(draw(seed, 1, 2 * idx)));
            const uint32_t b = pick(unit(draw(seed, 1, 2 * idx + 1)));
            cand[k] = a == b ? ~0ULL : (a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a);
        }
        for (uint64_t k = 0; k < chunk && keys.size() < m; k++)
            if (cand[k] != ~0ULL && set.add(cand[k])) keys.push_back(cand[k]);
    }
    if (permute) {
        std::vector<uint32_t> perm(n);
        for (uint32_t i = 0; i < n; i++) perm[i] = i;
        for (uint32_t i = n - 1; i > 0; i--) {
            const uint64_t j = (uint64_t)(((unsigned __int128)draw(seed, 2, i) * (i + 1)) >> 64);
            std::swap(perm[i], perm[j]);
        }
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)keys.size(); k++) {
            const uint32_t a = perm[keys[k] >> 32], b = perm[(uint32_t)keys[k]];
            keys[k] = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
        }
    }
    *out = csr_from_keys(n, keys);
    return SGC_OK;
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): edge_factor * 2^scale draws,
// each descending `scale` levels into quadrant a / b / c / d; symmetrised,
// self-loops dropped, deduplicated.
int sgc_gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c, uint64_t seed, sgc_csr** out) {
    if (!out || scale == 0 || scale > 31 || edge_factor <= 0 || a < 0 || b < 0 || c < 0 || a + b + c >= 1.0)
        return SGC_ERR_INVALID;
    const uint32_t n = 1u << scale;
    const uint64_t draws = (uint64_t)std::llround(edge_factor * (double)n);
    std::vector<uint64_t> keys(draws);
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)draws; k++) {
        uint32_t s = 0, d = 0;
        for (uint32_t lvl = 0; lvl < scale; lvl++) {
            const double r = unit(draw(seed, 3, (uint64_t)k * scale + lvl));
            const uint32_t bit = 1u << (scale - 1 - lvl);
            if (r < a) {
            } else if (r < ab) {
                d |= bit;
            } else if (r < abc) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
        keys[k] = s == d ? ~0ULL : (s < d ? ((uint64_t)s << 32) | d : ((uint64_t)d << 32) | s);
    }
    keys.erase(std::remove(keys.begin(), keys.end(), ~0ULL), keys.end());
    *out = csr_from_keys(n, keys);
    return SGC_OK;
}

// Core numbers by the Batagelj-Zaversnik bucket algorithm (O(m)).
int sgc_core_numbers(uint32_t n, const uint32_t* rp, const uint32_t* col, uint32_t* core) {
    if (!rp || !col || !core) return SGC_ERR_INVALID;
    uint32_t md = 0;
    std::vector<uint32_t> deg(n);
    for (uint32_t v = 0; v < n; v++) {
        deg[v] = rp[v + 1] - rp[v];
        md = std::max(md, deg[v]);
    }
    std::vector<uint32_t> bin(md + 2, 0), pos(n), vert(n);
    for (uint32_t v = 0; v < n; v++) bin[deg[v]]++;
    uint32_t start = 0;
    for (uint32_t d = 0; d <= md; d++) {
        const uint32_t num = bin[d];
        bin[d] = start;
        start += num;
    }
    for (uint32_t v = 0; v < n; v++) {
        pos[v] = bin[deg[v]]++;
        vert[pos[v]] = v;
    }
    for (uint32_t d = md; d >= 1; d--) bin[d] = bin[d - 1];
    bin[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t v = vert[i];
        for (uint32_t j = rp[v]; j < rp[v + 1]; j++) {
            const uint32_t u = col[j];
            if (deg[u] > deg[v]) {
                const uint32_t du = deg[u], pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                bin[du]++;
                deg[u]--;
            }
        }
    }
    memcpy(core, deg.data(), (size_t)n * 4);
    return SGC_OK;
}

// Random-pair streams over a region H of the vertices: each event inserts
// with probability ins/(ins+1) -- always, for an insert-only stream -- a
// uniform non-adjacent pair in H, redrawn until new, or deletes a uniform
// live edge inside H (swap-remove).  mt19937-free: counter-based draws of
// (seed, stream id, event).
static int region_stream(uint32_t n, const uint32_t* rp, const uint32_t* col, const std::vector<uint8_t>& inH,
                         uint64_t size, uint32_t ins_per_del, bool inserts_only, uint64_t seed, uint64_t sid,
                         sgc_edge_update* out) {
    std::vector<uint32_t> H;
    for (uint32_t v = 0; v < n; v++)
        if (inH[v]) H.push_back(v);
    if (H.size() < 2) return SGC_ERR_RANGE;
    std::vector<uint64_t> live_list;
    // all live edges (membership), and the deletable list inside H
    std::vector<uint64_t> all_keys;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = rp[u]; j < rp[u + 1]; j++)
            if (u < col[j]) {
                const uint64_t k = ((uint64_t)u << 32) | col[j];
                if (!inserts_only && inH[u] && inH[col[j]]) live_list.push_back(k);
                all_keys.push_back(k);
            }
    // open-addressing set with backward-shift delete
    std::vector<uint64_t> slot;
    uint64_t cap = 16;
    while (cap < (all_keys.size() + size) * 2) cap <<= 1;
    slot.assign(cap, ~0ULL);
    const uint64_t mask = cap - 1;
    auto find = [&](uint64_t k) {
        uint64_t i = mix64(k) & mask;
        while (slot[i] != ~0ULL && slot[i] != k) i = (i + 1) & mask;
        return i;
    };
    for (uint64_t k : all_keys) slot[find(k)] = k;
    std::vector<uint64_t>().swap(all_keys);
    auto erase = [&](uint64_t k) {
        uint64_t hole = find(k);
        if (slot[hole] != k) return;
        for (uint64_t j = (hole + 1) & mask; slot[j] != ~0ULL; j = (j + 1) & mask) {
            const uint64_t home = mix64(slot[j]) & mask;
            const bool stays = hole <= j ? (hole < home && home <= j) : (hole < home || home <= j);
            if (!stays) {
                slot[hole] = slot[j];
                hole = j;
            }
        }
        slot[hole] = ~0ULL;
    };
    uint64_t draw_k = 0;
    auto next = [&]() { return draw(seed, sid, draw_k++); };
    auto below = [&](uint64_t r) { return (uint64_t)(((unsigned __int128)next() * r) >> 64); };
    const uint64_t hs = H.size();
    for (uint64_t i = 0; i < size; i++) {
        const bool ins = inserts_only || live_list.empty() || below(ins_per_del + 1) < ins_per_del;
        std::memset(&out[i], 0, sizeof out[i]);
        if (ins) {
            for (int tries = 0;; tries++) {
                const uint32_t a = H[below(hs)], b = H[below(hs)];
                if (a == b) continue;
                const uint64_t k = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
                const uint64_t s = find(k);
                if (slot[s] == k) {
                    if (tries > 1000000) return SGC_ERR_RANGE;   // the region is complete
                    continue;
                }
                slot[s] = k;
                if (!inserts_only) live_list.push_back(k);
                out[i].u = (uint32_t)(k >> 32);
                out[i].v = (uint32_t)k;
                out[i].type = 0;
                break;
            }
        } else {
            const uint64_t j = below(live_list.size());
            const uint64_t k = live_list[j];
            live_list[j] = live_list.back();
            live_list.pop_back();
            erase(k);
            out[i].u = (uint32_t)(k >> 32);
            out[i].v = (uint32_t)k;
            out[i].type = 1;
        }
    }
    return SGC_OK;
}

// Dense-growth stream over the dense region H = {v : core(v) >= core_min}.
int sgc_dense_growth_stream(uint32_t n, const uint32_t* rp, const uint32_t* col, const uint32_t* core,
                            uint32_t core_min, uint64_t size, uint32_t ins_per_del, uint64_t seed,
                            sgc_edge_update* out) {
    if (!rp || !col || !core || (!out && size)) return SGC_ERR_INVALID;
    std::vector<uint8_t> inH(n, 0);
    for (uint32_t v = 0; v < n; v++) inH[v] = core[v] >= core_min;
    return region_stream(n, rp, col, inH, size, ins_per_del, false, seed, 4, out);
}

// Uniform stream (README.md "Uniform"): inserts of uniform non-adjacent
// vertex pairs, each new.
int sgc_uniform_stream(uint32_t n, const uint32_t* rp, const uint32_t* col, uint64_t size, uint64_t seed,
                       sgc_edge_update* out) {
    if (!rp || !col || (!out && size)) return SGC_ERR_INVALID;
    return region_stream(n, rp, col, std::vector<uint8_t>(n, 1), size, 0, true, seed, 6, out);
}

// Hold-out stream (README.md "Hold-out"): round(m * pct / 100) of the m
// edges, drawn uniformly (a partial Fisher-Yates shuffle of the edge list),
// leave the base graph and come back as inserts in the drawn order.
int sgc_holdout_stream(uint32_t n, const uint32_t* rp, const uint32_t* col, uint32_t pct, uint64_t seed,
                       sgc_csr** base_out, sgc_edge_update* out, uint64_t* count) {
    if (!rp || !col || !count || pct > 100) return SGC_ERR_INVALID;
    std::vector<uint64_t> keys;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = rp[u]; j < rp[u + 1]; j++)
            if (u < col[j] && col[j] < n) keys.push_back(((uint64_t)u << 32) | col[j]);
    const uint64_t m = keys.size(), h = (m * pct + 50) / 100;
    *count = h;
    if (!out && !base_out) return SGC_OK;
    for (uint64_t i = 0; i < h; i++) {
        const uint64_t j = i + (uint64_t)(((unsigned __int128)draw(seed, 7, i) * (m - i)) >> 64);
        std::swap(keys[i], keys[j]);
    }
    if (out)
        for (uint64_t i = 0; i < h; i++) {
            std::memset(&out[i], 0, sizeof out[i]);
            out[i].u = (uint32_t)(keys[i] >> 32);
            out[i].v = (uint32_t)keys[i];
            out[i].type = 0;
        }
    if (base_out) {
        std::vector<uint64_t> kept(keys.begin() + h, keys.end());
        *base_out = csr_from_keys(n, kept);
    }
    return SGC_OK;
}

}  // extern "C"
