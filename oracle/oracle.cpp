// oracle.cpp -- CPU restatement of the reference StreamGC hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
// entry point and the bench's CPU-baseline leg load it.  The product (the HIP
// engine under streamgc_amd/) never calls, links or falls back to it.
//
// Every function states the reference file:line it restates.  Paths are
// relative to the reference checkout, with A/ = algorithms/streamgc/.
//
// Semantics: the reference runs one warp per update with all real work on
// lane 0 (A/src/kernels/local_update.cu:81-123).  This oracle applies the
// updates one at a time in stream order, which is exactly the behaviour of a
// single warp processing the batch; the repair sweep (Tier-1) runs in vertex
// order on one thread (Gauss-Seidel), Tier-2 follows the host code verbatim.
// All arithmetic is integer; u16 versions wrap like the reference's uint16_t.
//
// Pinning (what checks each part, so a plausible slip fails a test):
//   mt19937_64 ............ the C++ standard's 10000th-output value; golden
//                           streams made by the reference's own stream_gen.cpp
//   uniform (Lemire) ...... brute-force distribution identities on tiny ranges;
//                           golden streams
//   init / bench stream ... golden fixtures from the reference host code
//   mtx loader ............ golden CSR hashes from the reference loader
//   find_available_color .. the reference unit-test vectors
//   engine ................ closed forms (K4, C100, star, triangle-delete),
//                           invariants (proper after repair, version +1 per
//                           recolor, only the lower-priority endpoint recolors)
//                           and brute-force chromatic bounds on tiny graphs
//   device quirks ......... hand-computed states, one per behaviour: colour
//                           reduction on delete and which freed colour each
//                           endpoint gets, propagation of a recolour to the
//                           neighbours' bitmaps and histograms, stable
//                           compaction, Tier-1's own-colour bit, the plateau
//                           stop, drain in buffer order, hub bitmaps left to
//                           the histogram.  tests/test_oracle_mutants.py plants
//                           a slip in each of these (and others) in a scratch
//                           copy of this file and requires a test to fail.
//   tier-2 / validator .... brute force on tiny graphs; static greedy against a
//                           textbook first-fit written independently in Python
//
// Build: oracle/Makefile (g++ -O2 -fopenmp).

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t INVALID = 0xFFFFFFFFu;   // A/include/types.h:24-25
constexpr uint32_t HIST_BINS = 1024;        // A/include/partition.cuh:6-9

// ---------------------------------------------------------------------------
// mt19937_64 as specified by [rand.eng.mt] / [rand.predef]; the reference uses
// std::mt19937_64 (A/src/host/stream_gen.cpp:73).
struct Mt64 {
    uint64_t mt[312];
    int idx;
    explicit Mt64(uint64_t seed) {
        mt[0] = seed;
        for (int i = 1; i < 312; i++)
            mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
        idx = 312;
    }
    void twist() {
        const uint64_t upper = ~0ULL << 31, lower = (1ULL << 31) - 1;
        for (int i = 0; i < 312; i++) {
            uint64_t x = (mt[i] & upper) | (mt[(i + 1) % 312] & lower);
            uint64_t xa = x >> 1;
            if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
            mt[i] = mt[(i + 156) % 312] ^ xa;
        }
        idx = 0;
    }
    uint64_t next() {
        if (idx >= 312) twist();
        uint64_t y = mt[idx++];
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= y >> 43;
        return y;
    }
};

// Uniform integer in [a, b] drawn the way libstdc++ 11 (GCC 11.4) does for a
// 64-bit engine: Lemire's nearly-divisionless method with a 128-bit product
// (Lemire 2019, "Fast Random Integer Generation in an Interval").  Call sites:
// A/src/host/stream_gen.cpp:105,114,129,155.
uint64_t uniform(Mt64& g, uint64_t a, uint64_t b) {
    const uint64_t range = b - a + 1;  // b - a < 2^64 - 1 at every call site
    unsigned __int128 prod = (unsigned __int128)g.next() * range;
    uint64_t low = (uint64_t)prod;
    if (low < range) {
        const uint64_t threshold = (0 - range) % range;
        while (low < threshold) {
            prod = (unsigned __int128)g.next() * range;
            low = (uint64_t)prod;
        }
    }
    return (uint64_t)(prod >> 64) + a;
}

inline uint16_t color_of(uint32_t w) { return (uint16_t)(w & 0xFFFF); }    // A/include/vertex_state.cuh:46-48
inline uint16_t version_of(uint32_t w) { return (uint16_t)(w >> 16); }     // A/include/vertex_state.cuh:50-52
inline uint32_t pack(uint16_t c, uint16_t ver) { return ((uint32_t)ver << 16) | c; }  // :42-44

struct Config {              // same field order as sgc_config in include/streamgc.h
    uint32_t epoch_size;
    uint32_t batch_size;
    uint32_t histogram_threshold;
    uint32_t slack_pct;
    uint32_t min_row_capacity;
    uint32_t overflow_capacity;
    uint32_t max_repair_iters;
    uint32_t plateau_iters;
    uint32_t tier2_passes;
    uint32_t flags;
};
constexpr uint32_t FLAG_REPAIR_EVERY_EPOCH = 1u << 1;
constexpr uint32_t FLAG_EXACT_HISTOGRAM = 1u << 2;
// The reference's compile-time ablations (Makefile:58-78) as flags.
constexpr uint32_t FLAG_ABL_FORCE_BITMAP = 1u << 5;     // A/include/color_select.cuh:98-100
constexpr uint32_t FLAG_ABL_NO_OVERFLOW = 1u << 7;      // A/src/kernels/local_update.cu:28-38
constexpr uint32_t FLAG_ABL_LIVE_DEGREE = 1u << 8;      // A/src/kernels/conflict_repair.cu:46-58

struct Stats {               // same layout as sgc_epoch_stats in include/streamgc.h
    uint32_t epoch_index;
    uint32_t overflow_count;
    uint32_t overflow_drained;
    uint32_t overflow_failed;
    uint32_t t1_iterations;
    uint32_t t1_conflicts_last;
    uint32_t t2_fixed;
    uint32_t t2_remaining;
    uint32_t max_color;
    uint32_t snapshot_published;
    uint32_t t2_ran;
    uint32_t reserved;
};

struct Engine {
    Config cfg;
    uint32_t n = 0;
    std::vector<uint32_t> deg0;      // loaded-graph degree: fixes the priority
    std::vector<uint32_t> row_ptr;   // n + 1, slack capacities
    std::vector<uint32_t> col;       // S slots, INVALID = free
    std::vector<uint32_t> word;      // packed (version << 16) | color
    std::vector<uint64_t> bitmap;
    std::vector<uint32_t> hd_map;    // histogram row or INVALID
    std::vector<uint32_t> hist;      // num_hd * 1024 counters
    std::vector<std::pair<uint32_t, uint32_t>> overflow;
    uint32_t overflow_count = 0;
    std::vector<uint16_t> snapshot;
    uint32_t epochs = 0;
    Stats last{};
    int threads = 1;

    uint64_t pi(uint32_t v) const { return ((uint64_t)deg0[v] << 32) | v; }  // A/include/priority.cuh:8-10
    uint16_t color(uint32_t v) const { return color_of(word[v]); }

    // A/include/color_select.cuh:10-53 (the ghost scan is empty on one GPU)
    uint16_t fas_bitmap(uint32_t v) const {
        uint64_t free_bits = ~bitmap[v];
        if (free_bits) return (uint16_t)(__builtin_ctzll(free_bits) + 1);
        bool used[256] = {false};
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
            uint32_t nb = col[j];
            if (nb == INVALID || nb >= n) continue;
            uint16_t c = color(nb);
            if (c > 0 && c < 256) used[c] = true;
        }
        for (uint16_t c = 1; c < 256; c++)
            if (!used[c]) return c;
        return 256;
    }
    // A/include/color_select.cuh:59-89
    uint16_t fas_hist(uint32_t v) const {
        const uint32_t* h = &hist[(size_t)hd_map[v] * HIST_BINS];
        for (uint32_t c = 0; c < HIST_BINS; c++)
            if (h[c] == 0) return (uint16_t)(c + 1);
        uint32_t mc = h[0], mi = 0;
        for (uint32_t c = 1; c < HIST_BINS; c++)
            if (h[c] < mc) { mc = h[c]; mi = c; }
        return (uint16_t)(mi + 1);
    }
    // A/include/color_select.cuh:93-115
    uint16_t fas(uint32_t v) const {
        if (cfg.flags & FLAG_ABL_FORCE_BITMAP) return fas_bitmap(v);
        return hd_map[v] != INVALID ? fas_hist(v) : fas_bitmap(v);
    }

    // A/include/color_select.cuh:160-183
    void on_new_neighbor(uint32_t v, uint16_t c) {
        if (c == 0) return;
        if (c <= 64) bitmap[v] |= 1ULL << (c - 1);
        if (hd_map[v] != INVALID && c <= HIST_BINS) hist[(size_t)hd_map[v] * HIST_BINS + c - 1] += 1;
    }
    // A/include/color_select.cuh:188-209 (bitmap bits are left set)
    void on_removed_neighbor(uint32_t v, uint16_t c) {
        if (c == 0) return;
        if (hd_map[v] != INVALID && c <= HIST_BINS) hist[(size_t)hd_map[v] * HIST_BINS + c - 1] -= 1;
    }
    // A/include/color_select.cuh:120-157
    void on_neighbor_recolor(uint32_t v, uint16_t oldc, uint16_t newc) {
        if (oldc == newc) return;
        if (newc != 0 && newc <= 64) bitmap[v] |= 1ULL << (newc - 1);
        if (hd_map[v] != INVALID) {
            uint32_t* h = &hist[(size_t)hd_map[v] * HIST_BINS];
            if (oldc != 0 && oldc <= HIST_BINS) h[oldc - 1] -= 1;
            if (newc != 0 && newc <= HIST_BINS) h[newc - 1] += 1;
        }
    }
    // A/src/kernels/local_update.cu:12-39: first free slot or an existing copy;
    // a full row goes to the overflow buffer, which drops entries past capacity.
    void add_neighbor(uint32_t v, uint32_t nb) {
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
            if (col[j] == INVALID) { col[j] = nb; return; }
            if (col[j] == nb) return;
        }
        if (cfg.flags & FLAG_ABL_NO_OVERFLOW) return;   // the ablation drops the edge
        uint32_t pos = overflow_count++;
        if (pos < cfg.overflow_capacity) overflow[pos] = {v, nb};
    }
    // A/src/kernels/local_update.cu:42-56: first match becomes INVALID
    void remove_neighbor(uint32_t v, uint32_t nb) {
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
            if (col[j] == nb) { col[j] = INVALID; return; }
    }
    // A/src/kernels/recolor.cu:50-82
    void propagate(uint32_t v, uint16_t oldc, uint16_t newc) {
        if (oldc == newc) return;
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
            uint32_t nb = col[j];
            if (nb == INVALID || nb >= n) continue;
            on_neighbor_recolor(nb, oldc, newc);
        }
    }
    // A/src/kernels/recolor.cu:87-141 (uncontended: the optimistic store wins)
    void recolor(uint32_t v) {
        uint16_t oldc = color(v);
        uint16_t newc = fas(v);
        if (newc == oldc) return;
        word[v] = pack(newc, (uint16_t)(version_of(word[v]) + 1));
        propagate(v, oldc, newc);
    }
    // A/src/kernels/recolor.cu:145-163
    void try_reduce(uint32_t v, uint16_t freed) {
        uint16_t cur = color(v);
        if (cur == 0) return;
        if (freed >= cur) return;
        if (fas(v) < cur) recolor(v);
    }
    // A/src/kernels/local_update.cu:71-146, one update
    void apply(uint32_t u, uint32_t v, uint8_t type) {
        if (u >= n || v >= n) return;                      // hash miss -> skipped (:77)
        if (type == 0) {                                   // ADD
            add_neighbor(u, v);
            add_neighbor(v, u);
            if (color(u) == 0) word[u] = pack(fas(u), 1);
            if (color(v) == 0) {
                on_new_neighbor(v, color(u));
                word[v] = pack(fas(v), 1);
            }
            on_new_neighbor(u, color(v));
            on_new_neighbor(v, color(u));
            if (color(u) == color(v) && color(u) != 0) recolor(pi(u) < pi(v) ? u : v);
        } else if (type == 1) {                            // DEL
            uint16_t fu = color(v), fv = color(u);
            remove_neighbor(u, v);
            remove_neighbor(v, u);
            on_removed_neighbor(u, fu);
            on_removed_neighbor(v, fv);
            try_reduce(u, fu);
            try_reduce(v, fv);
        }
    }

    // A/src/kernels/epoch_maintenance.cu:5-16
    void take_snapshot() {
        #pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t v = 0; v < (int64_t)n; v++) snapshot[v] = color((uint32_t)v);
    }
    // A/src/kernels/epoch_maintenance.cu:20-61 (stable left-pack per row)
    void compact() {
        #pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
        for (int64_t vi = 0; vi < (int64_t)n; vi++) {
            uint32_t v = (uint32_t)vi, w = row_ptr[v];
            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
                if (col[j] != INVALID) col[w++] = col[j];
            for (; w < row_ptr[v + 1]; w++) col[w] = INVALID;
        }
    }
    // A/src/streamgc.cu:972-983 + A/src/kernels/epoch_maintenance.cu:66-87,
    // entries retried in buffer order.
    void drain(Stats& st) {
        uint32_t count = overflow_count;
        st.overflow_count = count;
        if (count == 0) return;
        overflow_count = 0;
        uint32_t m = std::min(count, cfg.overflow_capacity);
        st.overflow_drained = m;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t v = overflow[i].first, nb = overflow[i].second;
            bool done = false;
            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1] && !done; j++) {
                if (col[j] == INVALID) { col[j] = nb; done = true; }
                else if (col[j] == nb) done = true;
            }
            if (!done) st.overflow_failed++;
        }
    }
    // A/include/color_select.cuh:213-251 + A/src/kernels/epoch_maintenance.cu:113-128.
    // With FLAG_EXACT_HISTOGRAM (no reference counterpart: the reference never
    // rebuilds histograms, A/src/kernels/epoch_maintenance.cu:120-123) the
    // histogram rows are recounted from the live rows at the same point.
    void rebuild_bitmaps() {
        const bool exact = (cfg.flags & FLAG_EXACT_HISTOGRAM) != 0;
        #pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
        for (int64_t vi = 0; vi < (int64_t)n; vi++) {
            uint32_t v = (uint32_t)vi;
            if (hd_map[v] != INVALID) {
                if (!exact) continue;
                uint32_t* h = &hist[(size_t)hd_map[v] * HIST_BINS];
                std::fill(h, h + HIST_BINS, 0u);
                for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
                    uint32_t nb = col[j];
                    if (nb == INVALID || nb >= n) continue;
                    uint16_t c = color(nb);
                    if (c != 0 && c <= HIST_BINS) h[c - 1]++;
                }
                continue;
            }
            uint64_t bm = 0;
            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
                uint32_t nb = col[j];
                if (nb == INVALID || nb >= n) continue;
                uint16_t c = color(nb);
                if (c != 0 && c <= 64) bm |= 1ULL << (c - 1);
            }
            bitmap[v] = bm;
        }
    }
    // A/src/kernels/conflict_repair.cu:10-114, one thread in vertex order
    uint32_t tier1_sweep() {
        uint32_t conflicts = 0;
        for (uint32_t v = 0; v < n; v++) {
            uint16_t mine = color(v);
            bool has_nb = false;
            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
                if (col[j] != INVALID) { has_nb = true; break; }
            if (mine == 0) {
                if (has_nb) {
                    word[v] = pack(fas(v), (uint16_t)(version_of(word[v]) + 1));
                    conflicts++;
                }
                continue;
            }
            uint64_t my_pi = pi(v);
            if (cfg.flags & FLAG_ABL_LIVE_DEGREE) {   // own priority from the live row, neighbours' from deg0
                uint64_t live = 0;
                for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) live += col[j] != INVALID;
                my_pi = (live << 32) | v;
            }
            bool need = false;
            for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
                uint32_t nb = col[j];
                if (nb == INVALID || nb >= n) continue;
                if (color(nb) == mine && my_pi < pi(nb)) { need = true; break; }
            }
            if (need) {
                conflicts++;
                if (mine <= 64) bitmap[v] |= 1ULL << (mine - 1);
                word[v] = pack(fas(v), (uint16_t)(version_of(word[v]) + 1));
            }
        }
        return conflicts;
    }
    // A/src/streamgc.cu:1317-1393
    void tier1(Stats& st) {
        uint32_t prev = 0xFFFFFFFFu;
        uint32_t plateau = 0;
        st.t1_iterations = 0;
        for (uint32_t it = 0; it < cfg.max_repair_iters; it++) {
            rebuild_bitmaps();
            uint32_t c = tier1_sweep();
            st.t1_iterations = it + 1;
            st.t1_conflicts_last = c;
            if (c == 0) return;
            if (c == prev) {
                if (++plateau >= cfg.plateau_iters) return;
            } else {
                plateau = 0;
            }
            prev = c;
        }
    }
    // A/src/streamgc.cu:1057-1240 (Tier-2), adjacency from both row directions
    void tier2(Stats& st) {
        // Tier-2 runs unconditionally here, as in the reference; t2_ran
        // reports whether it had anything to do (an uncoloured endpoint or a
        // conflict on some row slot).
        st.t2_ran = row_conflicts() > 0 ? 1 : 0;
        std::vector<uint64_t> off(n + 1, 0);
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++) {
                uint32_t v = col[j];
                if (v == INVALID || v >= n || v == u) continue;
                off[u + 1]++;
                off[v + 1]++;
            }
        for (uint32_t u = 0; u < n; u++) off[u + 1] += off[u];
        std::vector<uint32_t> adj(off[n]);
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++) {
                uint32_t v = col[j];
                if (v == INVALID || v >= n || v == u) continue;
                adj[pos[u]++] = v;
                adj[pos[v]++] = u;
            }
        auto greedy = [&](uint32_t v) -> bool {   // :1169-1187
            bool used[1024] = {false};
            for (uint64_t k = off[v]; k < off[v + 1]; k++) {
                uint16_t c = color(adj[k]);
                if (c != 0 && c < 1024) used[c] = true;
            }
            for (uint16_t c = 1; c < 1024; c++) {
                if (!used[c]) {
                    if (color(v) != c) {
                        word[v] = pack(c, (uint16_t)(version_of(word[v]) + 1));
                        return true;
                    }
                    return false;
                }
            }
            return false;
        };
        uint32_t fixed = 0;
        for (uint32_t u = 0; u < n; u++) {                 // :1190-1194
            if (color(u) != 0 || off[u] == off[u + 1]) continue;
            if (greedy(u)) fixed++;
        }
        for (uint32_t pass = 0; pass < cfg.tier2_passes; pass++) {   // :1197-1218
            std::vector<uint32_t> to_fix;
            for (uint32_t u = 0; u < n; u++)
                for (uint64_t k = off[u]; k < off[u + 1]; k++) {
                    uint32_t v = adj[k];
                    if (v <= u) continue;
                    if (color(u) == color(v) && color(u) != 0) to_fix.push_back(v);
                }
            if (to_fix.empty()) break;
            std::sort(to_fix.begin(), to_fix.end());
            to_fix.erase(std::unique(to_fix.begin(), to_fix.end()), to_fix.end());
            uint32_t f = 0;
            for (uint32_t v : to_fix)
                if (greedy(v)) f++;
            fixed += f;
            if (f == 0) break;
        }
        st.t2_fixed = fixed;
    }
    // Conflicting row edges (u, v), counted once per stored slot.
    uint32_t row_conflicts() const {
        uint64_t bad = 0;
        #pragma omp parallel for num_threads(threads) reduction(+ : bad) schedule(dynamic, 1024)
        for (int64_t ui = 0; ui < (int64_t)n; ui++) {
            uint32_t u = (uint32_t)ui;
            for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++) {
                uint32_t v = col[j];
                if (v == INVALID || v >= n || v == u) continue;
                if (color(u) == 0 || color(v) == 0 || color(u) == color(v)) bad++;
            }
        }
        return (uint32_t)std::min<uint64_t>(bad, 0xFFFFFFFFu);
    }
    uint32_t max_color() const {
        uint32_t m = 0;
        for (uint32_t v = 0; v < n; v++) m = std::max<uint32_t>(m, color(v));
        return m;
    }

    // One epoch step.  Reference order (A/src/streamgc.cu:949-1036): snapshot,
    // compact, drain, rebuild.  With REPAIR_EVERY_EPOCH the method order of
    // README.md:170-176: compact, drain, Tier-1 (its first step is the
    // rebuild), Tier-2, then publish the snapshot.
    void epoch(Stats& st) {
        st = Stats{};
        st.epoch_index = epochs++;
        if (cfg.flags & FLAG_REPAIR_EVERY_EPOCH) {
            compact();
            drain(st);
            tier1(st);
            tier2(st);
            st.t2_remaining = row_conflicts();
            take_snapshot();
            st.snapshot_published = 1;
        } else {
            take_snapshot();
            st.snapshot_published = 1;
            compact();
            drain(st);
            rebuild_bitmaps();
        }
        st.max_color = max_color();
        last = st;
    }
    // End of a phase: A/src/streamgc.cu:1243-1314, 1317-1393, 1057-1240.
    void finish(Stats& st) {
        st = Stats{};
        st.epoch_index = epochs;
        compact();
        drain(st);
        rebuild_bitmaps();
        tier1(st);
        tier2(st);
        st.t2_remaining = row_conflicts();
        take_snapshot();
        st.snapshot_published = 1;
        st.max_color = max_color();
        last = st;
    }
    // A/src/streamgc.cu:809-1053, single-GPU branch: batches of batch_size,
    // the epoch check after every batch.
    uint32_t submit(const uint32_t* u, const uint32_t* v, const uint8_t* t, uint64_t count, Stats* per_epoch,
                    uint32_t max_epochs) {
        uint64_t done = 0, since = 0;
        uint32_t ne = 0;
        while (done < count) {
            uint64_t b = std::min<uint64_t>(cfg.batch_size, count - done);
            for (uint64_t i = 0; i < b; i++) apply(u[done + i], v[done + i], t ? t[done + i] : 0);
            done += b;
            since += b;
            if (since >= cfg.epoch_size) {
                Stats st;
                epoch(st);
                if (per_epoch && ne < max_epochs) per_epoch[ne] = st;
                ne++;
                since = 0;
            }
        }
        return ne;
    }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// RNG access for tests.
uint64_t orc_mt64_nth(uint64_t seed, uint64_t k) {
    Mt64 g(seed);
    uint64_t x = 0;
    for (uint64_t i = 0; i < k; i++) x = g.next();
    return x;
}
void orc_uniform_draws(uint64_t seed, uint64_t a, uint64_t b, uint64_t count, uint64_t* out) {
    Mt64 g(seed);
    for (uint64_t i = 0; i < count; i++) out[i] = uniform(g, a, b);
}

// ---------------------------------------------------------------------------
// A/src/host/stream_gen.cpp:30-61.  out arrays hold row_ptr[n]/2 entries.
uint64_t orc_init_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* deg,
                         uint32_t* out_u, uint32_t* out_v) {
    std::vector<std::pair<uint32_t, uint32_t>> e;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++)
            if (u < col[j]) e.push_back({u, col[j]});
    std::sort(e.begin(), e.end(), [deg](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
        uint32_t la = std::min(deg[a.first], deg[a.second]), lb = std::min(deg[b.first], deg[b.second]);
        if (la != lb) return la > lb;
        uint32_t ha = std::max(deg[a.first], deg[a.second]), hb = std::max(deg[b.first], deg[b.second]);
        return ha > hb;
    });
    for (size_t i = 0; i < e.size(); i++) { out_u[i] = e[i].first; out_v[i] = e[i].second; }
    return e.size();
}

// A/src/host/stream_gen.cpp:65-177 (churn stream); returns live edges at the end.
uint64_t orc_bench_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, uint64_t size, uint32_t pct,
                          uint64_t seed, uint32_t* out_u, uint32_t* out_v, uint8_t* out_t) {
    Mt64 g(seed);
    std::unordered_set<uint64_t> live;
    std::vector<std::pair<uint32_t, uint32_t>> live_vec, pool;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = row_ptr[u]; j < row_ptr[u + 1]; j++)
            if (u < col[j]) {
                live.insert(((uint64_t)u << 32) | col[j]);
                live_vec.push_back({u, col[j]});
            }
    uint64_t k = 0;
    for (uint64_t i = 0; i < size; i++) {
        bool ins = uniform(g, 0, 99) < pct;
        if (ins) {
            if (!pool.empty()) {
                size_t idx = (size_t)uniform(g, 0, pool.size() - 1);
                auto e = pool[idx];
                pool[idx] = pool.back();
                pool.pop_back();
                live.insert(((uint64_t)e.first << 32) | e.second);
                live_vec.push_back(e);
                out_u[k] = e.first; out_v[k] = e.second; out_t[k] = 0; k++;
            } else {
                uint32_t a = (uint32_t)uniform(g, 0, n - 1);
                uint32_t b = (uint32_t)uniform(g, 0, n - 1);
                if (a == b) b = (b + 1) % n;
                uint32_t lo = std::min(a, b), hi = std::max(a, b);
                uint64_t key = ((uint64_t)lo << 32) | hi;
                if (live.find(key) == live.end()) {
                    live.insert(key);
                    live_vec.push_back({lo, hi});
                    out_u[k] = lo; out_v[k] = hi; out_t[k] = 0; k++;
                } else {
                    i--;
                    continue;
                }
            }
        } else {
            if (live_vec.empty()) { i--; continue; }
            size_t idx = (size_t)uniform(g, 0, live_vec.size() - 1);
            auto e = live_vec[idx];
            live.erase(((uint64_t)e.first << 32) | e.second);
            live_vec[idx] = live_vec.back();
            live_vec.pop_back();
            pool.push_back(e);
            out_u[k] = e.first; out_v[k] = e.second; out_t[k] = 1; k++;
        }
    }
    return live.size();
}

// ---------------------------------------------------------------------------
// A/src/host/graph_loader.cpp:29-169.  Two calls: the first (out arrays null)
// reports n and nnz; the second fills row_ptr[n+1], col[nnz], deg[n].
struct OrcCsr { uint32_t n; std::vector<uint32_t> rp, col, deg; };
void* orc_load_mtx(const char* path, uint32_t* n_out, uint64_t* nnz_out) {
    FILE* f = fopen(path, "r");
    if (!f) return nullptr;
    char line[4096];
    bool pattern = false;
    if (fgets(line, sizeof line, f) && strstr(line, "pattern")) pattern = true;
    while (fgets(line, sizeof line, f))
        if (line[0] != '%') break;
    uint32_t rows = 0, cols = 0;
    unsigned long long nnz = 0;
    if (sscanf(line, "%u %u %llu", &rows, &cols, &nnz) != 3) { fclose(f); return nullptr; }
    uint32_t n = rows > cols ? rows : cols;
    std::vector<std::pair<uint32_t, uint32_t>> e;
    e.reserve(nnz);
    for (unsigned long long i = 0; i < nnz; i++) {
        uint32_t u = 0, v = 0;
        if (fscanf(f, "%u %u", &u, &v) != 2) break;
        if (!pattern) {
            int c;
            while ((c = fgetc(f)) != '\n' && c != EOF) {}
        }
        u--; v--;
        if (u == v) continue;
        if (u > v) std::swap(u, v);
        e.push_back({u, v});
    }
    fclose(f);
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    OrcCsr* g = new OrcCsr;
    g->n = n;
    g->deg.assign(n, 0);
    for (auto& p : e) { g->deg[p.first]++; g->deg[p.second]++; }
    g->rp.assign(n + 1, 0);
    for (uint32_t v = 0; v < n; v++) g->rp[v + 1] = g->rp[v] + g->deg[v];
    g->col.assign(e.size() * 2, 0);
    std::vector<uint32_t> o(n, 0);
    for (auto& p : e) {
        g->col[g->rp[p.first] + o[p.first]++] = p.second;
        g->col[g->rp[p.second] + o[p.second]++] = p.first;
    }
    for (uint32_t v = 0; v < n; v++) std::sort(g->col.begin() + g->rp[v], g->col.begin() + g->rp[v + 1]);
    *n_out = n;
    *nnz_out = g->col.size();
    return g;
}
void orc_csr_get(void* h, uint32_t* rp, uint32_t* col, uint32_t* deg) {
    OrcCsr* g = (OrcCsr*)h;
    memcpy(rp, g->rp.data(), g->rp.size() * 4);
    if (!g->col.empty()) memcpy(col, g->col.data(), g->col.size() * 4);
    memcpy(deg, g->deg.data(), g->deg.size() * 4);
}
void orc_csr_free(void* h) { delete (OrcCsr*)h; }

// ---------------------------------------------------------------------------
// Engine.  Construction follows A/src/streamgc.cu:144-444 (1-GPU branch):
// row capacity max(d + d*slack/100, min_cap), rows prefilled with the loaded
// (sorted) neighbours, histogram rows for deg0 > threshold in vertex order.
void* orc_engine_create(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* deg0,
                        const Config* cfg) {
    Engine* e = new Engine;
    e->cfg = *cfg;
    e->n = n;
    e->deg0.assign(deg0, deg0 + n);
    e->row_ptr.assign(n + 1, 0);
    for (uint32_t v = 0; v < n; v++) {
        uint64_t d = row_ptr[v + 1] - row_ptr[v];
        uint64_t cap = std::max<uint64_t>(d + d * cfg->slack_pct / 100, cfg->min_row_capacity);
        uint64_t end = e->row_ptr[v] + cap;
        if (end > 0xFFFFFFFFull) { delete e; return nullptr; }
        e->row_ptr[v + 1] = (uint32_t)end;
    }
    e->col.assign(e->row_ptr[n], INVALID);
    for (uint32_t v = 0; v < n; v++)
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) e->col[e->row_ptr[v] + (j - row_ptr[v])] = col[j];
    e->word.assign(n, 0);
    e->bitmap.assign(n, 0);
    e->hd_map.assign(n, INVALID);
    uint32_t nhd = 0;
    for (uint32_t v = 0; v < n; v++)
        if (deg0[v] > cfg->histogram_threshold) e->hd_map[v] = nhd++;
    e->hist.assign((size_t)nhd * HIST_BINS, 0);
    e->overflow.assign(cfg->overflow_capacity, {0, 0});
    e->snapshot.assign(n, 0);
    return e;
}
void orc_engine_destroy(void* h) { delete (Engine*)h; }
void orc_set_flags(void* h, uint32_t flags) { ((Engine*)h)->cfg.flags = flags; }
void orc_set_threads(void* h, int t) { ((Engine*)h)->threads = t > 0 ? t : 1; }

uint32_t orc_submit(void* h, const uint32_t* u, const uint32_t* v, const uint8_t* t, uint64_t count, Stats* per_epoch,
                    uint32_t max_epochs) {
    return ((Engine*)h)->submit(u, v, t, count, per_epoch, max_epochs);
}
void orc_apply(void* h, const uint32_t* u, const uint32_t* v, const uint8_t* t, uint64_t count) {
    Engine* e = (Engine*)h;
    for (uint64_t i = 0; i < count; i++) e->apply(u[i], v[i], t ? t[i] : 0);
}
void orc_epoch(void* h, Stats* st) { ((Engine*)h)->epoch(*st); }
void orc_finish(void* h, Stats* st) { ((Engine*)h)->finish(*st); }
void orc_tier1(void* h, Stats* st) { ((Engine*)h)->tier1(*st); }
// One Tier-1 sweep on the current bitmaps (no rebuild first); returns its conflict count.
uint32_t orc_tier1_sweep(void* h) { return ((Engine*)h)->tier1_sweep(); }
void orc_tier2(void* h, Stats* st) { ((Engine*)h)->tier2(*st); }
void orc_rebuild_bitmaps(void* h) { ((Engine*)h)->rebuild_bitmaps(); }
void orc_compact(void* h) { ((Engine*)h)->compact(); }

uint64_t orc_num_slots(void* h) { return ((Engine*)h)->row_ptr[((Engine*)h)->n]; }
uint32_t orc_num_hd(void* h) { return (uint32_t)(((Engine*)h)->hist.size() / HIST_BINS); }
void orc_get_words(void* h, uint32_t* out) { Engine* e = (Engine*)h; memcpy(out, e->word.data(), e->n * 4ull); }
void orc_set_words(void* h, const uint32_t* in) { Engine* e = (Engine*)h; memcpy(e->word.data(), in, e->n * 4ull); }
void orc_get_snapshot(void* h, uint16_t* out) { Engine* e = (Engine*)h; memcpy(out, e->snapshot.data(), e->n * 2ull); }
void orc_get_bitmaps(void* h, uint64_t* out) { Engine* e = (Engine*)h; memcpy(out, e->bitmap.data(), e->n * 8ull); }
void orc_set_bitmaps(void* h, const uint64_t* in) { Engine* e = (Engine*)h; memcpy(e->bitmap.data(), in, e->n * 8ull); }
void orc_get_hist(void* h, uint32_t* out) {
    Engine* e = (Engine*)h;
    if (!e->hist.empty()) memcpy(out, e->hist.data(), e->hist.size() * 4);
}
void orc_set_hist(void* h, const uint32_t* in) {
    Engine* e = (Engine*)h;
    if (!e->hist.empty()) memcpy(e->hist.data(), in, e->hist.size() * 4);
}
void orc_get_csr(void* h, uint32_t* rp, uint32_t* col) {
    Engine* e = (Engine*)h;
    memcpy(rp, e->row_ptr.data(), (e->n + 1) * 4ull);
    memcpy(col, e->col.data(), e->col.size() * 4);
}
void orc_set_col(void* h, const uint32_t* col) { Engine* e = (Engine*)h; memcpy(e->col.data(), col, e->col.size() * 4); }
uint32_t orc_overflow_count(void* h) { return ((Engine*)h)->overflow_count; }
uint16_t orc_find_color(void* h, uint32_t v) { return ((Engine*)h)->fas(v); }

// ---------------------------------------------------------------------------
// Validator on a plain edge list (u[i], v[i]) and a color array: counts edges
// whose endpoints share a non-zero color and vertices left uncolored with an
// edge.  Follows A/src/streamgc.cu:448-556 in what it counts; the edge list
// it is given comes from the caller (the live set), not from the engine.
void orc_validate_edges(uint32_t n, const uint32_t* eu, const uint32_t* ev, uint64_t m, const uint16_t* colors,
                        uint64_t* violations, uint64_t* uncolored_with_edges) {
    uint64_t bad = 0;
    std::vector<uint8_t> touched(n, 0);
    for (uint64_t i = 0; i < m; i++) {
        uint32_t a = eu[i], b = ev[i];
        if (a >= n || b >= n || a == b) continue;
        touched[a] = touched[b] = 1;
        if (colors[a] == colors[b] && colors[a] != 0) bad++;
    }
    uint64_t unc = 0;
    for (uint32_t v = 0; v < n; v++)
        if (touched[v] && colors[v] == 0) unc++;
    *violations = bad;
    *uncolored_with_edges = unc;
}

// A/src/streamgc.cu:559-592 (first fit in vertex-id order on the base graph)
uint32_t orc_static_greedy(uint32_t n, const uint32_t* row_ptr, const uint32_t* col) {
    std::vector<uint16_t> c(n, 0);
    std::vector<uint8_t> used(n + 2, 0);
    uint32_t mx = 0;
    for (uint32_t v = 0; v < n; v++) {
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
            if (c[col[j]]) used[c[col[j]]] = 1;
        uint32_t k = 1;
        while (used[k]) k++;
        c[v] = (uint16_t)k;
        mx = std::max(mx, k);
        for (uint32_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
            if (c[col[j]]) used[c[col[j]]] = 0;
    }
    return mx;
}

}  // extern "C"
