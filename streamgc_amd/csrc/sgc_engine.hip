// sgc_engine.hip -- the StreamGC engine: a pool of graphs resident in HBM.
//
// Every graph has its own slack CSR, packed vertex words, bitmaps, hub
// histograms, overflow buffer and double-buffered snapshot (layout below).
// Every phase of the hot path is ONE launch over all graphs taking part: the
// kernels walk the vertices of the graphs taking part laid end to end (a
// prefix of their vertex counts and their ids), so 256 small graphs cost the
// same number of launches as one large graph, and a phase for a few graphs of
// a large pool touches only their vertices.  The single-graph C ABI (sgc_create ...) is a
// pool of one.
//
// Per-graph layout: word[n] u32, deg0[n] u32, row_ptr[n+1] u32, fill[n] u32, col[S] u32,
// bitmap[n] u64, hd_map[n] u32, hubs[nhd] u32, hist[nhd*1024] u32,
// overflow 2*cap u32, snapshot 2 x n u16, 5 mark arrays of n (rounded to 4) u8;
// counters of all graphs in one array (one copy reads them all).
//
// Maintenance sweeps (compaction, rebuild, Tier-1, Tier-2 check) visit only
// the vertices marked dirty since the sweep last ran: a worklist is built on
// the device from the dirty bytes, and the sweep reads its length there.
//
// All work of a pool is enqueued on one HIP stream; the host synchronises only
// for data-dependent decisions (Tier-1 convergence, the Tier-2 check).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sgc_device.hpp"
#include "sgc_internal.hpp"
#include "sgc_launch.hpp"
#include "sgc_round_control.hpp"
#include "streamgc.h"

using namespace sgc;

namespace {

std::string g_last_error;
std::mutex g_err_mu;

int global_fail(int code, const std::string& m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = m;
    return code;
}

struct KTime {
    double ms = 0;
    uint64_t launches = 0;
};
struct PendingTiming {
    const char* name;
    hipEvent_t a, b;
};

// Mark arrays are padded to a multiple of 16 vertices: sweeps read the marks
// of 4 vertices per 4-byte word, the serial sweep 16 per lane.
uint32_t mark_stride(uint32_t n) { return (n + 15u) & ~15u; }
constexpr uint32_t ABL_MASK = SGC_ABL_FORCE_BITMAP | SGC_ABL_FORCE_HISTOGRAM | SGC_ABL_NO_OVERFLOW | SGC_ABL_LIVE_DEGREE;

// Bytes one graph occupies on the device (admission control).
uint64_t graph_bytes(uint32_t n, uint64_t S, uint32_t nhd, uint32_t ovf_cap) {
    return 28ull * n + 4ull * mark_stride(n) + 4 + (uint64_t)M_ARRAYS * mark_stride(n) + 4 * S + 4ull * nhd +
           4096ull * nhd +
           8ull * (ovf_cap ? ovf_cap : 1);
}

// Slack capacities (A/src/streamgc.cu:373-378) and histogram rows
// (A/src/streamgc.cu:226-250) of a graph.
int plan_graph(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* deg0,
               std::vector<uint32_t>& rp, std::vector<uint32_t>& hd, std::vector<uint32_t>& hubs, uint64_t& S) {
    rp.assign((size_t)n + 1, 0);
    S = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint64_t d = row_ptr[v + 1] - row_ptr[v];
        uint64_t cap = d + d * cfg->slack_pct / 100;
        if (cap < cfg->min_row_capacity) cap = cfg->min_row_capacity;
        S += cap;
        if (S >= 0xFFFFFFFFull) return SGC_ERR_RANGE;
        rp[v + 1] = (uint32_t)S;
    }
    hd.assign(n, INVALID);
    hubs.clear();
    for (uint32_t v = 0; v < n; v++)
        if (deg0[v] > cfg->histogram_threshold) {
            hd[v] = (uint32_t)hubs.size();
            hubs.push_back(v);
        }
    return SGC_OK;
}

}  // namespace

struct GraphState {
    bool used = false;
    sgc_config cfg{};
    uint32_t n = 0, nhd = 0;
    uint64_t S = 0;
    uint64_t bytes = 0;
    uint32_t *word = nullptr, *deg0 = nullptr, *row_ptr = nullptr, *col = nullptr, *hd_map = nullptr, *fill = nullptr;
    uint32_t *hubs = nullptr, *hist = nullptr, *ovf = nullptr;
    uint64_t* bitmap = nullptr;
    uint16_t* snap = nullptr;  // 2 * n
    uint8_t* dirty = nullptr;  // M_ARRAYS x mark_stride(n)
    int published = -1;
    uint64_t since_epoch = 0;
    uint32_t epochs = 0;
    uint16_t* history = nullptr;
    uint32_t hist_cap = 0, hist_count = 0;
    hipEvent_t pub_ev[2] = {nullptr, nullptr};
    hipEvent_t read_ev[2] = {nullptr, nullptr};
    // Tier-1 loop state of the current repair
    uint32_t prev = 0, plateau = 0;
    bool iterating = false;
    bool t1_clean = false;     // the last Tier-1 sweep found nothing to repair
    bool settled = false;      // a repair has completed since the graph was loaded or its state uploaded
    sgc_epoch_stats st{};
    double t2_ms = 0;          // host Tier-2 time and runs
    uint32_t t2_count = 0;

    void release() {
        for (void* p : {(void*)word, (void*)deg0, (void*)row_ptr, (void*)col, (void*)hd_map, (void*)hubs, (void*)hist,
                        (void*)ovf, (void*)bitmap, (void*)snap, (void*)history, (void*)dirty, (void*)fill})
            if (p) (void)hipFree(p);
        for (int i = 0; i < 2; i++) {
            if (pub_ev[i]) (void)hipEventDestroy(pub_ev[i]);
            if (read_ev[i]) (void)hipEventDestroy(read_ev[i]);
        }
        *this = GraphState();
    }
};

// A pinned host mirror of a small device array, uploaded in stream order.
// The mirror rotates over STAGED_RING pinned copies: writable() moves to the
// copy uploaded STAGED_RING uploads ago once that copy's upload has completed
// (nearly always long done), so the host rarely waits for the GPU to catch
// up before it writes the next parameters.  keep = false when the caller
// rewrites everything it uploads (the copy's old contents are then left as
// they are); keep = true carries the current contents over.
constexpr int STAGED_RING = 4;
template <class T>
struct Staged {
    T* h = nullptr;
    T* d = nullptr;
    size_t count = 0;
    T* ring[STAGED_RING] = {};
    hipEvent_t done[STAGED_RING] = {};
    int cur = 0;
    hipError_t alloc(size_t c) {
        count = c;
        hipError_t e = hipMalloc((void**)&d, c * sizeof(T));
        for (int k = 0; k < STAGED_RING && e == hipSuccess; k++) {
            e = hipHostMalloc((void**)&ring[k], c * sizeof(T));
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
            if (e == hipSuccess) memset(ring[k], 0, c * sizeof(T));
        }
        cur = 0;
        h = ring[0];
        return e;
    }
    hipError_t writable(bool keep = true) {
        const int next = (cur + 1) % STAGED_RING;
        hipError_t e = hipEventSynchronize(done[next]);
        if (e != hipSuccess) return e;
        if (keep) memcpy(ring[next], ring[cur], count * sizeof(T));
        cur = next;
        h = ring[cur];
        return hipSuccess;
    }
    hipError_t upload(hipStream_t s, size_t first = 0, size_t n = (size_t)-1) {
        if (n == (size_t)-1) n = count - first;
        if (n == 0) return hipSuccess;
        hipError_t e = hipMemcpyAsync(d + first, h + first, n * sizeof(T), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(done[cur], s);
        return e;
    }
    void release() {
        for (int k = 0; k < STAGED_RING; k++) {
            if (done[k]) (void)hipEventSynchronize(done[k]);
            if (ring[k]) (void)hipHostFree(ring[k]);
            if (done[k]) (void)hipEventDestroy(done[k]);
            ring[k] = nullptr;
            done[k] = nullptr;
        }
        if (d) (void)hipFree(d);
        h = nullptr;
        d = nullptr;
    }
};

// A prepared batched run (sgc_pool_batch_prepare): the mixed stream resident
// on the device in round order, the round plan, the device-side control of
// the captured rounds and one instantiated hipGraph per launch-shape
// candidate.
struct BatchRun {
    sgc_batch_config cfg{};
    uint32_t rounds = 0, ng = 0, max_count = 0, sweeps = 0, ovf_cap = 0, t2_passes = 0;
    uint32_t sweeps_max = 0;                   // most Tier-1 sweeps a capture may hold
    uint64_t updates = 0;
    bool multi = false, any_history = false, all_settled = true;
    bool lean = false;                         // captures leave the Tier-2 check and repair out
    std::vector<uint32_t> tab;                 // per round: first record, record count
    std::vector<uint8_t> inround, full;        // rounds x ng: takes part; reaches its epoch boundary
    std::vector<uint32_t> counts;              // rounds x ng: records
    std::vector<std::vector<uint32_t>> gids;   // graphs of each round
    sgc_edge_update* d_stream = nullptr;
    uint32_t *d_tab = nullptr, *d_seg = nullptr, *d_ctl = nullptr, *d_fl = nullptr, *d_pub = nullptr;
    uint32_t* d_space = nullptr;               // vertex space of the graphs running device Tier-2 (dspace_words)
    uint8_t* d_inround = nullptr;
    RoundState* d_state = nullptr;
    RoundGraph* d_rg = nullptr;
    uint16_t** d_snapbase = nullptr;
    HistDev* d_hist = nullptr;
    uint32_t* h_ctl = nullptr;                 // pinned copy of the control block
    uint32_t* h_rd = nullptr;                  // pinned: the control block after each of two chunks in flight
    hipEvent_t rd_ev[2] = {nullptr, nullptr};  // each chunk's reading has arrived
    std::vector<sgc_tuning> cands;
    std::vector<hipGraphExec_t> execs;

    void release() {
        for (hipGraphExec_t x : execs)
            if (x) (void)hipGraphExecDestroy(x);
        execs.clear();
        for (void* p : {(void*)d_stream, (void*)d_tab, (void*)d_seg, (void*)d_ctl, (void*)d_fl, (void*)d_pub,
                        (void*)d_space, (void*)d_inround, (void*)d_state, (void*)d_rg, (void*)d_snapbase,
                        (void*)d_hist})
            if (p) (void)hipFree(p);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (h_rd) (void)hipHostFree(h_rd);
        for (hipEvent_t x : rd_ev)
            if (x) (void)hipEventDestroy(x);
        *this = BatchRun();
    }
};

struct sgc_pool {
    uint32_t max_graphs = 0, batch_cap = 0, flags = 0;
    sgc_tuning tune{1024, 2048, 16384, 65536, 64, 2};
    uint64_t budget = 0, used_bytes = 0;
    hipStream_t stream = nullptr, qstream = nullptr;
    hipEvent_t batch_up = nullptr;        // a host batch's upload has been read
    std::vector<GraphState> gs;
    Staged<GraphDev> graphs;              // descriptors
    Staged<uint32_t> vbase, gbase, vtile;    // all graphs: vertex prefix, edge-key offsets, 4096-vertex tile prefix
    // The graphs taking part in the next launches, uploaded at once: per-graph
    // flags (max_graphs), then the active graphs' vertex prefix
    // (max_graphs + 1), ids (max_graphs) and tile prefix (max_graphs + 1).
    Staged<uint32_t> act;
    std::vector<uint8_t> act_last;           // the flags on the device (empty: unknown)
    uint32_t act_tiles = 0, all_tiles = 0;
    size_t a_sbase() const { return max_graphs; }
    size_t a_sgid() const { return 2 * (size_t)max_graphs + 1; }
    size_t a_stile() const { return 3 * (size_t)max_graphs + 1; }
    Staged<uint32_t> perm, seg;           // deterministic per-graph replay
    uint32_t* d_ctr = nullptr;            // max_graphs * C_COUNT
    uint32_t* h_ctr = nullptr;            // pinned copy
    uint64_t* d_totals = nullptr;
    uint64_t* h_totals = nullptr;
    EdgeGroups eg;
    sgc_edge_update* d_batch = nullptr;   // staging of host batches
    void* d_rev = nullptr;                // device Tier-2: reverse-only neighbour list
    uint32_t* d_rev_count = nullptr;      // its length (2 words)
    uint32_t rev_cap = 0;
    uint32_t* d_t2_items = nullptr;       // (graph, vertex, old colour) recoloured by the host Tier-2
    size_t t2_items_cap = 0;
    uint32_t* d_list = nullptr;           // sweep worklist (virtual vertex ids)
    uint32_t* d_list_count = nullptr;
    uint32_t list_cap = 0;
    uint32_t ng = 0, V = 0;               // slots in use, vertices
    uint32_t act_n = 0, act_V = 0;        // active graphs and their vertices
    uint16_t* d_tmp16 = nullptr;
    uint32_t* d_qids = nullptr;
    uint16_t* d_qout = nullptr;
    uint32_t qcap = 0;
    int device = 0;                       // the HIP device current when the pool was created
    std::mutex qmu;
    std::mutex run_mu;                    // held by a batched run; queries wait for it (taken before qmu)
    uint64_t t1_sweeps = 0, t2_runs = 0;
    std::string err;
    bool ktiming = false;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, KTime> ktimes;
    double last_ms[4] = {0, 0, 0, 0};
    BatchRun* batch = nullptr;            // prepared batched run

    bool deterministic() const { return (flags & SGC_DETERMINISTIC) != 0; }
    bool exact_hist() const { return (flags & SGC_EXACT_HISTOGRAM) != 0; }
    bool repair_every_epoch() const { return (flags & SGC_REPAIR_EVERY_EPOCH) != 0; }
    bool tracking() const { return (flags & SGC_FULL_SWEEPS) == 0; }
    bool rebuild_each_sweep() const { return deterministic() || tune.fresh_repair < 2; }
    // Tier-1 selection of a graph: exact neighbour colours (tune.fresh_repair)
    // once it has been repaired; until then -- its first repair colours and
    // settles every vertex of its rows at once -- the reference's selection
    // from rebuilt bitmaps and histograms, whose result does not depend on
    // the order in which the wavefronts run.
    int fresh_level(uint32_t g) const { return deterministic() || !gs[g].settled ? 0 : (int)tune.fresh_repair; }

    int fail(int code, const std::string& msg) {
        err = msg;
        return global_fail(code, msg);
    }
    int hip_fail(hipError_t e, const char* what) { return fail(SGC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)); }

    LaunchCtx ctx(bool use_active) const {
        LaunchCtx L;
        L.stream = stream;
        L.graphs = graphs.d;
        L.ng = ng ? ng : 1;
        L.active = use_active ? act.d : nullptr;
        if (use_active && act_n != ng) {
            L.sbase = act.d + a_sbase();
            L.sgid = act.d + a_sgid();
            L.ns = act_n;
            L.V = act_V;
            L.stile = act.d + a_stile();
            L.tiles = act_tiles;
        } else {
            L.sbase = ng > 1 ? vbase.d : nullptr;
            L.sgid = nullptr;
            L.ns = ng ? ng : 1;
            L.V = V;
            L.stile = vtile.d;
            L.tiles = all_tiles;
        }
        L.gbase = gbase.d;
        L.totals = d_totals;
        L.max_blocks = tune.apply_blocks;
        L.sweep_blocks = tune.sweep_blocks;
        L.long_row = tune.long_row;
        return L;
    }

    // ------------------------------------------------------------ timing
    hipEvent_t get_event() {
        if (!free_events.empty()) {
            hipEvent_t e = free_events.back();
            free_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    template <class F>
    hipError_t timed(const char* name, F&& launch) {
        if (!ktiming) return launch();
        hipEvent_t a = get_event(), b = get_event();
        (void)hipEventRecord(a, stream);
        hipError_t r = launch();
        (void)hipEventRecord(b, stream);
        pending.push_back({name, a, b});
        return r;
    }
    void resolve_timing() {
        for (auto& p : pending) {
            (void)hipEventSynchronize(p.b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, p.a, p.b);
            KTime& k = ktimes[p.name];
            k.ms += ms;
            k.launches++;
            free_events.push_back(p.a);
            free_events.push_back(p.b);
        }
        pending.clear();
    }

    // ------------------------------------------------------------ host mirrors
    hipError_t read_ctr() {
        hipError_t e = hipMemcpyAsync(h_ctr, d_ctr, (size_t)ng * C_COUNT * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        return e;
    }
    uint32_t ctr(uint32_t gid, uint32_t idx) const { return h_ctr[(size_t)gid * C_COUNT + idx]; }

    // The graphs taking part in the next launches: active flags and their
    // vertex space (uploaded in stream order).  A LaunchCtx made with
    // ctx(true) describes the set current when it was made.
    hipError_t set_active(const std::vector<uint32_t>& gids) {
        std::vector<uint8_t> fl(ng, 0);
        for (uint32_t g : gids) fl[g] = 1;
        if (fl == act_last) return hipSuccess;   // the same set is on the device already
        hipError_t e = act.writable(false);      // rewritten below
        if (e != hipSuccess) return e;
        uint32_t* flags_h = act.h;
        uint32_t* sbase_h = act.h + a_sbase();
        uint32_t* sgid_h = act.h + a_sgid();
        uint32_t* stile_h = act.h + a_stile();
        uint64_t v = 0;
        uint32_t tiles = 0;
        act_n = 0;
        for (uint32_t g = 0; g < ng; g++) {
            flags_h[g] = fl[g];
            if (!fl[g]) continue;
            sbase_h[act_n] = (uint32_t)v;
            stile_h[act_n] = tiles;
            sgid_h[act_n] = g;
            act_n++;
            v += gs[g].n;
            tiles += (gs[g].n + 4095) / 4096;
        }
        sbase_h[act_n] = (uint32_t)v;
        stile_h[act_n] = tiles;
        act_V = (uint32_t)v;
        act_tiles = tiles;
        e = act.upload(stream);
        act_last = e == hipSuccess ? fl : std::vector<uint8_t>();
        return e;
    }
    hipError_t set_active_iterating(const std::vector<uint32_t>& gids) {
        std::vector<uint32_t> it;
        for (uint32_t g : gids)
            if (gs[g].iterating) it.push_back(g);
        return set_active(it);
    }

    hipError_t refresh_layout() {
        hipError_t e = vbase.writable(false);   // rewritten below
        if (e == hipSuccess) e = gbase.writable(false);
        if (e == hipSuccess) e = vtile.writable(false);
        if (e != hipSuccess) return e;
        ng = 0;
        for (uint32_t g = 0; g < max_graphs; g++)
            if (gs[g].used) ng = g + 1;
        uint64_t v = 0;
        uint32_t tiles = 0;
        for (uint32_t g = 0; g < ng; g++) {
            vbase.h[g] = (uint32_t)v;
            gbase.h[g] = (uint32_t)v;
            vtile.h[g] = tiles;
            if (gs[g].used) {
                v += gs[g].n;
                tiles += (gs[g].n + 4095) / 4096;
            }
        }
        vbase.h[ng] = (uint32_t)v;
        vtile.h[ng] = tiles;
        V = (uint32_t)v;
        all_tiles = tiles;
        act_n = ng;   // every graph until the next set_active
        act_V = V;
        act_tiles = tiles;
        act_last.clear();
        e = vbase.upload(stream, 0, ng + 1);
        if (e == hipSuccess) e = gbase.upload(stream, 0, ng + 1);
        if (e == hipSuccess) e = vtile.upload(stream, 0, ng + 1);
        if (e == hipSuccess && V > list_cap) {
            e = hipStreamSynchronize(stream);
            if (d_list) (void)hipFree(d_list);
            d_list = nullptr;
            list_cap = 0;
            if (e == hipSuccess) e = hipMalloc((void**)&d_list, list_words(V, max_graphs) * 4);
            if (e == hipSuccess) list_cap = V;
        }
        return e;
    }

    // The sweep context: the vertices of the active graphs marked for sweep
    // `kind` (their marks are cleared) or sticky; every vertex without tracking.
    hipError_t sweep(LaunchCtx& L, uint32_t kind) {
        if (!tracking()) return hipSuccess;
        L.list = d_list;
        L.list_count = d_list_count;
        L.list_cap = list_cap;
        return timed("build_list", [&] { return launch_build_list(L, kind, d_list, list_cap, d_list_count); });
    }
    // All sweeps must visit every vertex of graph gid (its state changed from the host).
    hipError_t mark_all(uint32_t gid) {
        GraphState& G = gs[gid];
        G.settled = false;
        return hipMemsetAsync(G.dirty, 1, (size_t)M_STICKY * mark_stride(G.n), stream);
    }
    // Descriptors follow the tracking flag (dirty == null: no marking).
    hipError_t set_tracking(bool on) {
        hipError_t e = graphs.writable();
        for (uint32_t g = 0; g < ng && e == hipSuccess; g++) {
            if (!gs[g].used) continue;
            graphs.h[g].dirty = on ? gs[g].dirty : nullptr;
            if (on) e = mark_all(g);
        }
        if (e == hipSuccess && ng) e = graphs.upload(stream, 0, ng);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        return e;
    }

    hipError_t upload_descriptor(uint32_t gid) { return graphs.upload(stream, gid, 1); }
    // The ablation flags travel in every descriptor.
    hipError_t refresh_flags() {
        hipError_t e = graphs.writable();
        for (uint32_t g = 0; g < ng && e == hipSuccess; g++)
            if (gs[g].used) graphs.h[g].flags = flags & ABL_MASK;
        if (e == hipSuccess && ng) e = graphs.upload(stream, 0, ng);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        return e;
    }

    // ------------------------------------------------------------ graphs
    int add_graph(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                  const uint32_t* deg0, uint32_t* gid_out);
    int remove_graph(uint32_t gid);

    // ------------------------------------------------------------ hot path
    int apply(const sgc_edge_update* upd, uint32_t count, bool on_device, std::vector<uint64_t>& per_graph);
    int epoch(const std::vector<uint32_t>& gids, bool finish, std::vector<sgc_epoch_stats>* out);
    hipError_t rebuild(LaunchCtx L) {
        hipError_t e = sweep(L, M_BITS);
        if (e == hipSuccess) e = timed("rebuild", [&] { return launch_rebuild(L, exact_hist()); });
        return e;
    }
    int tier1(const std::vector<uint32_t>& gids, uint32_t start_it = 0);
    int tier2(const std::vector<uint32_t>& gids);
    int host_tier2(const std::vector<uint32_t>& list, std::vector<uint32_t>& ran);
    // Scratch of the device Tier-2's reverse-only neighbour list (12 bytes an entry).
    int ensure_rev_scratch() {
        const uint32_t want = std::max<uint32_t>(1u << 20, V);
        if (rev_cap >= want) return SGC_OK;
        hipError_t e = hipStreamSynchronize(stream);
        if (d_rev) (void)hipFree(d_rev);
        d_rev = nullptr;
        rev_cap = 0;
        if (e == hipSuccess) e = hipMalloc(&d_rev, (size_t)want * 12);
        if (e == hipSuccess) rev_cap = want;
        if (e != hipSuccess) return hip_fail(e, "tier-2 scratch");
        return SGC_OK;
    }
    int device_tier2(const std::vector<uint32_t>& list, std::vector<uint32_t>& host_jobs, std::vector<uint32_t>& ran);
    int publish(const std::vector<uint32_t>& gids);
    int read_drain(const std::vector<uint32_t>& gids);
    int max_colors(const std::vector<uint32_t>& gids);

    // ------------------------------------------------------------ batched runs
    int batch_prepare(const sgc_edge_update* upd, uint64_t count, const sgc_batch_config& cfg, uint32_t* rounds);
    int batch_plan(const sgc_edge_update* upd, uint64_t count, const sgc_batch_config& cfg, uint32_t* rounds);
    int batch_capture(const sgc_tuning& t, hipGraphExec_t* out);
    int batch_run(sgc_batch_stats* out);
    int batch_run_host(BatchRun& B);
    int batch_host_round(BatchRun& B, uint32_t r);
    hipError_t batch_push(BatchRun& B);
    hipError_t batch_pull(BatchRun& B);
    void batch_drop();

    void release();
};

// ============================================================== graph admission
int sgc_pool::add_graph(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                        const uint32_t* deg0, uint32_t* gid_out) {
    batch_drop();   // a prepared plan names the graphs of the pool
    if (!cfg || !row_ptr || !deg0 || n == 0) return fail(SGC_ERR_INVALID, "add graph: bad argument");
    if (cfg->batch_size == 0 || cfg->epoch_size == 0) return fail(SGC_ERR_INVALID, "add graph: zero batch/epoch");
    const uint64_t nnz = row_ptr[n];
    if (nnz && !col_idx) return fail(SGC_ERR_INVALID, "add graph: col_idx missing");
    for (uint32_t v = 0; v < n; v++)
        if (row_ptr[v + 1] < row_ptr[v]) return fail(SGC_ERR_INVALID, "add graph: row_ptr not monotone");
    for (uint64_t j = 0; j < nnz; j++)
        if (col_idx[j] >= n) return fail(SGC_ERR_INVALID, "add graph: column id out of range");
    uint32_t gid = max_graphs;
    for (uint32_t g = 0; g < max_graphs; g++)
        if (!gs[g].used) {
            gid = g;
            break;
        }
    if (gid == max_graphs) return fail(SGC_ERR_NOMEM, "add graph: pool is full");
    if ((uint64_t)V + n >= 0xFFFFFFFFull) return fail(SGC_ERR_RANGE, "add graph: pool vertex space exhausted");
    std::vector<uint32_t> rp, hd, hub_list;
    uint64_t S = 0;
    if (plan_graph(cfg, n, row_ptr, deg0, rp, hd, hub_list, S) != SGC_OK)
        return fail(SGC_ERR_RANGE, "add graph: slot count exceeds 32-bit offsets");
    const uint32_t nhd = (uint32_t)hub_list.size();
    const uint64_t bytes = graph_bytes(n, S, nhd, cfg->overflow_capacity);
    if (budget && used_bytes + bytes > budget)
        return fail(SGC_ERR_NOMEM, "add graph: over the pool's HBM budget");

    GraphState& G = gs[gid];
    G = GraphState();
    G.cfg = *cfg;
    G.n = n;
    G.S = S;
    G.nhd = nhd;
    G.bytes = bytes;
    hipStream_t s = stream;
    hipError_t e = hipSuccess;
    auto A = [&](void** p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b ? b : 4);
    };
    A((void**)&G.word, (size_t)n * 4);
    A((void**)&G.deg0, (size_t)n * 4);
    A((void**)&G.row_ptr, ((size_t)n + 1) * 4);
    A((void**)&G.col, S * 4);
    A((void**)&G.fill, (size_t)mark_stride(n) * 4);   // padded: sweeps read fill 4 vertices at a time
    A((void**)&G.bitmap, (size_t)n * 8);
    A((void**)&G.hd_map, (size_t)n * 4);
    A((void**)&G.hubs, (size_t)nhd * 4);
    A((void**)&G.hist, (size_t)nhd * HIST_BINS * 4);
    A((void**)&G.ovf, (size_t)cfg->overflow_capacity * 8);
    A((void**)&G.snap, (size_t)n * 4);
    A((void**)&G.dirty, (size_t)M_ARRAYS * mark_stride(n));
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipEventCreateWithFlags(&G.pub_ev[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&G.read_ev[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipMemsetAsync(G.word, 0, (size_t)n * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(G.bitmap, 0, (size_t)n * 8, s);
    if (e == hipSuccess && nhd) e = hipMemsetAsync(G.hist, 0, (size_t)nhd * HIST_BINS * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(G.snap, 0, (size_t)n * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(G.dirty, 1, (size_t)M_STICKY * mark_stride(n), s);
    if (e == hipSuccess)
        e = hipMemsetAsync(G.dirty + (size_t)M_STICKY * mark_stride(n), 0, (size_t)(M_ARRAYS - M_STICKY) * mark_stride(n), s);
    if (e == hipSuccess) e = hipMemsetAsync(d_ctr + (size_t)gid * C_COUNT, 0, C_COUNT * 4, s);
    if (e == hipSuccess) e = hipMemcpyAsync(G.deg0, deg0, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(G.row_ptr, rp.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(G.hd_map, hd.data(), (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nhd) e = hipMemcpyAsync(G.hubs, hub_list.data(), (size_t)nhd * 4, hipMemcpyHostToDevice, s);
    // prefill the slack rows on the device from the loaded CSR (A/src/streamgc.cu:400-432)
    uint32_t *rp0 = nullptr, *col0 = nullptr;
    A((void**)&rp0, ((size_t)n + 1) * 4);
    A((void**)&col0, nnz * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(rp0, row_ptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(col0, col_idx, nnz * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_fill_rows(s, n, rp0, col0, G.row_ptr, G.col, G.fill);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (rp0) (void)hipFree(rp0);
    if (col0) (void)hipFree(col0);
    if (e == hipSuccess) e = graphs.writable();
    if (e != hipSuccess) {
        G.release();
        return hip_fail(e, "add graph");
    }
    GraphDev& d = graphs.h[gid];
    d = GraphDev();
    d.word = G.word;
    d.deg0 = G.deg0;
    d.row_ptr = G.row_ptr;
    d.col = G.col;
    d.fill = G.fill;
    d.bitmap = G.bitmap;
    d.hd_map = G.hd_map;
    d.hubs = G.hubs;
    d.hist = G.hist;
    d.ovf = G.ovf;
    d.ctr = d_ctr + (size_t)gid * C_COUNT;
    d.snap_write = G.snap;
    d.dirty = tracking() ? G.dirty : nullptr;
    d.dstride = mark_stride(n);
    d.n = n;
    d.ovf_cap = cfg->overflow_capacity;
    d.nhd = nhd;
    d.flags = flags & ABL_MASK;
    G.used = true;
    used_bytes += bytes;
    e = upload_descriptor(gid);
    if (e == hipSuccess) e = refresh_layout();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "add graph");
    *gid_out = gid;
    return SGC_OK;
}

int sgc_pool::remove_graph(uint32_t gid) {
    batch_drop();
    if (gid >= max_graphs || !gs[gid].used) return fail(SGC_ERR_INVALID, "remove graph: no such graph");
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipStreamSynchronize(qstream);
    if (e != hipSuccess) return hip_fail(e, "remove graph");
    used_bytes -= gs[gid].bytes;
    gs[gid].release();
    e = graphs.writable();
    if (e == hipSuccess) {
        graphs.h[gid] = GraphDev();
        e = upload_descriptor(gid);
    }
    if (e == hipSuccess) e = refresh_layout();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "remove graph");
    return SGC_OK;
}

// ============================================================== updates
// One launch over a batch (local_update_kernel, A/src/streamgc.cu:847-854).
// Parallel: one wavefront per update, same-edge updates grouped in stream
// order.  Deterministic: one wavefront per graph, that graph's updates in
// stream order.  per_graph receives the updates per graph.
int sgc_pool::apply(const sgc_edge_update* upd, uint32_t count, bool on_device, std::vector<uint64_t>& per_graph) {
    per_graph.assign(ng, 0);
    if (count == 0) return SGC_OK;
    if (count > batch_cap) return fail(SGC_ERR_INVALID, "apply: batch larger than the pool's batch capacity");
    const bool multi = ng > 1;
    const sgc_edge_update* src = upd;
    hipError_t e = hipSuccess;
    if (!on_device) {
        e = hipMemcpyAsync(d_batch, upd, (size_t)count * sizeof(sgc_edge_update), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(batch_up, stream);
        if (e != hipSuccess) return hip_fail(e, "batch upload");
        src = d_batch;
    }
    // per-graph counts (and the deterministic order) need the graph ids
    std::vector<uint32_t> host_gid;
    if (multi) {
        if (on_device) {
            std::vector<sgc_edge_update> tmp(count);
            e = hipMemcpyAsync(tmp.data(), upd, (size_t)count * sizeof(sgc_edge_update), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "batch read");
            host_gid.resize(count);
            for (uint32_t i = 0; i < count; i++) memcpy(&host_gid[i], tmp[i].pad + 3, 4);
        } else {
            host_gid.resize(count);
            for (uint32_t i = 0; i < count; i++) memcpy(&host_gid[i], upd[i].pad + 3, 4);
        }
        for (uint32_t i = 0; i < count; i++) {
            if (host_gid[i] >= ng || !gs[host_gid[i]].used) return fail(SGC_ERR_INVALID, "apply: unknown graph id");
            per_graph[host_gid[i]]++;
        }
    } else {
        if (!on_device)
            for (uint32_t i = 0; i < count; i++) {
                uint32_t gid;
                memcpy(&gid, upd[i].pad + 3, 4);
                if (gid != 0) return fail(SGC_ERR_INVALID, "apply: graph id set on a single-graph batch");
            }
        per_graph[0] = count;
    }
    LaunchCtx L = ctx(false);
    if (deterministic()) {
        e = perm.writable(false);   // rewritten below, as far as uploaded
        if (e == hipSuccess) e = seg.writable(false);
        if (e != hipSuccess) return hip_fail(e, "segments");
        uint32_t nseg = 0;
        if (!multi) {
            seg.h[0] = 0;
            seg.h[1] = count;
            nseg = 1;
        } else {
            std::vector<uint32_t> start(ng + 1, 0);
            for (uint32_t g = 0; g < ng; g++) start[g + 1] = start[g] + (uint32_t)per_graph[g];
            std::vector<uint32_t> at(start.begin(), start.end() - 1);
            for (uint32_t i = 0; i < count; i++) perm.h[at[host_gid[i]]++] = i;
            for (uint32_t g = 0; g < ng; g++)
                if (start[g + 1] > start[g]) {
                    seg.h[2 * nseg] = start[g];
                    seg.h[2 * nseg + 1] = start[g + 1];
                    nseg++;
                }
            e = perm.upload(stream, 0, count);
        }
        if (e == hipSuccess) e = seg.upload(stream, 0, 2 * nseg);
        if (e == hipSuccess)
            e = timed("apply", [&] { return launch_apply_seg(L, src, seg.d, nseg, multi ? perm.d : nullptr, multi); });
    } else {
        e = timed("group", [&] { return launch_group(L, src, count, multi, &eg); });
        if (e == hipSuccess) e = timed("apply", [&] { return launch_apply(L, src, count, multi, &eg); });
    }
    if (e != hipSuccess) return hip_fail(e, "apply launch");
    if (!on_device) {
        // the caller may reuse its records once the upload has read them; the
        // launches run on (the next batch's upload waits for them in stream order)
        e = hipEventSynchronize(batch_up);
        if (e != hipSuccess) return hip_fail(e, "apply");
    }
    for (uint32_t g = 0; g < ng; g++) gs[g].since_epoch += per_graph[g];
    return SGC_OK;
}

// ============================================================== epoch pieces
int sgc_pool::read_drain(const std::vector<uint32_t>& gids) {
    hipError_t e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "counters");
    for (uint32_t g : gids) {
        gs[g].st.overflow_count = ctr(g, C_OVF_RAW);
        gs[g].st.overflow_drained = ctr(g, C_OVF_RAW) ? ctr(g, C_OVF_SNAP) : 0;
        gs[g].st.overflow_failed = ctr(g, C_DRAIN_FAILED);
    }
    return SGC_OK;
}

// run_conflict_repair (A/src/streamgc.cu:1317-1393) for every graph of the set
// at once: each iteration is one rebuild and one sweep over the graphs still
// iterating; each graph stops on its own (converged, plateau, iteration cap).
// start_it > 0 continues a loop whose state (iterating, prev, plateau,
// statistics) the caller has set: the sweeps a captured round ran.
int sgc_pool::tier1(const std::vector<uint32_t>& gids, uint32_t start_it) {
    uint32_t max_iters = 0;
    for (uint32_t g : gids) {
        if (start_it == 0) {
            gs[g].prev = 0xFFFFFFFFu;
            gs[g].plateau = 0;
            gs[g].iterating = gs[g].cfg.max_repair_iters > 0;
            gs[g].t1_clean = false;
            gs[g].st.t1_iterations = 0;
        }
        max_iters = std::max(max_iters, gs[g].cfg.max_repair_iters);
    }
    for (uint32_t it = start_it; it < max_iters; it++) {
        bool any = false;
        for (uint32_t g : gids) any |= gs[g].iterating;
        if (!any) break;
        hipError_t e = hipSuccess;
        // The reference rebuilds before every sweep (A/src/streamgc.cu:1317-1393).
        // A parallel-mode sweep that selects from exact neighbour colours for
        // every row (fresh level 2) reads no bitmap or histogram, so there the
        // rebuild runs once, after the repair.
        std::vector<uint32_t> reb;
        for (uint32_t g : gids)
            if (gs[g].iterating && fresh_level(g) < 2) reb.push_back(g);
        if (!reb.empty()) {
            e = set_active(reb);
            if (e == hipSuccess) e = rebuild(ctx(true));
        }
        if (e == hipSuccess) e = set_active_iterating(gids);
        LaunchCtx L = ctx(true);
        if (e == hipSuccess) e = launch_zero_ctr(L, C_CONFLICTS);
        // A graph sweeps in vertex order on one wavefront (Gauss-Seidel) when
        // deterministic, or when it is small and two parallel sweeps have left
        // a few conflicts (a racy sweep over a small conflict set can chase the
        // same conflicts for many iterations); otherwise one wavefront per
        // listed row.  Graphs of different selection levels take separate launches.
        std::vector<uint32_t> part[2][3];   // [serial][fresh level]
        for (uint32_t g : gids) {
            const GraphState& G = gs[g];
            if (!G.iterating) continue;
            const bool serial = deterministic() || (G.n <= tune.serial_repair && G.prev <= tune.serial_conflicts &&
                                                    G.st.t1_iterations >= 2);
            part[serial][fresh_level(g)].push_back(g);
        }
        for (int serial = 0; serial < 2 && e == hipSuccess; serial++)
            for (int lvl = 0; lvl < 3 && e == hipSuccess; lvl++) {
                if (part[serial][lvl].empty()) continue;
                e = set_active(part[serial][lvl]);
                LaunchCtx L1 = ctx(true);
                if (!serial) {
                    if (e == hipSuccess) e = sweep(L1, M_CHECK);
                    if (e == hipSuccess) e = timed("tier1", [&] { return launch_tier1(L1, false, lvl); });
                } else if (e == hipSuccess) {
                    e = timed("tier1_serial", [&] { return launch_tier1(L1, true, lvl); });
                }
            }
        if (e == hipSuccess) e = read_ctr();
        if (e != hipSuccess) return hip_fail(e, "tier-1");
        t1_sweeps++;
        for (uint32_t g : gids) {
            GraphState& G = gs[g];
            if (!G.iterating) continue;
            const uint32_t c = ctr(g, C_CONFLICTS);
            G.st.t1_iterations = it + 1;
            G.st.t1_conflicts_last = c;
            G.t1_clean = c == 0;
            if (c == 0 || it + 1 >= G.cfg.max_repair_iters) {
                G.iterating = false;
            } else if (c == G.prev) {
                if (++G.plateau >= G.cfg.plateau_iters) G.iterating = false;
            } else {
                G.plateau = 0;
            }
            G.prev = c;
        }
    }
    return SGC_OK;
}

// cpu_serial_repair (A/src/streamgc.cu:1057-1240), run on the host only for
// the graphs whose row check finds an uncoloured endpoint or a conflict
// (elsewhere it would change nothing).
//
// A Tier-1 sweep that found nothing to repair proves the colouring proper
// when every edge sits in both of its rows: a same-colour edge has a loser,
// whose check would have recoloured it, and an uncoloured endpoint with a live
// neighbour would have been coloured.  Those graphs skip the check (its count
// is zero) unless an overflow ever left an edge in one row only.
int sgc_pool::tier2(const std::vector<uint32_t>& gids) {
    std::vector<uint32_t> check;
    for (uint32_t g : gids) {
        if (gs[g].t1_clean && ctr(g, C_ASYM) == 0) gs[g].st.t2_remaining = 0;
        else check.push_back(g);
    }
    if (check.empty()) return SGC_OK;
    hipError_t e = set_active(check);
    LaunchCtx L = ctx(true), Lc = L;
    if (e == hipSuccess) e = launch_zero_ctr(L, C_T2BAD);
    if (e == hipSuccess) e = sweep(Lc, M_T2);
    if (e == hipSuccess) e = timed("t2_check", [&] { return launch_t2_check(Lc); });
    if (e == hipSuccess) e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "tier-2 check");
    std::vector<uint32_t> bad, host_jobs, ran;
    for (uint32_t g : check) {
        gs[g].st.t2_remaining = ctr(g, C_T2BAD);
        if (gs[g].st.t2_remaining) bad.push_back(g);
    }
    if (bad.empty()) return SGC_OK;
    int rc;
    if (tracking() && !(flags & SGC_HOST_TIER2)) {
        if ((rc = device_tier2(bad, host_jobs, ran))) return rc;
    } else {
        host_jobs = bad;
    }
    if ((rc = host_tier2(host_jobs, ran))) return rc;
    if (ran.empty()) return SGC_OK;
    e = set_active(ran);
    L = ctx(true);
    Lc = L;
    if (e == hipSuccess) e = launch_zero_ctr(L, C_T2BAD);
    if (e == hipSuccess) e = sweep(Lc, M_T2);
    if (e == hipSuccess) e = timed("t2_check", [&] { return launch_t2_check(Lc); });
    if (e == hipSuccess) e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "tier-2 check");
    for (uint32_t g : ran) gs[g].st.t2_remaining = ctr(g, C_T2BAD);
    return SGC_OK;
}

// cpu_serial_repair on the host for the graphs of `list`, all at once: one
// download, the repairs in parallel on the host cores, one upload; then only
// the recoloured vertices and their neighbourhoods are marked for the next
// sweeps (as a recolour on the device marks them).
int sgc_pool::host_tier2(const std::vector<uint32_t>& list, std::vector<uint32_t>& ran) {
    struct Job {
        uint32_t g;
        std::vector<uint32_t> w, w0, rp, c;
        uint32_t fixed = 0;
        double ms = 0;
    };
    std::vector<Job> jobs;
    for (uint32_t g : list) jobs.push_back(Job{g, {}, {}, {}, {}, 0, 0});
    if (jobs.empty()) return SGC_OK;
    hipError_t e = hipSuccess;
    auto t0 = std::chrono::steady_clock::now();
    for (Job& J : jobs) {
        const GraphState& G = gs[J.g];
        J.w.resize(G.n);
        J.rp.resize((size_t)G.n + 1);
        J.c.resize(G.S);
        if (e == hipSuccess) e = hipMemcpyAsync(J.w.data(), G.word, (size_t)G.n * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(J.rp.data(), G.row_ptr, ((size_t)G.n + 1) * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(J.c.data(), G.col, G.S * 4, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "tier-2 download");
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t k = 0; k < jobs.size(); k++) {
        Job& J = jobs[k];
        auto j0 = std::chrono::steady_clock::now();
        J.w0 = J.w;
        J.fixed = tier2_host(gs[J.g].n, J.rp.data(), J.c.data(), J.w.data(), gs[J.g].cfg.tier2_passes);
        J.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - j0).count();
    }
    std::vector<uint32_t> items;   // (graph, vertex, old colour) of every recoloured vertex
    for (Job& J : jobs)
        for (uint32_t v = 0; v < gs[J.g].n; v++)
            if (J.w[v] != J.w0[v]) {
                items.push_back(J.g);
                items.push_back(v);
                items.push_back(J.w0[v] & 0xFFFFu);
            }
    const uint32_t nitems = (uint32_t)(items.size() / 3);
    if (items.size() > t2_items_cap) {
        e = hipStreamSynchronize(stream);
        if (d_t2_items) (void)hipFree(d_t2_items);
        d_t2_items = nullptr;
        t2_items_cap = 0;
        if (e == hipSuccess) e = hipMalloc((void**)&d_t2_items, items.size() * 4);
        if (e == hipSuccess) t2_items_cap = items.size();
    }
    for (Job& J : jobs)
        if (e == hipSuccess)
            e = hipMemcpyAsync(gs[J.g].word, J.w.data(), (size_t)gs[J.g].n * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && nitems)
        e = hipMemcpyAsync(d_t2_items, items.data(), items.size() * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && nitems) e = launch_mark_recolored(stream, graphs.d, d_t2_items, nitems);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "tier-2 upload");
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    double cpu = 0;
    for (const Job& J : jobs) cpu += J.ms;
    for (const Job& J : jobs) {
        GraphState& G = gs[J.g];
        G.st.t2_fixed = J.fixed;
        G.st.t2_ran = 1;
        t2_runs++;
        G.t2_count++;
        // the transfers are shared; each graph is charged its own repair time and a share of the rest
        G.t2_ms += J.ms + std::max(0.0, total - cpu) / (double)jobs.size();
        ran.push_back(J.g);
    }
    last_ms[2] += total;
    return SGC_OK;
}

// The same repair on the device (k_t2_* kernels) for the graphs of `list`;
// a graph whose reverse-only neighbours do not fit the scratch list goes to
// the host instead (host_jobs).
int sgc_pool::device_tier2(const std::vector<uint32_t>& list, std::vector<uint32_t>& host_jobs,
                           std::vector<uint32_t>& ran) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = ensure_rev_scratch();
    if (rc) return rc;
    hipError_t e = set_active(list);
    LaunchCtx L = ctx(true);
    if (e == hipSuccess) e = launch_zero_ctr(L, C_T2TOTAL);
    if (e == hipSuccess) e = launch_zero_ctr(L, C_REVOVF);
    if (e == hipSuccess) e = timed("t2_device", [&] { return launch_t2_begin(L, d_rev, d_rev_count, rev_cap); });
    if (e == hipSuccess) e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "tier-2");
    std::vector<uint32_t> dev;
    for (uint32_t g : list) (ctr(g, C_REVOVF) ? host_jobs : dev).push_back(g);
    if (dev.empty()) return SGC_OK;
    if (dev.size() != list.size()) e = set_active(dev);
    L = ctx(true);
    if (e == hipSuccess)
        e = timed("t2_device", [&] { return launch_t2_pass(L, d_rev, d_rev_count, rev_cap, true); });
    std::vector<uint32_t> act = dev;
    for (uint32_t pass = 0; e == hipSuccess && !act.empty(); pass++) {
        e = set_active(act);
        L = ctx(true);
        if (e == hipSuccess) e = launch_zero_ctr(L, C_T2TODO);
        if (e == hipSuccess) e = launch_zero_ctr(L, C_T2FIXED);
        if (e == hipSuccess)
            e = timed("t2_device", [&] { return launch_t2_pass(L, d_rev, d_rev_count, rev_cap, false); });
        if (e == hipSuccess) e = read_ctr();
        std::vector<uint32_t> next;
        for (uint32_t g : act)
            if (pass + 1 < gs[g].cfg.tier2_passes && ctr(g, C_T2TODO) && ctr(g, C_T2FIXED)) next.push_back(g);
        act.swap(next);
    }
    if (e != hipSuccess) return hip_fail(e, "tier-2");
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (uint32_t g : dev) {
        GraphState& G = gs[g];
        G.st.t2_fixed = ctr(g, C_T2TOTAL);
        G.st.t2_ran = 1;
        t2_runs++;
        G.t2_count++;
        G.t2_ms += total / (double)dev.size();
        ran.push_back(g);
    }
    last_ms[2] += total;
    return SGC_OK;
}

// epoch_snapshot_kernel (A/src/streamgc.cu:957-958) into each graph's back
// buffer, then the buffers flip: queries move to the new snapshot.
int sgc_pool::publish(const std::vector<uint32_t>& gids) {
    hipError_t e = graphs.writable();
    if (e != hipSuccess) return hip_fail(e, "descriptors");
    std::vector<int> half(gids.size());
    for (size_t k = 0; k < gids.size(); k++) {
        GraphState& G = gs[gids[k]];
        half[k] = G.published == 0 ? 1 : 0;
        graphs.h[gids[k]].snap_write = G.snap + (size_t)half[k] * G.n;
        // queries still reading this half must finish first
        e = hipStreamWaitEvent(stream, G.read_ev[half[k]], 0);
        if (e != hipSuccess) return hip_fail(e, "snapshot wait");
    }
    e = graphs.upload(stream, 0, ng);
    if (e == hipSuccess) e = set_active(gids);
    LaunchCtx L = ctx(true);
    if (e == hipSuccess) e = timed("snapshot", [&] { return launch_snapshot(L); });
    if (e != hipSuccess) return hip_fail(e, "snapshot");
    std::lock_guard<std::mutex> lk(qmu);
    for (size_t k = 0; k < gids.size(); k++) {
        GraphState& G = gs[gids[k]];
        e = hipEventRecord(G.pub_ev[half[k]], stream);
        if (e == hipSuccess && G.history && G.hist_count < G.hist_cap) {
            e = hipMemcpyAsync(G.history + (size_t)G.hist_count * G.n, G.snap + (size_t)half[k] * G.n,
                               (size_t)G.n * 2, hipMemcpyDeviceToDevice, stream);
            G.hist_count++;
        }
        if (e != hipSuccess) return hip_fail(e, "publish");
        G.published = half[k];
        G.st.snapshot_published = 1;
    }
    return SGC_OK;
}

int sgc_pool::max_colors(const std::vector<uint32_t>& gids) {
    hipError_t e = set_active(gids);
    LaunchCtx L = ctx(true);
    if (e == hipSuccess) e = launch_zero_ctr(L, C_MAXC);
    if (e == hipSuccess) e = launch_max_color(L);
    if (e == hipSuccess) e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "max colour");
    for (uint32_t g : gids) gs[g].st.max_color = ctr(g, C_MAXC);
    return SGC_OK;
}

// One epoch step for a set of graphs, all in the same launches.  Reference
// order (A/src/streamgc.cu:949-1036): publish, compact, drain, rebuild.
// Repair-every-epoch order (README.md:170-176): compact, drain, Tier-1 (whose
// first step is the rebuild), Tier-2, publish.  finish = drain_and_compact +
// run_conflict_repair + cpu_serial_repair (A/src/streamgc.cu:1243-1423) and a
// publish.  Statistics (out != null) cost a few extra synchronisations.
int sgc_pool::epoch(const std::vector<uint32_t>& gids, bool finish, std::vector<sgc_epoch_stats>* out) {
    if (gids.empty()) return SGC_OK;
    for (uint32_t g : gids) {
        gs[g].st = sgc_epoch_stats();
        gs[g].st.epoch_index = gs[g].epochs;
        if (!finish) gs[g].epochs++;
        gs[g].since_epoch = 0;
    }
    int r;
    hipError_t e = set_active(gids);
    LaunchCtx L = ctx(true);
    const bool repair = finish || repair_every_epoch();
    if (!repair) {
        if ((r = publish(gids))) return r;
        e = set_active(gids);
        L = ctx(true);
    }
    LaunchCtx Lc = L;
    if (e == hipSuccess) e = sweep(Lc, M_ROW);
    if (e == hipSuccess) e = timed("compact", [&] { return launch_compact(Lc, true); });
    uint32_t ovf_cap = 0;
    for (uint32_t g : gids) ovf_cap = std::max(ovf_cap, gs[g].cfg.overflow_capacity);
    if (e == hipSuccess) e = timed("drain", [&] { return launch_drain(L, deterministic(), ovf_cap, true); });
    if (e != hipSuccess) return hip_fail(e, "compact/drain");
    if (out)   // drain statistics (the repair below reads the counters it needs itself)
        if ((r = read_drain(gids))) return r;
    if (repair) {
        if (finish && rebuild_each_sweep()) {
            e = rebuild(L);
            if (e != hipSuccess) return hip_fail(e, "rebuild");
        }
        if ((r = tier1(gids))) return r;
        if ((r = tier2(gids))) return r;
        if (!deterministic()) {   // bitmaps and histograms exact again for the next updates
            e = set_active(gids);
            L = ctx(true);
            if (e == hipSuccess) e = rebuild(L);
            if (e != hipSuccess) return hip_fail(e, "rebuild");
        }
        if ((r = publish(gids))) return r;
        for (uint32_t g : gids) gs[g].settled = true;
    } else {
        e = rebuild(L);
        if (e != hipSuccess) return hip_fail(e, "rebuild");
    }
    if (out) {
        if ((r = max_colors(gids))) return r;
        out->clear();
        for (uint32_t g : gids) out->push_back(gs[g].st);
    }
    return SGC_OK;
}

void sgc_pool::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (qstream) (void)hipStreamSynchronize(qstream);
    batch_drop();
    for (auto& G : gs) G.release();
    graphs.release();
    vbase.release();
    act.release();
    vtile.release();
    gbase.release();
    perm.release();
    seg.release();
    for (void* p : {(void*)d_ctr, (void*)d_totals, (void*)d_batch, (void*)d_tmp16, (void*)d_qids, (void*)d_qout,
                    (void*)eg.keys, (void*)eg.head, (void*)eg.next, (void*)eg.heads, (void*)eg.nheads, (void*)eg.pack,
                    (void*)d_list, (void*)d_list_count,
                    (void*)d_t2_items, d_rev, (void*)d_rev_count})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)h_ctr, (void*)h_totals})
        if (p) (void)hipHostFree(p);
    for (auto& p : pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t ev : free_events) (void)hipEventDestroy(ev);
    if (batch_up) (void)hipEventDestroy(batch_up);
    batch_up = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    if (qstream) (void)hipStreamDestroy(qstream);
    stream = qstream = nullptr;
}

// ============================================================== batched runs
void sgc_pool::batch_drop() {
    if (!batch) return;
    if (stream) (void)hipStreamSynchronize(stream);
    batch->release();
    delete batch;
    batch = nullptr;
}

// The round plan.  Round r gives every graph with records left its next
// min(E_g - pending, left) records (E_g = its epoch size), laid out graph by
// graph in stream order; every graph of the round then takes its epoch step.
int sgc_pool::batch_prepare(const sgc_edge_update* upd, uint64_t count, const sgc_batch_config& c,
                            uint32_t* rounds_out) {
    batch_drop();
    const int rc = batch_plan(upd, count, c, rounds_out);
    if (rc) batch_drop();
    return rc;
}

int sgc_pool::batch_plan(const sgc_edge_update* upd, uint64_t count, const sgc_batch_config& c,
                         uint32_t* rounds_out) {
    if (!repair_every_epoch() || !tracking())
        return fail(SGC_ERR_INVALID, "batch: needs SGC_REPAIR_EVERY_EPOCH and dirty tracking");
    if (ng == 0) return fail(SGC_ERR_INVALID, "batch: the pool holds no graph");
    if (count >= 0xFFFFFFFFull) return fail(SGC_ERR_RANGE, "batch: stream longer than 2^32 - 1 records");
    std::vector<std::vector<uint32_t>> q(ng);
    for (uint64_t i = 0; i < count; i++) {
        uint32_t gid;
        memcpy(&gid, upd[i].pad + 3, 4);
        if (gid >= ng || !gs[gid].used) return fail(SGC_ERR_INVALID, "batch: unknown graph id");
        q[gid].push_back((uint32_t)i);
    }
    BatchRun* B = new BatchRun();
    batch = B;
    B->cfg = c;
    B->ng = ng;
    B->multi = ng > 1;
    B->updates = count;
    std::vector<sgc_edge_update> ordered;
    ordered.reserve(count);
    std::vector<uint32_t> seg;
    std::vector<size_t> head(ng, 0);
    std::vector<uint64_t> since(ng, 0);
    for (uint32_t g = 0; g < ng; g++) since[g] = gs[g].used ? gs[g].since_epoch : 0;
    for (;;) {
        const uint32_t off = (uint32_t)ordered.size();
        std::vector<uint32_t> in;
        const size_t at = B->inround.size();
        B->inround.resize(at + ng, 0);
        B->full.resize(at + ng, 0);
        B->counts.resize(at + ng, 0);
        seg.resize(2 * (at + ng), 0);
        for (uint32_t g = 0; g < ng; g++) {
            if (!gs[g].used || head[g] == q[g].size()) continue;
            const uint64_t E = gs[g].cfg.epoch_size;
            const uint64_t room = E - std::min<uint64_t>(since[g], E - 1);
            const uint64_t take = std::min<uint64_t>(q[g].size() - head[g], room);
            seg[2 * (at + g)] = (uint32_t)ordered.size();
            for (uint64_t k = 0; k < take; k++) ordered.push_back(upd[q[g][head[g] + k]]);
            seg[2 * (at + g) + 1] = (uint32_t)ordered.size();
            head[g] += take;
            B->inround[at + g] = 1;
            B->full[at + g] = since[g] + take >= E;
            B->counts[at + g] = (uint32_t)take;
            since[g] = 0;
            in.push_back(g);
        }
        if (in.empty()) {
            B->inround.resize(at);
            B->full.resize(at);
            B->counts.resize(at);
            seg.resize(2 * at);
            break;
        }
        B->tab.push_back(off);
        B->tab.push_back((uint32_t)ordered.size() - off);
        B->max_count = std::max(B->max_count, (uint32_t)ordered.size() - off);
        B->gids.push_back(in);
        B->rounds++;
    }
    if ((uint64_t)B->rounds * 2 * ng >= 0xFFFFFFFFull) return fail(SGC_ERR_RANGE, "batch: round table too large");
    if (!deterministic() && (B->max_count > eg.next_cap || 2ull * B->max_count > eg.size))
        return fail(SGC_ERR_INVALID, "batch: a round holds more records than the pool's batch capacity (" +
                                         std::to_string(B->max_count) + " > " + std::to_string(batch_cap) + ")");
    uint32_t max_iters = 0;
    std::vector<RoundGraph> rg(ng);
    std::vector<uint16_t*> base(ng, nullptr);
    for (uint32_t g = 0; g < ng; g++) {
        if (!gs[g].used) continue;
        rg[g].max_iters = gs[g].cfg.max_repair_iters;
        rg[g].plateau_iters = gs[g].cfg.plateau_iters;
        rg[g].n = gs[g].n;
        rg[g].t2_passes = gs[g].cfg.tier2_passes;
        base[g] = gs[g].snap;
        max_iters = std::max(max_iters, gs[g].cfg.max_repair_iters);
        B->ovf_cap = std::max(B->ovf_cap, gs[g].cfg.overflow_capacity);
        B->any_history |= gs[g].history != nullptr;
        B->all_settled &= gs[g].settled;
    }
    // By default the rounds start with two captured sweeps (most rounds need
    // one or two); each round that needs more is finished by the host and
    // doubles the sweeps captured for the rounds left, up to the graphs'
    // iteration cap (a slot costs a few launches, a host-finished round a
    // synchronisation and a capture).
    B->sweeps_max = c.t1_sweeps ? std::min(c.t1_sweeps, max_iters) : max_iters;
    B->sweeps = c.t1_sweeps ? B->sweeps_max : std::min(2u, B->sweeps_max);
    // Tier-2 passes captured per round (host placement or no dirty tracking: none, the host repairs)
    uint32_t max_passes = 0;
    for (uint32_t g = 0; g < ng; g++)
        if (gs[g].used) max_passes = std::max(max_passes, gs[g].cfg.tier2_passes);
    B->t2_passes = (flags & SGC_HOST_TIER2) ? 0 : std::min(c.t2_passes ? c.t2_passes : 4u, max_passes);
    // No graph has ever had an edge in one row only: the Tier-2 check can only
    // be needed after a Tier-1 loop that ended unclean or a new overflow, so
    // the rounds are captured without it and stop for the host when it is.
    {
        std::vector<uint32_t> hc((size_t)ng * C_COUNT);
        hipError_t ec = hipMemcpyAsync(hc.data(), d_ctr, hc.size() * 4, hipMemcpyDeviceToHost, stream);
        if (ec == hipSuccess) ec = hipStreamSynchronize(stream);
        if (ec != hipSuccess) return hip_fail(ec, "batch: counters");
        B->lean = true;
        for (uint32_t g = 0; g < ng; g++)
            if (gs[g].used && hc[(size_t)g * C_COUNT + C_ASYM] != 0) B->lean = false;
    }
    hipError_t e = hipMalloc((void**)&B->d_stream, std::max<size_t>(ordered.size(), 1) * sizeof(sgc_edge_update));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_tab, std::max<size_t>(B->tab.size(), 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_seg, std::max<size_t>(seg.size(), 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_inround, std::max<size_t>(B->inround.size(), 1));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_ctl, RC_WORDS * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_fl, (size_t)F_KINDS * ng * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_pub, (size_t)ng * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_space, (size_t)dspace_words(ng) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_state, (size_t)ng * sizeof(RoundState));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_rg, (size_t)ng * sizeof(RoundGraph));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_snapbase, (size_t)ng * sizeof(uint16_t*));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_hist, (size_t)ng * sizeof(HistDev));
    if (e == hipSuccess) e = hipHostMalloc((void**)&B->h_ctl, RC_WORDS * 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&B->h_rd, 2 * RC_WORDS * 4);
    for (hipEvent_t& x : B->rd_ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
    auto up = [&](void* d, const void* h, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream);
    };
    up(B->d_stream, ordered.data(), ordered.size() * sizeof(sgc_edge_update));
    up(B->d_tab, B->tab.data(), B->tab.size() * 4);
    up(B->d_seg, seg.data(), seg.size() * 4);
    up(B->d_inround, B->inround.data(), B->inround.size());
    up(B->d_rg, rg.data(), rg.size() * sizeof(RoundGraph));
    up(B->d_snapbase, base.data(), base.size() * sizeof(uint16_t*));
    if (e == hipSuccess) e = hipMemsetAsync(B->d_fl, 0, (size_t)F_KINDS * ng * 4, stream);
    if (e == hipSuccess) e = hipMemsetAsync(B->d_state, 0, (size_t)ng * sizeof(RoundState), stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "batch prepare");
    if (c.use_graph && B->t2_passes) {
        int r = ensure_rev_scratch();
        if (r) return r;
    }
    // launch-shape candidates: the pool's shapes, then one knob changed at a time
    B->cands.push_back(tune);
    if (c.use_graph && c.autotune_rounds) {
        sgc_tuning t = tune;
        t.long_row = tune.long_row > 256 ? 256 : 4096;
        B->cands.push_back(t);
        t = tune;
        t.sweep_blocks = tune.sweep_blocks > 512 ? 512 : 4096;
        B->cands.push_back(t);
        t = tune;
        t.apply_blocks = tune.apply_blocks > 4096 ? 4096 : 16384;
        B->cands.push_back(t);
    }
    if (c.use_graph) {
        for (const sgc_tuning& t : B->cands) {
            hipGraphExec_t x = nullptr;
            int r = batch_capture(t, &x);
            if (r) return r;
            B->execs.push_back(x);
        }
    }
    if (rounds_out) *rounds_out = B->rounds;
    return SGC_OK;
}

// One round, captured: apply the round's slice, compact, drain, the Tier-1
// sweeps with device-side control, the Tier-2 check, rebuild, publish.  Every
// launch takes its per-round inputs from device memory, so one instantiated
// graph serves every round.
int sgc_pool::batch_capture(const sgc_tuning& t, hipGraphExec_t* out) {
    BatchRun& B = *batch;
    const bool det = deterministic();
    LaunchCtx base = ctx(false);
    base.max_blocks = t.apply_blocks;
    base.sweep_blocks = t.sweep_blocks;
    base.long_row = t.long_row;
    base.list = d_list;
    base.list_count = d_list_count;
    base.list_cap = list_cap;
    auto F = [&](uint32_t k) { return (const uint32_t*)(B.d_fl + (size_t)k * ng); };
    const uint32_t* t1_any = B.d_ctl + RC_T1_ANY;
    // graphs not yet repaired once take the reference's selection (fresh_level)
    const int fresh = B.all_settled ? (int)t.fresh_repair : 0;
    const bool rebuild_sweep = det || fresh < 2;
    bool serial_ok = false;   // some graph is small enough for serial sweeps
    for (uint32_t g = 0; g < ng; g++) serial_ok |= gs[g].used && gs[g].n <= t.serial_repair;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "batch capture");
    e = launch_round_begin(stream, B.d_ctl, B.d_tab, B.d_inround, ng, B.d_fl, B.d_state, B.d_rg);
    LaunchCtx La = base;
    La.dyn = B.d_ctl + RC_OFF;
    if (e == hipSuccess) {
        if (det) {
            e = launch_apply_seg(La, B.d_stream, B.d_seg, ng, nullptr, B.multi ? 1 : 0);
        } else {
            e = launch_group(La, B.d_stream, B.max_count, B.multi ? 1 : 0, &eg);
            if (e == hipSuccess) e = launch_apply(La, B.d_stream, B.max_count, B.multi ? 1 : 0, &eg);
        }
    }
    LaunchCtx Lr = base;
    Lr.active = F(F_ROUND);
    if (e == hipSuccess) e = launch_build_list(Lr, M_ROW, d_list, list_cap, d_list_count);
    if (e == hipSuccess) e = launch_compact(Lr, true);
    if (e == hipSuccess) e = launch_drain(Lr, det, B.ovf_cap, true);
    for (uint32_t it = 0; it < B.sweeps && e == hipSuccess; it++) {
        e = launch_t1_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_state, B.d_rg, it > 0, true, det,
                              t.serial_repair, t.serial_conflicts);
        if (e == hipSuccess && rebuild_sweep) {
            LaunchCtx Lb = base;
            Lb.active = F(F_ITER);
            Lb.gate = t1_any;
            e = launch_build_list(Lb, M_BITS, d_list, list_cap, d_list_count);
            if (e == hipSuccess) e = launch_rebuild(Lb, exact_hist());
        }
        if (e == hipSuccess && !det) {
            LaunchCtx Lp = base;
            Lp.active = F(F_PAR);
            Lp.gate = t1_any;
            e = launch_build_list(Lp, M_CHECK, d_list, list_cap, d_list_count);
            if (e == hipSuccess) e = launch_tier1(Lp, false, fresh);
        }
        if (e == hipSuccess && (det || (it >= 2 && serial_ok))) {   // parallel runs: serial after two sweeps
            LaunchCtx Ls = base;
            Ls.active = F(F_SER);
            e = launch_tier1(Ls, true, det ? 0 : fresh);
        }
    }
    if (e == hipSuccess)   // the last sweep's counts, the Tier-2 check's graphs (lean: and the round's end)
        e = launch_t1_finish(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_state, B.d_rg, det, t.serial_repair,
                             t.serial_conflicts, B.lean);
    LaunchCtx Lt = base;
    Lt.active = F(F_T2);
    if (e == hipSuccess && !B.lean) e = launch_build_list(Lt, M_T2, d_list, list_cap, d_list_count);
    if (e == hipSuccess && !B.lean) e = launch_t2_check(Lt);
    if (B.t2_passes && !B.lean) {   // Tier-2 on the device for the graphs whose check found bad slots
        LaunchCtx L2 = base;   // the graphs with bad slots, laid out by k_t2_control
        L2.active = F(F_T2RUN);
        L2.gate = B.d_ctl + RC_T2_ANY;
        L2.dspace = B.d_space;
        if (e == hipSuccess) e = launch_t2_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_rg, 0, 0, B.d_space);
        if (e == hipSuccess) e = launch_t2_begin(L2, d_rev, d_rev_count, rev_cap);
        if (e == hipSuccess) e = launch_t2_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_rg, 1, 0, B.d_space);
        if (e == hipSuccess) e = launch_t2_pass(L2, d_rev, d_rev_count, rev_cap, true);
        for (uint32_t p = 0; p < B.t2_passes && e == hipSuccess; p++) {
            e = launch_t2_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_rg, 2, p, B.d_space);
            if (e == hipSuccess) e = launch_t2_pass(L2, d_rev, d_rev_count, rev_cap, false);
            if (e == hipSuccess)
                e = launch_t2_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_rg, 3, p, B.d_space);
        }
        if (e == hipSuccess) e = launch_t2_control(stream, B.d_ctl, graphs.d, ng, B.d_fl, B.d_rg, 4, 0, B.d_space);
        LaunchCtx Lc = base;   // the remaining bad slots, as the host's check after a repair counts them
        Lc.active = F(F_T2RAN);
        Lc.gate = B.d_ctl + RC_T2_RAN;
        if (e == hipSuccess) e = launch_build_list(Lc, M_T2, d_list, list_cap, d_list_count);
        if (e == hipSuccess) e = launch_t2_check(Lc);
    }
    if (e == hipSuccess && !B.lean) e = launch_round_end(stream, B.d_ctl, graphs.d, ng, B.d_fl);
    if (e == hipSuccess && !det) {   // bitmaps and histograms exact again for the next updates
        e = launch_build_list(Lr, M_BITS, d_list, list_cap, d_list_count);
        if (e == hipSuccess) e = launch_rebuild(Lr, exact_hist());
    }
    // the snapshot also fills each round graph's next history slot
    if (e == hipSuccess) e = launch_snapshot(Lr, B.any_history ? F(F_ROUND) : nullptr, B.d_hist);
    if (e == hipSuccess)
        e = launch_round_publish(stream, graphs.d, ng, F(F_ROUND), B.d_snapbase, B.d_pub, B.d_hist);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "batch capture");
    return SGC_OK;
}

// Host state -> device: each graph's next snapshot buffer, its published
// half and its history slot.
hipError_t sgc_pool::batch_push(BatchRun& B) {
    hipError_t e = graphs.writable();
    if (e != hipSuccess) return e;
    std::vector<uint32_t> pub(ng, 0xFFFFFFFFu);
    std::vector<HistDev> hist(ng);
    for (uint32_t g = 0; g < ng; g++) {
        hist[g] = HistDev{nullptr, 0, 0};
        if (!gs[g].used) continue;
        const GraphState& G = gs[g];
        graphs.h[g].snap_write = G.snap + (size_t)(G.published == 0 ? 1 : 0) * G.n;
        if (G.published >= 0) pub[g] = (uint32_t)G.published;
        hist[g] = HistDev{G.history, G.hist_cap, G.hist_count};
    }
    e = graphs.upload(stream, 0, ng);
    if (e == hipSuccess) e = hipMemcpyAsync(B.d_pub, pub.data(), (size_t)ng * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(B.d_hist, hist.data(), (size_t)ng * sizeof(HistDev), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(B.d_ctl, B.h_ctl, RC_WORDS * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

// Device state -> host: published halves and history counts.
hipError_t sgc_pool::batch_pull(BatchRun& B) {
    std::vector<uint32_t> pub(ng);
    std::vector<HistDev> hist(ng);
    hipError_t e = hipMemcpyAsync(pub.data(), B.d_pub, (size_t)ng * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hist.data(), B.d_hist, (size_t)ng * sizeof(HistDev), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    for (uint32_t g = 0; g < ng; g++) {
        if (!gs[g].used) continue;
        if (pub[g] != 0xFFFFFFFFu) gs[g].published = (int)pub[g];
        if (gs[g].history) gs[g].hist_count = hist[g].count;
    }
    return hipSuccess;
}

// Finish round r on the host from where the captured round stopped: more
// Tier-1 sweeps for graphs still iterating, the Tier-2 check and repair,
// rebuild, publish.
int sgc_pool::batch_host_round(BatchRun& B, uint32_t r) {
    std::vector<RoundState> st(ng);
    hipError_t e = hipMemcpyAsync(st.data(), B.d_state, (size_t)ng * sizeof(RoundState), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = batch_pull(B);
    if (e != hipSuccess) return hip_fail(e, "batch state");
    const std::vector<uint32_t>& gids = B.gids[r];
    bool resume = false;
    for (uint32_t g : gids) {
        GraphState& G = gs[g];
        G.iterating = st[g].iterating != 0;
        G.prev = st[g].prev;
        G.plateau = st[g].plateau;
        G.t1_clean = st[g].clean != 0;
        G.st.t1_iterations = st[g].iters;
        G.st.t1_conflicts_last = st[g].last;
        resume |= G.iterating;
    }
    int rc;
    if (resume && (rc = tier1(gids, B.sweeps))) return rc;
    e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "counters");
    if ((rc = tier2(gids))) return rc;
    if (!deterministic()) {
        e = set_active(gids);
        if (e == hipSuccess) e = rebuild(ctx(true));
        if (e != hipSuccess) return hip_fail(e, "rebuild");
    }
    if ((rc = publish(gids))) return rc;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "batch round");
    return SGC_OK;
}

// The same rounds with host-driven launches (no capture).
int sgc_pool::batch_run_host(BatchRun& B) {
    for (uint32_t r = 0; r < B.rounds; r++) {
        LaunchCtx L = ctx(false);
        const uint32_t off = B.tab[2 * r], cnt = B.tab[2 * r + 1];
        hipError_t e;
        if (deterministic()) {
            e = timed("apply", [&] {
                return launch_apply_seg(L, B.d_stream, B.d_seg + (size_t)r * 2 * ng, ng, nullptr, B.multi ? 1 : 0);
            });
        } else {
            e = timed("group", [&] { return launch_group(L, B.d_stream + off, cnt, B.multi ? 1 : 0, &eg); });
            if (e == hipSuccess)
                e = timed("apply", [&] { return launch_apply(L, B.d_stream + off, cnt, B.multi ? 1 : 0, &eg); });
        }
        if (e != hipSuccess) return hip_fail(e, "batch apply");
        for (uint32_t g : B.gids[r]) gs[g].since_epoch += B.counts[(size_t)r * ng + g];
        int rc = epoch(B.gids[r], false, nullptr);
        if (rc) return rc;
        for (uint32_t g : B.gids[r])
            if (!B.full[(size_t)r * ng + g]) gs[g].epochs--;
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "batch run");
    resolve_timing();
    return SGC_OK;
}

// Rounds per replay chunk once the autotuner has chosen (two chunks in flight).
constexpr uint32_t ROUND_CHUNK = 2;

int sgc_pool::batch_run(sgc_batch_stats* out) {
    if (!batch) return fail(SGC_ERR_INVALID, "batch run: nothing prepared");
    BatchRun& B = *batch;
    if (B.ng != ng) return fail(SGC_ERR_INVALID, "batch run: graphs were added or removed since the plan was made");
    sgc_batch_stats S;
    memset(&S, 0, sizeof S);
    S.rounds = B.rounds;
    S.updates = B.updates;
    std::lock_guard<std::mutex> run(run_mu);   // queries wait for the run
    hipError_t e = hipStreamSynchronize(qstream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "batch run");
    auto t0 = std::chrono::steady_clock::now();
    if (!B.cfg.use_graph) {
        int rc = batch_run_host(B);
        if (rc) return rc;
    } else {
        memset(B.h_ctl, 0, RC_WORDS * 4);
        B.h_ctl[RC_ROUNDS] = B.rounds;
        e = batch_push(B);
        if (e != hipSuccess) return hip_fail(e, "batch run");
        // Autotuning.  Every candidate is timed (HIP events around its replay)
        // on cfg.autotune_rounds rounds that ran captured to the end: a round
        // the host finished gives no sample, and a recapture (the captured
        // body changed) discards the samples taken so far.  Each tuning
        // replay gives one round to every candidate still short of samples,
        // fewest samples first.  The fastest per update is kept for the rest
        // of the run; the tuning rounds are reported apart (tune_*).
        const uint32_t C = (uint32_t)B.execs.size();
        const uint32_t need = C > 1 ? B.cfg.autotune_rounds : 0;
        std::vector<hipEvent_t> ev(2 * (size_t)C, nullptr);
        for (auto& x : ev)
            if (e == hipSuccess) e = hipEventCreate(&x);
        std::vector<double> cand_ms(C, 0.0);
        std::vector<uint64_t> cand_upd(C, 0);
        std::vector<uint32_t> cand_n(C, 0), who;
        uint32_t best = 0, r = 0;
        bool tuned = need == 0;
        int rc = SGC_OK;
        while (e == hipSuccess && rc == SGC_OK && r < B.rounds) {
            who.clear();
            if (!tuned) {
                for (uint32_t k = 0; k < C; k++)
                    if (cand_n[k] < need) who.push_back(k);
                std::stable_sort(who.begin(), who.end(), [&](uint32_t a, uint32_t b) { return cand_n[a] < cand_n[b]; });
            }
            RoundReading rd;
            uint32_t end = r;
            if (!tuned) {   // one round per candidate still short of samples, each between events
                end = std::min<uint32_t>(B.rounds, r + (uint32_t)who.size());
                for (uint32_t k = r; k < end && e == hipSuccess; k++) {
                    e = hipEventRecord(ev[2 * (k - r)], stream);
                    if (e == hipSuccess) e = hipGraphLaunch(B.execs[who[k - r]], stream);
                    if (e == hipSuccess) e = hipEventRecord(ev[2 * (k - r) + 1], stream);
                }
                if (e == hipSuccess) e = hipMemcpyAsync(B.h_ctl, B.d_ctl, RC_WORDS * 4, hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) break;
                // the replays either all ran or stopped at one round in [r, end);
                // any other reading ends the run (each accepted one moves r on)
                if (!read_rounds(r, end, B.rounds, B.h_ctl[RC_ROUND], B.h_ctl[RC_STOP], B.h_ctl[RC_STOP_ROUND], &rd)) {
                    rc = fail(SGC_ERR_STATE, "batch run: inconsistent round control after replays of rounds " +
                                                 std::to_string(r) + ".." + std::to_string(end));
                    break;
                }
            } else {
                // The chosen body in chunks of ROUND_CHUNK rounds, two chunks in
                // flight: each chunk's control block comes back while the next
                // one runs, so the GPU never waits for the host, and a round
                // that stops costs at most the rest of its chunk and the next
                // chunk as replays that find the stop and end at once.
                uint32_t c_end[2] = {0, 0}, next = r;
                auto chunk = [&](int slot) {
                    c_end[slot] = std::min<uint32_t>(B.rounds, next + ROUND_CHUNK);
                    for (uint32_t k = next; k < c_end[slot] && e == hipSuccess; k++)
                        e = hipGraphLaunch(B.execs[best], stream);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(B.h_rd + slot * RC_WORDS, B.d_ctl, RC_WORDS * 4, hipMemcpyDeviceToHost,
                                           stream);
                    if (e == hipSuccess) e = hipEventRecord(B.rd_ev[slot], stream);
                    next = c_end[slot];
                };
                chunk(0);
                if (next < B.rounds) chunk(1);
                int slot = 0;
                bool ok = true;
                while (e == hipSuccess) {
                    e = hipEventSynchronize(B.rd_ev[slot]);
                    if (e != hipSuccess) break;
                    const uint32_t* c = B.h_rd + slot * RC_WORDS;
                    end = c_end[slot];
                    ok = read_rounds(r, end, B.rounds, c[RC_ROUND], c[RC_STOP], c[RC_STOP_ROUND], &rd);
                    if (!ok || rd.stopped || end >= B.rounds) {
                        // the chunk queued behind a stop finds it and ends; the host
                        // round below reads the control block as the stop left it
                        memcpy(B.h_ctl, c, RC_WORDS * 4);
                        e = hipStreamSynchronize(stream);
                        break;
                    }
                    r = end;
                    if (next < B.rounds) chunk(slot);   // this slot's reading is consumed
                    slot ^= 1;
                }
                if (e != hipSuccess) break;
                if (!ok) {
                    rc = fail(SGC_ERR_STATE, "batch run: inconsistent round control after replays of rounds " +
                                                 std::to_string(r) + ".." + std::to_string(end));
                    break;
                }
            }
            const bool stopped = rd.stopped;
            const uint32_t done = rd.done;
            if (!tuned)
                for (uint32_t k = r; k < done; k++) {   // rounds that ran captured to the end
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, ev[2 * (k - r)], ev[2 * (k - r) + 1]);
                    const uint32_t q = who[k - r];
                    cand_ms[q] += ms;
                    cand_upd[q] += B.tab[2 * k + 1];
                    cand_n[q]++;
                }
            if (stopped) {
                const uint32_t s = rd.host;
                S.host_rounds++;
                rc = batch_host_round(B, s);
                if (rc) break;
                const bool more_t2 = B.h_ctl[RC_T2_MISSING] && B.lean;   // the Tier-2 check is needed now
                const bool more_t1 = B.h_ctl[RC_T1_SHORT] && B.sweeps < B.sweeps_max;   // more sweeps too
                if (more_t2 || more_t1) {   // capture them for the rounds left
                    if (more_t2) B.lean = false;
                    if (more_t1) B.sweeps = std::min(2 * B.sweeps, B.sweeps_max);
                    for (uint32_t k = 0; k < C && rc == SGC_OK; k++) {
                        if (tuned && k != best) continue;   // only the chosen body replays from here on
                        hipGraphExec_t x = nullptr;
                        rc = batch_capture(B.cands[k], &x);
                        if (rc == SGC_OK) {
                            (void)hipGraphExecDestroy(B.execs[k]);
                            B.execs[k] = x;
                        }
                    }
                    if (rc) break;
                    if (!tuned) {   // samples of the old bodies no longer apply
                        std::fill(cand_ms.begin(), cand_ms.end(), 0.0);
                        std::fill(cand_upd.begin(), cand_upd.end(), 0);
                        std::fill(cand_n.begin(), cand_n.end(), 0);
                    }
                }
                r = rd.next;
                B.h_ctl[RC_ROUND] = r;
                B.h_ctl[RC_STOP] = 0;
                B.h_ctl[RC_T2_MISSING] = 0;
                B.h_ctl[RC_T1_SHORT] = 0;
                e = batch_push(B);
            } else {
                r = end;
            }
            if (!tuned) {
                bool all = true;
                for (uint32_t k = 0; k < C; k++) all = all && cand_n[k] >= need;
                if (all || r >= B.rounds) {
                    tuned = true;
                    double best_v = 0;
                    for (uint32_t k = 0; k < C; k++) {
                        if (!cand_n[k]) continue;
                        const double v = cand_ms[k] / (double)std::max<uint64_t>(cand_upd[k], 1);
                        if (best_v == 0 || v < best_v) {
                            best_v = v;
                            best = k;
                        }
                    }
                    tune = B.cands[best];
                    S.tune_rounds = r;
                    for (uint32_t k = 0; k < r; k++) S.tune_updates += B.tab[2 * k + 1];
                    S.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                }
            }
        }
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
        if (rc) return rc;
        if (e == hipSuccess) e = batch_pull(B);
        for (uint32_t g = 0; g < ng && e == hipSuccess; g++)
            if (gs[g].used && gs[g].published >= 0) e = hipEventRecord(gs[g].pub_ev[gs[g].published], stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "batch run");
        for (uint32_t r2 = 0; r2 < B.rounds; r2++)
            for (uint32_t g : B.gids[r2]) gs[g].epochs += B.full[(size_t)r2 * ng + g];
        for (uint32_t g = 0; g < ng; g++)
            if (gs[g].used) gs[g].since_epoch = 0;
        for (uint32_t r2 = 0; r2 < B.rounds; r2++)
            for (uint32_t g : B.gids[r2]) gs[g].settled = true;
        S.candidates = C > 1 ? C : 0;
        S.chosen = best;
        for (uint32_t k = 0; k < C && k < SGC_BATCH_CANDIDATES; k++) {
            S.candidate[k] = B.cands[k];
            S.candidate_ns_per_update[k] = cand_upd[k] ? cand_ms[k] * 1e6 / (double)cand_upd[k] : 0.0;
            S.candidate_samples[k] = cand_n[k];
        }
    }
    S.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (out) *out = S;
    return SGC_OK;
}

// The single-graph engine is a pool of one.
struct sgc_engine : sgc_pool {};

namespace {

__global__ void k_gather16(const uint16_t* __restrict__ src, uint32_t n, const uint32_t* __restrict__ ids,
                           uint32_t count, uint16_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t v = ids[i];
        out[i] = v < n ? src[v] : 0;
    }
}

int pool_init(sgc_pool* P, uint32_t max_graphs, uint32_t batch_cap, uint32_t flags, uint64_t budget) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return P->fail(SGC_ERR_HIP, "no HIP device");
    if (hipGetDevice(&P->device) != hipSuccess) return P->fail(SGC_ERR_HIP, "no current HIP device");
    P->max_graphs = max_graphs;
    P->batch_cap = batch_cap;
    P->flags = flags;
    P->budget = budget;
    P->gs.resize(max_graphs);
    hipError_t e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->qstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&P->batch_up, hipEventDisableTiming);
    if (e == hipSuccess) e = P->graphs.alloc(max_graphs);
    if (e == hipSuccess) e = P->vbase.alloc(max_graphs + 1);
    if (e == hipSuccess) e = P->act.alloc(4 * (size_t)max_graphs + 2);
    if (e == hipSuccess) e = P->vtile.alloc(max_graphs + 1);
    if (e == hipSuccess) e = P->gbase.alloc(max_graphs + 1);
    if (e == hipSuccess) e = P->perm.alloc(batch_cap);
    if (e == hipSuccess) e = P->seg.alloc(2 * (size_t)max_graphs);
    if (e == hipSuccess) e = hipMalloc(&P->d_ctr, (size_t)max_graphs * C_COUNT * 4);
    if (e == hipSuccess) e = hipMemsetAsync(P->d_ctr, 0, (size_t)max_graphs * C_COUNT * 4, P->stream);
    if (e == hipSuccess) e = hipHostMalloc(&P->h_ctr, (size_t)max_graphs * C_COUNT * 4);
    if (e == hipSuccess) e = hipMalloc(&P->d_totals, (size_t)T_STRIPES * T_STRIDE * 8);
    if (e == hipSuccess) e = hipMemsetAsync(P->d_totals, 0, (size_t)T_STRIPES * T_STRIDE * 8, P->stream);
    if (e == hipSuccess) e = hipHostMalloc(&P->h_totals, (size_t)T_STRIPES * T_STRIDE * 8);
    if (e == hipSuccess) e = hipMalloc(&P->d_batch, (size_t)batch_cap * sizeof(sgc_edge_update));
    if (e == hipSuccess) e = hipMalloc(&P->d_tmp16, 64);
    if (e == hipSuccess) e = hipMalloc(&P->d_list_count, 8);
    if (e == hipSuccess) e = hipMalloc(&P->d_rev_count, 8);
    P->eg.size = 16;
    while (P->eg.size < 2u * batch_cap) P->eg.size <<= 1;
    P->eg.next_cap = batch_cap;
    if (e == hipSuccess) e = hipMalloc(&P->eg.keys, (size_t)P->eg.size * 8);
    if (e == hipSuccess) e = hipMalloc(&P->eg.head, (size_t)P->eg.size * 4);
    if (e == hipSuccess) e = hipMalloc(&P->eg.next, (size_t)batch_cap * 4);
    if (e == hipSuccess) e = hipMalloc(&P->eg.heads, (size_t)batch_cap * 4);
    if (e == hipSuccess) e = hipMalloc(&P->eg.nheads, 8);
    if (e == hipSuccess) e = hipMalloc(&P->eg.pack, (size_t)batch_cap * 4 * sizeof(uint4));
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
    if (e != hipSuccess) return P->hip_fail(e, "pool create");
    return SGC_OK;
}

int submit_impl(sgc_pool* P, const sgc_edge_update* updates, uint64_t count, int on_device,
                sgc_epoch_stats* stats_out, uint32_t max_stats, uint32_t* nstats) {
    auto t0 = std::chrono::steady_clock::now();
    P->last_ms[2] = 0;
    uint32_t produced = 0;
    uint64_t done = 0;
    // batch size of a pool = the smallest batch_size among its graphs (one batch for a graph of one)
    uint32_t bsz = P->batch_cap;
    for (uint32_t g = 0; g < P->ng; g++)
        if (P->gs[g].used) bsz = std::min(bsz, P->gs[g].cfg.batch_size);
    std::vector<uint64_t> per_graph;
    std::vector<sgc_epoch_stats> st;
    while (done < count) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(bsz, count - done);
        int r = P->apply(updates + done, b, on_device != 0, per_graph);
        if (r) return r;
        done += b;
        std::vector<uint32_t> due;
        for (uint32_t g = 0; g < P->ng; g++)
            if (P->gs[g].used && P->gs[g].since_epoch >= P->gs[g].cfg.epoch_size) due.push_back(g);
        if (!due.empty()) {
            const bool want = stats_out && produced < max_stats;
            r = P->epoch(due, false, want ? &st : nullptr);
            if (r) return r;
            if (want)
                for (auto& s : st)
                    if (produced < max_stats) stats_out[produced++] = s;
        }
    }
    hipError_t e = hipStreamSynchronize(P->stream);
    if (e != hipSuccess) return P->hip_fail(e, "submit");
    P->resolve_timing();
    P->last_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (nstats) *nstats = produced;
    return SGC_OK;
}

// process_stream for a pool: the mixed stream is split per graph (stream order
// kept inside each graph) and applied in rounds.  A round takes, from every
// graph with pending updates, at most what is left until that graph's next
// epoch (E_g updates per epoch), within the pool's batch capacity; the graphs
// that reach their boundary then share one epoch.  Each graph therefore sees
// exactly the batches and epochs it would see alone with batch = epoch = E_g.
int pool_submit_impl(sgc_pool* P, const sgc_edge_update* updates, uint64_t count, sgc_epoch_stats* stats_out,
                     uint32_t max_stats, uint32_t* nstats) {
    auto t0 = std::chrono::steady_clock::now();
    P->last_ms[2] = 0;
    std::vector<std::vector<uint32_t>> queue(P->ng);
    for (uint64_t i = 0; i < count; i++) {
        uint32_t gid;
        memcpy(&gid, updates[i].pad + 3, 4);
        if (gid >= P->ng || !P->gs[gid].used) return P->fail(SGC_ERR_INVALID, "submit: unknown graph id");
        if (i > 0xFFFFFFFFull) return P->fail(SGC_ERR_RANGE, "submit: stream longer than 2^32 records");
        queue[gid].push_back((uint32_t)i);
    }
    std::vector<size_t> head(P->ng, 0);
    std::vector<sgc_edge_update> batch;
    batch.reserve(std::min<uint64_t>(count, P->batch_cap));
    std::vector<uint64_t> per_graph;
    std::vector<sgc_epoch_stats> st;
    uint32_t produced = 0;
    for (;;) {
        batch.clear();
        for (uint32_t g = 0; g < P->ng && batch.size() < P->batch_cap; g++) {
            const GraphState& G = P->gs[g];
            if (!G.used || head[g] == queue[g].size()) continue;
            const uint64_t room = G.cfg.epoch_size - std::min<uint64_t>(G.since_epoch, G.cfg.epoch_size - 1);
            uint64_t take = std::min<uint64_t>(queue[g].size() - head[g], room);
            take = std::min<uint64_t>(take, P->batch_cap - batch.size());
            for (uint64_t k = 0; k < take; k++) batch.push_back(updates[queue[g][head[g] + k]]);
            head[g] += take;
        }
        if (batch.empty()) break;
        int r = P->apply(batch.data(), (uint32_t)batch.size(), false, per_graph);
        if (r) return r;
        std::vector<uint32_t> due;
        for (uint32_t g = 0; g < P->ng; g++)
            if (P->gs[g].used && P->gs[g].since_epoch >= P->gs[g].cfg.epoch_size) due.push_back(g);
        if (!due.empty()) {
            const bool want = stats_out && produced < max_stats;
            r = P->epoch(due, false, want ? &st : nullptr);
            if (r) return r;
            if (want)
                for (auto& s : st)
                    if (produced < max_stats) stats_out[produced++] = s;
        }
    }
    hipError_t e = hipStreamSynchronize(P->stream);
    if (e != hipSuccess) return P->hip_fail(e, "submit");
    P->resolve_timing();
    P->last_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (nstats) *nstats = produced;
    return SGC_OK;
}

int download(sgc_pool* P, void* dst, const void* src, size_t bytes) {
    hipError_t r = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, P->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(P->stream);
    if (r != hipSuccess) return P->hip_fail(r, "download");
    return SGC_OK;
}
int upload(sgc_pool* P, void* dst, const void* src, size_t bytes) {
    hipError_t r = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, P->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(P->stream);
    if (r != hipSuccess) return P->hip_fail(r, "upload");
    return SGC_OK;
}

// State uploaded from the host: every sweep must visit every vertex again.
int mark_all_sync(sgc_pool* P) {
    hipError_t r = hipSuccess;
    for (uint32_t g = 0; g < P->ng && r == hipSuccess; g++)
        if (P->gs[g].used) r = P->mark_all(g);
    if (r == hipSuccess) r = hipStreamSynchronize(P->stream);
    if (r != hipSuccess) return P->hip_fail(r, "mark");
    return SGC_OK;
}

GraphState* graph_of(sgc_pool* P, uint32_t gid) {
    if (!P || gid >= P->max_graphs || !P->gs[gid].used) return nullptr;
    return &P->gs[gid];
}

// The published colours of `count` ids through device buffers d_ids (count
// u32) and d_out (count u16), on the pool's query stream after the publish.
// Caller holds run_mu and qmu.
int query_run(sgc_pool* P, GraphState* G, const uint32_t* ids, uint32_t count, uint16_t* out, uint32_t* d_ids,
              uint16_t* d_out) {
    const int half = G->published;
    hipError_t r = hipStreamWaitEvent(P->qstream, G->pub_ev[half], 0);
    if (r == hipSuccess) r = hipMemcpyAsync(d_ids, ids, (size_t)count * 4, hipMemcpyHostToDevice, P->qstream);
    if (r == hipSuccess) {
        const uint32_t grid = std::min<uint32_t>((count + 255) / 256, 1024);
        k_gather16<<<grid, 256, 0, P->qstream>>>(G->snap + (size_t)half * G->n, G->n, d_ids, count, d_out);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipEventRecord(G->read_ev[half], P->qstream);
    if (r == hipSuccess) r = hipMemcpyAsync(out, d_out, (size_t)count * 2, hipMemcpyDeviceToHost, P->qstream);
    if (r == hipSuccess) r = hipStreamSynchronize(P->qstream);
    if (r != hipSuccess) return P->hip_fail(r, "query");
    return SGC_OK;
}

int query_impl(sgc_pool* P, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out) {
    GraphState* G = graph_of(P, gid);
    if (!G || (!ids && count) || (!out && count)) return SGC_ERR_INVALID;
    if (count == 0) return SGC_OK;
    std::lock_guard<std::mutex> run(P->run_mu);
    std::lock_guard<std::mutex> lk(P->qmu);
    if (G->published < 0) {
        memset(out, 0, (size_t)count * 2);
        return SGC_OK;
    }
    if (count > P->qcap) {
        if (P->d_qids) (void)hipFree(P->d_qids);
        if (P->d_qout) (void)hipFree(P->d_qout);
        P->d_qids = nullptr;
        P->d_qout = nullptr;
        P->qcap = 0;
        hipError_t r = hipMalloc(&P->d_qids, (size_t)count * 4);
        if (r == hipSuccess) r = hipMalloc(&P->d_qout, (size_t)count * 2);
        if (r != hipSuccess) return P->hip_fail(r, "query buffers");
        P->qcap = count;
    }
    return query_run(P, G, ids, count, out, P->d_qids, P->d_qout);
}

// ---------------------------------------------------------------- allocation-free query
// Staging of sgc_pool_query_noalloc: SGC_QUERY_NOALLOC_MAX ids (u32) followed
// by as many colours (u16), in a file-scope device array that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel or a copy: it passes the device address that
// hipGetSymbolAddress returns for the device, looked up once per device.
}  // namespace

static constexpr uint32_t QSCRATCH_WORDS = SGC_QUERY_NOALLOC_MAX + SGC_QUERY_NOALLOC_MAX / 2;
static __device__ __align__(256) uint32_t scratch[QSCRATCH_WORDS];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;                  // guards g_scratch_ptr and serialises the calls that use it
void* g_scratch_ptr[SCRATCH_DEVICES] = {};

// The scratch array's address on `device` (looked up on first use, with that
// device current; the caller's current device is restored).  Caller holds
// g_scratch_mu.
hipError_t scratch_on(int device, uint32_t** out) {
    if (device < 0 || device >= SCRATCH_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t r = hipGetDevice(&cur);
        if (r == hipSuccess && cur != device) r = hipSetDevice(device);
        void* ptr = nullptr;
        if (r == hipSuccess) r = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) {
            hipError_t r2 = hipSetDevice(cur);
            if (r == hipSuccess) r = r2;
        }
        if (r != hipSuccess) return r;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = ptr;
    }
    *out = (uint32_t*)g_scratch_ptr[device];
    return hipSuccess;
}

int query_noalloc_impl(sgc_pool* P, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out) {
    GraphState* G = graph_of(P, gid);
    if (!G || (!ids && count) || (!out && count)) return SGC_ERR_INVALID;
    if (count > SGC_QUERY_NOALLOC_MAX)
        return P->fail(SGC_ERR_RANGE, "query_noalloc: " + std::to_string(count) + " ids exceed the " +
                                          std::to_string(SGC_QUERY_NOALLOC_MAX) + " of its static staging");
    if (count == 0) return SGC_OK;
    std::lock_guard<std::mutex> run(P->run_mu);
    std::lock_guard<std::mutex> lk(P->qmu);
    if (G->published < 0) {
        memset(out, 0, (size_t)count * 2);
        return SGC_OK;
    }
    std::lock_guard<std::mutex> sk(g_scratch_mu);
    uint32_t* base = nullptr;
    hipError_t r = scratch_on(P->device, &base);
    if (r != hipSuccess) return P->hip_fail(r, "query_noalloc: scratch address");
    return query_run(P, G, ids, count, out, base, (uint16_t*)(base + SGC_QUERY_NOALLOC_MAX));
}

}  // namespace

// ======================================================================== C ABI
extern "C" {

const char* sgc_version(void) { return "streamgc_amd 0.2 (gfx950)"; }

void sgc_default_config(sgc_config* c) {
    if (!c) return;
    c->epoch_size = 32768;
    c->batch_size = 32768;
    c->histogram_threshold = 64;
    c->slack_pct = 300;
    c->min_row_capacity = 8;
    c->overflow_capacity = 1u << 18;
    c->max_repair_iters = 20;
    c->plateau_iters = 3;
    c->tier2_passes = 30;
    c->flags = 0;
}

const char* sgc_last_error(const sgc_engine* e) {
    if (e) return e->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error.c_str();
}
const char* sgc_pool_last_error(const sgc_pool* p) {
    if (p) return p->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error.c_str();
}

int sgc_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int sgc_device_name(int device, char* name, uint32_t cap) {
    if (!name || cap == 0) return SGC_ERR_INVALID;
    hipDeviceProp_t prop;
    hipError_t r = hipGetDeviceProperties(&prop, device);
    if (r != hipSuccess) return global_fail(SGC_ERR_HIP, hipGetErrorString(r));
    if (prop.name[0]) {
        snprintf(name, cap, "%s", prop.name);
    } else {   // some runtimes leave the marketing name empty: fall back to the ISA name
        std::string arch(prop.gcnArchName);
        snprintf(name, cap, "%s", arch.substr(0, arch.find(':')).c_str());
    }
    return SGC_OK;
}

int sgc_microbench_trc(int device, uint32_t waves, uint32_t shared, uint32_t iters, uint32_t runs, double* ns,
                       uint32_t* final_words) {
    if (!ns || waves == 0 || waves > 65536 || iters == 0 || runs == 0 || runs > 101)
        return global_fail(SGC_ERR_INVALID, "sgc_microbench_trc: bad argument");
    hipError_t e = hipSetDevice(device);
    const uint32_t stride = shared ? 0u : 32u;   // one 128-byte line per wavefront, or one word for all
    const size_t nwords = shared ? 1 : (size_t)waves * stride;
    uint32_t* words = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    std::vector<double> t;
    if (e == hipSuccess) e = hipMalloc(&words, nwords * 4);
    if (e == hipSuccess) e = hipMemset(words, 0, nwords * 4);
    if (e == hipSuccess) e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = launch_trc(st, words, waves, stride, 1024);   // warm-up
    if (e == hipSuccess) e = hipMemsetAsync(words, 0, nwords * 4, st);
    for (uint32_t r = 0; r < runs && e == hipSuccess; r++) {
        e = hipEventRecord(a, st);
        if (e == hipSuccess) e = launch_trc(st, words, waves, stride, iters);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        t.push_back((double)ms * 1e6 / iters);
    }
    if (e == hipSuccess && final_words) {   // the first wavefront's word and, apart, the last one's
        e = hipMemcpy(&final_words[0], words, 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(&final_words[1], words + (size_t)(waves - 1) * stride, 4, hipMemcpyDeviceToHost);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st) (void)hipStreamDestroy(st);
    if (words) (void)hipFree(words);
    if (e != hipSuccess) return global_fail(SGC_ERR_HIP, std::string("sgc_microbench_trc: ") + hipGetErrorString(e));
    std::sort(t.begin(), t.end());
    *ns = t[t.size() / 2];
    return SGC_OK;
}

// ---------------------------------------------------------------- single graph
int sgc_create(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
               const uint32_t* deg0, sgc_engine** out) {
    if (!cfg || !row_ptr || !deg0 || !out || n == 0) return global_fail(SGC_ERR_INVALID, "sgc_create: bad argument");
    if (cfg->batch_size == 0 || cfg->epoch_size == 0) return global_fail(SGC_ERR_INVALID, "sgc_create: zero batch/epoch");
    sgc_engine* E = new sgc_engine;
    int r = pool_init(E, 1, cfg->batch_size, cfg->flags, 0);
    uint32_t gid = 0;
    if (r == SGC_OK) r = E->add_graph(cfg, n, row_ptr, col_idx, deg0, &gid);
    if (r != SGC_OK) {
        std::string m = "sgc_create: " + E->err;
        E->release();
        delete E;
        return global_fail(r, m);
    }
    *out = E;
    return SGC_OK;
}

int sgc_destroy(sgc_engine* e) {
    if (!e) return SGC_ERR_INVALID;
    e->release();
    delete e;
    return SGC_OK;
}

void sgc_default_tuning(sgc_tuning* t) {
    if (t) *t = sgc_tuning{1024, 2048, 16384, 65536, 64, 2};
}
static int set_tuning(sgc_pool* p, const sgc_tuning* t) {
    if (!p || !t || t->sweep_blocks == 0 || t->apply_blocks == 0 || t->sweep_blocks > 65535 ||
        t->apply_blocks > 65535 || t->fresh_repair > 2)
        return SGC_ERR_INVALID;
    p->tune = *t;
    return SGC_OK;
}
int sgc_set_tuning(sgc_engine* e, const sgc_tuning* t) { return set_tuning(e, t); }
int sgc_pool_set_tuning(sgc_pool* p, const sgc_tuning* t) { return set_tuning(p, t); }

int sgc_set_flags(sgc_engine* e, uint32_t flags) {
    if (!e || (flags & ~SGC_ALL_FLAGS)) return SGC_ERR_INVALID;
    const bool was = e->tracking();
    if (flags != e->flags) e->batch_drop();
    e->flags = flags;
    e->gs[0].cfg.flags = flags;
    hipError_t r = e->refresh_flags();
    if (r == hipSuccess && e->tracking() != was) r = e->set_tracking(e->tracking());
    if (r != hipSuccess) return e->hip_fail(r, "set flags");
    return SGC_OK;
}

int sgc_apply(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device) {
    if (!e || (!updates && count)) return SGC_ERR_INVALID;
    std::vector<uint64_t> pg;
    uint64_t done = 0;
    while (done < count) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(std::min(e->batch_cap, e->gs[0].cfg.batch_size), count - done);
        int r = e->apply(updates + done, b, on_device != 0, pg);
        if (r) return r;
        done += b;
    }
    hipError_t h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return e->hip_fail(h, "apply");
    e->resolve_timing();
    return SGC_OK;
}

int sgc_submit(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device,
               sgc_epoch_stats* epochs_out, uint32_t max_epochs, uint32_t* epochs_done) {
    if (!e || (!updates && count)) return SGC_ERR_INVALID;
    return submit_impl(e, updates, count, on_device, epochs_out, max_epochs, epochs_done);
}

int sgc_epoch(sgc_engine* e, sgc_epoch_stats* out) {
    if (!e) return SGC_ERR_INVALID;
    std::vector<sgc_epoch_stats> st;
    int r = e->epoch({0}, false, &st);
    if (r) return r;
    hipError_t h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return e->hip_fail(h, "epoch");
    e->resolve_timing();
    if (out) *out = st[0];
    return SGC_OK;
}

int sgc_finish(sgc_engine* e, sgc_epoch_stats* out) {
    if (!e) return SGC_ERR_INVALID;
    std::vector<sgc_epoch_stats> st;
    int r = e->epoch({0}, true, &st);
    if (r) return r;
    hipError_t h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return e->hip_fail(h, "finish");
    e->resolve_timing();
    if (out) *out = st[0];
    return SGC_OK;
}

int sgc_query(sgc_engine* e, const uint32_t* ids, uint32_t count, uint16_t* out_colors) {
    if (!e) return SGC_ERR_INVALID;
    return query_impl(e, 0, ids, count, out_colors);
}

int sgc_download_words(sgc_engine* e, uint32_t* w) {
    if (!e || !w) return SGC_ERR_INVALID;
    return download(e, w, e->gs[0].word, (size_t)e->gs[0].n * 4);
}
int sgc_download_colors(sgc_engine* e, uint16_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    std::vector<uint32_t> w(e->gs[0].n);
    int r = sgc_download_words(e, w.data());
    if (r) return r;
    for (uint32_t i = 0; i < e->gs[0].n; i++) c[i] = (uint16_t)(w[i] & 0xFFFF);
    return SGC_OK;
}
int sgc_download_snapshot(sgc_engine* e, uint16_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    GraphState& G = e->gs[0];
    if (G.published < 0) {
        memset(c, 0, (size_t)G.n * 2);
        return SGC_OK;
    }
    return download(e, c, G.snap + (size_t)G.published * G.n, (size_t)G.n * 2);
}
int sgc_num_slots(sgc_engine* e, uint64_t* s) {
    if (!e || !s) return SGC_ERR_INVALID;
    *s = e->gs[0].S;
    return SGC_OK;
}
int sgc_download_csr(sgc_engine* e, uint32_t* rp, uint32_t* col) {
    if (!e || !rp || !col) return SGC_ERR_INVALID;
    int r = download(e, rp, e->gs[0].row_ptr, ((size_t)e->gs[0].n + 1) * 4);
    if (r) return r;
    return download(e, col, e->gs[0].col, e->gs[0].S * 4);
}
int sgc_download_bitmaps(sgc_engine* e, uint64_t* b) {
    if (!e || !b) return SGC_ERR_INVALID;
    return download(e, b, e->gs[0].bitmap, (size_t)e->gs[0].n * 8);
}
int sgc_num_hd(sgc_engine* e, uint32_t* k) {
    if (!e || !k) return SGC_ERR_INVALID;
    *k = e->gs[0].nhd;
    return SGC_OK;
}
int sgc_download_hist(sgc_engine* e, uint32_t* h) {
    if (!e) return SGC_ERR_INVALID;
    if (e->gs[0].nhd == 0) return SGC_OK;
    if (!h) return SGC_ERR_INVALID;
    return download(e, h, e->gs[0].hist, (size_t)e->gs[0].nhd * HIST_BINS * 4);
}
int sgc_upload_words(sgc_engine* e, const uint32_t* w) {
    if (!e || !w) return SGC_ERR_INVALID;
    int r = upload(e, e->gs[0].word, w, (size_t)e->gs[0].n * 4);
    return r ? r : mark_all_sync(e);
}
int sgc_upload_bitmaps(sgc_engine* e, const uint64_t* b) {
    if (!e || !b) return SGC_ERR_INVALID;
    int r = upload(e, e->gs[0].bitmap, b, (size_t)e->gs[0].n * 8);
    return r ? r : mark_all_sync(e);
}
int sgc_upload_hist(sgc_engine* e, const uint32_t* h) {
    if (!e) return SGC_ERR_INVALID;
    if (e->gs[0].nhd == 0) return SGC_OK;
    if (!h) return SGC_ERR_INVALID;
    int r = upload(e, e->gs[0].hist, h, (size_t)e->gs[0].nhd * HIST_BINS * 4);
    return r ? r : mark_all_sync(e);
}
int sgc_overflow_count(sgc_engine* e, uint32_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    return download(e, c, e->d_ctr + C_OVF, 4);
}

static int totals_impl(sgc_pool* P, sgc_totals* out) {
    int r = download(P, P->h_totals, P->d_totals, (size_t)T_STRIPES * T_STRIDE * 8);
    if (r) return r;
    for (uint32_t k = 0; k < T_COUNT; k++)   // the stripes' sum lands in stripe 0
        for (uint32_t q = 1; q < T_STRIPES; q++) P->h_totals[k] += P->h_totals[(size_t)q * T_STRIDE + k];
    memset(out, 0, sizeof *out);
    out->updates = P->h_totals[T_UPDATES];
    out->scan_slots = P->h_totals[T_SCAN];
    out->hd_endpoints = P->h_totals[T_HD];
    out->recolors = P->h_totals[T_RECOLOR];
    out->recolor_cap_slots = P->h_totals[T_RC_CAP];
    out->recolor_live = P->h_totals[T_RC_LIVE];
    out->overflow_events = P->h_totals[T_OVF_EVENTS];
    out->live_slots = P->h_totals[T_LIVE_SLOTS];
    out->grouped = P->h_totals[T_GROUPED];
    for (int k = 0; k < 5; k++) {
        out->sweep_rows[k] = P->h_totals[T_SWEEP_ROWS + k];
        out->sweep_slots[k] = P->h_totals[T_SWEEP_SLOTS + k];
    }
    for (auto& G : P->gs) out->epochs += G.epochs;
    out->t1_sweeps = P->t1_sweeps;
    out->t2_runs = P->t2_runs;
    return SGC_OK;
}
int sgc_totals_get(sgc_engine* e, sgc_totals* out) {
    if (!e || !out) return SGC_ERR_INVALID;
    return totals_impl(e, out);
}

int sgc_find_color(sgc_engine* e, uint32_t v, uint16_t* out) {
    if (!e || !out || v >= e->gs[0].n) return SGC_ERR_INVALID;
    hipError_t r = launch_find_color(e->ctx(false), 0, v, e->d_tmp16);
    if (r != hipSuccess) return e->hip_fail(r, "find colour");
    return download(e, out, e->d_tmp16, 2);
}

static int validate_impl(sgc_pool* P, uint32_t gid, int use_snapshot, uint64_t* violations, uint64_t* uncolored,
                         uint32_t* max_color) {
    GraphState* G = graph_of(P, gid);
    if (!G) return SGC_ERR_INVALID;
    if (use_snapshot && G->published < 0) return P->fail(SGC_ERR_STATE, "no snapshot published");
    hipError_t r = hipMemsetAsync(P->d_ctr + (size_t)gid * C_COUNT + C_VIOL, 0, 3 * 4, P->stream);
    const uint16_t* colors = use_snapshot ? G->snap + (size_t)G->published * G->n : nullptr;
    if (r == hipSuccess) r = launch_validate(P->ctx(false), gid, G->n, colors);
    if (r == hipSuccess) r = P->read_ctr();
    if (r != hipSuccess) return P->hip_fail(r, "validate");
    if (violations) *violations = P->ctr(gid, C_VIOL);
    if (uncolored) *uncolored = P->ctr(gid, C_UNCOL);
    if (max_color) *max_color = P->ctr(gid, C_MAXC);
    return SGC_OK;
}
int sgc_validate(sgc_engine* e, int use_snapshot, uint64_t* violations, uint64_t* uncolored, uint32_t* max_color) {
    if (!e) return SGC_ERR_INVALID;
    return validate_impl(e, 0, use_snapshot, violations, uncolored, max_color);
}

static int history_reserve_impl(sgc_pool* P, uint32_t gid, uint32_t capacity) {
    GraphState* G = graph_of(P, gid);
    if (!G) return SGC_ERR_INVALID;
    (void)hipStreamSynchronize(P->stream);
    if (G->history) (void)hipFree(G->history);
    G->history = nullptr;
    G->hist_cap = G->hist_count = 0;
    if (capacity == 0) return SGC_OK;
    hipError_t r = hipMalloc(&G->history, (size_t)capacity * G->n * 2);
    if (r != hipSuccess) return P->hip_fail(r, "history alloc");
    G->hist_cap = capacity;
    return SGC_OK;
}
int sgc_history_reserve(sgc_engine* e, uint32_t capacity) {
    if (!e) return SGC_ERR_INVALID;
    return history_reserve_impl(e, 0, capacity);
}
int sgc_history_count(sgc_engine* e, uint32_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    *c = e->gs[0].hist_count;
    return SGC_OK;
}

static int validate_history_impl(sgc_pool* P, uint32_t gid, const uint32_t* eu, const uint32_t* ev,
                                 const uint64_t* born, const uint64_t* died, uint64_t m, const uint64_t* pos,
                                 uint32_t count, uint64_t* bad_out) {
    GraphState* G = graph_of(P, gid);
    if (!G || !pos || !bad_out || count > G->hist_count) return SGC_ERR_INVALID;
    if (count == 0) return SGC_OK;
    uint32_t *deu = nullptr, *dev = nullptr;
    uint64_t *dborn = nullptr, *ddied = nullptr, *dpos = nullptr;
    unsigned long long* dbad = nullptr;
    const size_t mm = m ? m : 1;
    hipError_t r = hipMalloc(&deu, mm * 4);
    if (r == hipSuccess) r = hipMalloc(&dev, mm * 4);
    if (r == hipSuccess) r = hipMalloc(&dborn, mm * 8);
    if (r == hipSuccess) r = hipMalloc(&ddied, mm * 8);
    if (r == hipSuccess) r = hipMalloc(&dpos, (size_t)count * 8);
    if (r == hipSuccess) r = hipMalloc(&dbad, (size_t)count * 8);
    hipStream_t s = P->stream;
    if (r == hipSuccess && m) r = hipMemcpyAsync(deu, eu, m * 4, hipMemcpyHostToDevice, s);
    if (r == hipSuccess && m) r = hipMemcpyAsync(dev, ev, m * 4, hipMemcpyHostToDevice, s);
    if (r == hipSuccess && m) r = hipMemcpyAsync(dborn, born, m * 8, hipMemcpyHostToDevice, s);
    if (r == hipSuccess && m) r = hipMemcpyAsync(ddied, died, m * 8, hipMemcpyHostToDevice, s);
    if (r == hipSuccess) r = hipMemcpyAsync(dpos, pos, (size_t)count * 8, hipMemcpyHostToDevice, s);
    if (r == hipSuccess) r = hipMemsetAsync(dbad, 0, (size_t)count * 8, s);
    if (r == hipSuccess) r = launch_validate_history(s, G->history, G->n, deu, dev, dborn, ddied, m, dpos, count, dbad);
    if (r == hipSuccess) r = hipMemcpyAsync(bad_out, dbad, (size_t)count * 8, hipMemcpyDeviceToHost, s);
    if (r == hipSuccess) r = hipStreamSynchronize(s);
    for (void* p : {(void*)deu, (void*)dev, (void*)dborn, (void*)ddied, (void*)dpos, (void*)dbad})
        if (p) (void)hipFree(p);
    if (r != hipSuccess) return P->hip_fail(r, "validate history");
    return SGC_OK;
}
int sgc_validate_history(sgc_engine* e, const uint32_t* eu, const uint32_t* ev, const uint64_t* born,
                         const uint64_t* died, uint64_t m, const uint64_t* pos, uint32_t count, uint64_t* bad_out) {
    if (!e) return SGC_ERR_INVALID;
    return validate_history_impl(e, 0, eu, ev, born, died, m, pos, count, bad_out);
}

int sgc_device_upload(const void* host, uint64_t bytes, void** dev_out) {
    if (!dev_out || (!host && bytes)) return SGC_ERR_INVALID;
    void* p = nullptr;
    hipError_t r = hipMalloc(&p, bytes ? bytes : 1);
    if (r != hipSuccess) return global_fail(SGC_ERR_NOMEM, std::string("device upload: ") + hipGetErrorString(r));
    if (bytes) r = hipMemcpy(p, host, bytes, hipMemcpyHostToDevice);
    if (r != hipSuccess) {
        (void)hipFree(p);
        return global_fail(SGC_ERR_HIP, std::string("device upload: ") + hipGetErrorString(r));
    }
    *dev_out = p;
    return SGC_OK;
}
int sgc_device_free(void* dev) {
    if (dev) (void)hipFree(dev);
    return SGC_OK;
}

static int last_timing_impl(sgc_pool* P, double* ms4) {
    double upd = 0, ep = 0;
    for (auto& kv : P->ktimes) {
        if (kv.first == "apply") upd += kv.second.ms;
        else ep += kv.second.ms;
    }
    ms4[0] = upd;
    ms4[1] = ep;
    ms4[2] = P->last_ms[2];
    ms4[3] = P->last_ms[3];
    return SGC_OK;
}
int sgc_last_timing(sgc_engine* e, double* ms4) {
    if (!e || !ms4) return SGC_ERR_INVALID;
    return last_timing_impl(e, ms4);
}
static int timing_enable_impl(sgc_pool* P, int enable) {
    P->resolve_timing();
    P->ktiming = enable != 0;
    if (enable) P->ktimes.clear();
    return SGC_OK;
}
int sgc_kernel_timing_enable(sgc_engine* e, int enable) {
    if (!e) return SGC_ERR_INVALID;
    return timing_enable_impl(e, enable);
}
static int kernel_times_impl(sgc_pool* P, char* names, uint32_t stride, double* total_ms, uint64_t* launches,
                             uint32_t cap, uint32_t* count) {
    P->resolve_timing();
    uint32_t k = 0;
    for (auto& kv : P->ktimes) {
        if (k < cap) {
            if (names && stride) {
                strncpy(names + (size_t)k * stride, kv.first.c_str(), stride - 1);
                names[(size_t)k * stride + stride - 1] = 0;
            }
            if (total_ms) total_ms[k] = kv.second.ms;
            if (launches) launches[k] = kv.second.launches;
        }
        k++;
    }
    *count = k;
    return SGC_OK;
}
int sgc_kernel_times(sgc_engine* e, char* names, uint32_t stride, double* total_ms, uint64_t* launches, uint32_t cap,
                     uint32_t* count) {
    if (!e || !count) return SGC_ERR_INVALID;
    return kernel_times_impl(e, names, stride, total_ms, launches, cap, count);
}

// ---------------------------------------------------------------- pools
int sgc_pool_create(const sgc_pool_config* cfg, sgc_pool** out) {
    if (!cfg || !out || cfg->max_graphs == 0 || cfg->batch_capacity == 0)
        return global_fail(SGC_ERR_INVALID, "pool create: bad argument");
    if (cfg->flags & ~SGC_ALL_FLAGS) return global_fail(SGC_ERR_INVALID, "pool create: unknown flags");
    sgc_pool* P = new sgc_pool;
    int r = pool_init(P, cfg->max_graphs, cfg->batch_capacity, cfg->flags, cfg->hbm_budget);
    if (r != SGC_OK) {
        std::string m = P->err;
        P->release();
        delete P;
        return global_fail(r, m);
    }
    *out = P;
    return SGC_OK;
}
int sgc_pool_destroy(sgc_pool* p) {
    if (!p) return SGC_ERR_INVALID;
    p->release();
    delete p;
    return SGC_OK;
}
int sgc_pool_graph_bytes(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* deg0,
                         uint64_t* bytes) {
    if (!cfg || !row_ptr || !deg0 || !bytes) return SGC_ERR_INVALID;
    std::vector<uint32_t> rp, hd, hubs;
    uint64_t S = 0;
    int r = plan_graph(cfg, n, row_ptr, deg0, rp, hd, hubs, S);
    if (r) return r;
    *bytes = graph_bytes(n, S, (uint32_t)hubs.size(), cfg->overflow_capacity);
    return SGC_OK;
}
int sgc_pool_add(sgc_pool* p, const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                 const uint32_t* deg0, uint32_t* gid) {
    if (!p || !gid) return SGC_ERR_INVALID;
    return p->add_graph(cfg, n, row_ptr, col_idx, deg0, gid);   // the pool's flags govern its graphs
}
int sgc_pool_remove(sgc_pool* p, uint32_t gid) {
    if (!p) return SGC_ERR_INVALID;
    return p->remove_graph(gid);
}
int sgc_pool_usage(sgc_pool* p, uint64_t* used, uint64_t* budget, uint32_t* graphs) {
    if (!p) return SGC_ERR_INVALID;
    if (used) *used = p->used_bytes;
    if (budget) *budget = p->budget;
    if (graphs) {
        uint32_t k = 0;
        for (auto& G : p->gs) k += G.used;
        *graphs = k;
    }
    return SGC_OK;
}
int sgc_pool_set_flags(sgc_pool* p, uint32_t flags) {
    if (!p || (flags & ~SGC_ALL_FLAGS)) return SGC_ERR_INVALID;
    const bool was = p->tracking();
    if (flags != p->flags) p->batch_drop();   // captured rounds follow the flags they were made with
    p->flags = flags;
    for (auto& G : p->gs) G.cfg.flags = flags;
    hipError_t r = p->refresh_flags();
    if (r == hipSuccess && p->tracking() != was) r = p->set_tracking(p->tracking());
    if (r != hipSuccess) return p->hip_fail(r, "set flags");
    return SGC_OK;
}
int sgc_pool_apply(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, int on_device) {
    if (!p || (!updates && count)) return SGC_ERR_INVALID;
    std::vector<uint64_t> pg;
    uint64_t done = 0;
    while (done < count) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(p->batch_cap, count - done);
        int r = p->apply(updates + done, b, on_device != 0, pg);
        if (r) return r;
        done += b;
    }
    hipError_t h = hipStreamSynchronize(p->stream);
    if (h != hipSuccess) return p->hip_fail(h, "pool apply");
    p->resolve_timing();
    return SGC_OK;
}
int sgc_pool_due(sgc_pool* p, uint32_t* gids, uint32_t cap, uint32_t* count) {
    if (!p || !count) return SGC_ERR_INVALID;
    uint32_t k = 0;
    for (uint32_t g = 0; g < p->ng; g++)
        if (p->gs[g].used && p->gs[g].since_epoch >= p->gs[g].cfg.epoch_size) {
            if (gids && k < cap) gids[k] = g;
            k++;
        }
    *count = k;
    return SGC_OK;
}
static int pool_epoch_impl(sgc_pool* p, const uint32_t* gids, uint32_t count, sgc_epoch_stats* out, bool finish) {
    if (!p || (!gids && count)) return SGC_ERR_INVALID;
    std::vector<uint32_t> set;
    for (uint32_t k = 0; k < count; k++) {
        if (!graph_of(p, gids[k])) return p->fail(SGC_ERR_INVALID, "epoch: unknown graph");
        if (std::find(set.begin(), set.end(), gids[k]) == set.end()) set.push_back(gids[k]);
    }
    std::vector<sgc_epoch_stats> st;
    int r = p->epoch(set, finish, out ? &st : nullptr);
    if (r) return r;
    hipError_t h = hipStreamSynchronize(p->stream);
    if (h != hipSuccess) return p->hip_fail(h, "pool epoch");
    p->resolve_timing();
    if (out)
        for (uint32_t k = 0; k < count; k++) {
            const size_t at = std::find(set.begin(), set.end(), gids[k]) - set.begin();
            out[k] = st[at];
        }
    return SGC_OK;
}
int sgc_pool_epoch(sgc_pool* p, const uint32_t* gids, uint32_t count, sgc_epoch_stats* out) {
    return pool_epoch_impl(p, gids, count, out, false);
}
int sgc_pool_finish(sgc_pool* p, const uint32_t* gids, uint32_t count, sgc_epoch_stats* out) {
    return pool_epoch_impl(p, gids, count, out, true);
}
int sgc_pool_submit(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, int on_device,
                    sgc_epoch_stats* stats_out, uint32_t max_stats, uint32_t* nstats) {
    if (!p || (!updates && count)) return SGC_ERR_INVALID;
    if (on_device) return p->fail(SGC_ERR_INVALID, "pool submit takes host records (split per graph on the host)");
    return pool_submit_impl(p, updates, count, stats_out, max_stats, nstats);
}
int sgc_pool_query(sgc_pool* p, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out_colors) {
    if (!p) return SGC_ERR_INVALID;
    return query_impl(p, gid, ids, count, out_colors);
}
int sgc_pool_query_noalloc(sgc_pool* p, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out_colors) {
    if (!p) return SGC_ERR_INVALID;
    return query_noalloc_impl(p, gid, ids, count, out_colors);
}
int sgc_pool_download_words(sgc_pool* p, uint32_t gid, uint32_t* words) {
    GraphState* G = graph_of(p, gid);
    if (!G || !words) return SGC_ERR_INVALID;
    return download(p, words, G->word, (size_t)G->n * 4);
}
int sgc_pool_upload_words(sgc_pool* p, uint32_t gid, const uint32_t* words) {
    GraphState* G = graph_of(p, gid);
    if (!G || !words) return SGC_ERR_INVALID;
    int r = upload(p, G->word, words, (size_t)G->n * 4);
    if (r) return r;
    hipError_t e = p->mark_all(gid);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return p->hip_fail(e, "mark");
    return SGC_OK;
}
int sgc_pool_download_csr(sgc_pool* p, uint32_t gid, uint32_t* rp, uint32_t* col) {
    GraphState* G = graph_of(p, gid);
    if (!G || !rp || !col) return SGC_ERR_INVALID;
    int r = download(p, rp, G->row_ptr, ((size_t)G->n + 1) * 4);
    return r ? r : download(p, col, G->col, G->S * 4);
}
int sgc_pool_download_bitmaps(sgc_pool* p, uint32_t gid, uint64_t* bitmaps) {
    GraphState* G = graph_of(p, gid);
    if (!G || !bitmaps) return SGC_ERR_INVALID;
    return download(p, bitmaps, G->bitmap, (size_t)G->n * 8);
}
int sgc_pool_download_hist(sgc_pool* p, uint32_t gid, uint32_t* hist) {
    GraphState* G = graph_of(p, gid);
    if (!G || (!hist && G->nhd)) return SGC_ERR_INVALID;
    return G->nhd ? download(p, hist, G->hist, (size_t)G->nhd * HIST_BINS * 4) : SGC_OK;
}
int sgc_pool_graph_info(sgc_pool* p, uint32_t gid, sgc_graph_info* out) {
    GraphState* G = graph_of(p, gid);
    if (!G || !out) return SGC_ERR_INVALID;
    uint32_t c[C_COUNT];
    int r = download(p, c, p->d_ctr + (size_t)gid * C_COUNT, sizeof c);
    if (r) return r;
    memset(out, 0, sizeof *out);
    out->n = G->n;
    out->num_hd = G->nhd;
    out->slots = G->S;
    out->bytes = G->bytes;
    out->epoch_size = G->cfg.epoch_size;
    out->epochs = G->epochs;
    out->overflow_events = c[C_ASYM];
    out->lost = c[C_LOST];
    out->published = G->published;
    out->tier2_ms = G->t2_ms;
    out->tier2_runs = G->t2_count;
    return SGC_OK;
}
int sgc_pool_download_snapshot(sgc_pool* p, uint32_t gid, uint16_t* colors) {
    GraphState* G = graph_of(p, gid);
    if (!G || !colors) return SGC_ERR_INVALID;
    if (G->published < 0) {
        memset(colors, 0, (size_t)G->n * 2);
        return SGC_OK;
    }
    return download(p, colors, G->snap + (size_t)G->published * G->n, (size_t)G->n * 2);
}
int sgc_pool_validate(sgc_pool* p, uint32_t gid, int use_snapshot, uint64_t* violations, uint64_t* uncolored,
                      uint32_t* max_color) {
    if (!p) return SGC_ERR_INVALID;
    return validate_impl(p, gid, use_snapshot, violations, uncolored, max_color);
}
int sgc_pool_history_reserve(sgc_pool* p, uint32_t gid, uint32_t capacity) {
    if (!p) return SGC_ERR_INVALID;
    return history_reserve_impl(p, gid, capacity);
}
int sgc_pool_history_count(sgc_pool* p, uint32_t gid, uint32_t* count) {
    GraphState* G = graph_of(p, gid);
    if (!G || !count) return SGC_ERR_INVALID;
    *count = G->hist_count;
    return SGC_OK;
}
int sgc_pool_validate_history(sgc_pool* p, uint32_t gid, const uint32_t* eu, const uint32_t* ev, const uint64_t* born,
                              const uint64_t* died, uint64_t m, const uint64_t* pos, uint32_t count,
                              uint64_t* bad_out) {
    if (!p) return SGC_ERR_INVALID;
    return validate_history_impl(p, gid, eu, ev, born, died, m, pos, count, bad_out);
}
int sgc_pool_totals(sgc_pool* p, sgc_totals* out) {
    if (!p || !out) return SGC_ERR_INVALID;
    return totals_impl(p, out);
}
int sgc_pool_set_epoch_size(sgc_pool* p, uint32_t gid, uint32_t epoch_size) {
    GraphState* G = graph_of(p, gid);
    if (!G || epoch_size == 0) return SGC_ERR_INVALID;
    G->cfg.epoch_size = epoch_size;
    return SGC_OK;
}
void sgc_default_batch_config(sgc_batch_config* c) {
    if (!c) return;
    c->use_graph = 1;
    c->t1_sweeps = 0;
    c->autotune_rounds = 0;
    c->t2_passes = 0;
}
int sgc_pool_batch_prepare(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, const sgc_batch_config* cfg,
                           uint32_t* rounds) {
    if (!p || (!updates && count)) return SGC_ERR_INVALID;
    sgc_batch_config c;
    sgc_default_batch_config(&c);
    if (cfg) c = *cfg;
    return p->batch_prepare(updates, count, c, rounds);
}
int sgc_pool_batch_run(sgc_pool* p, sgc_batch_stats* out) {
    if (!p) return SGC_ERR_INVALID;
    return p->batch_run(out);
}
int sgc_batch_prepare(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, const sgc_batch_config* cfg,
                      uint32_t* rounds) {
    return sgc_pool_batch_prepare(e, updates, count, cfg, rounds);
}
int sgc_batch_run(sgc_engine* e, sgc_batch_stats* out) { return sgc_pool_batch_run(e, out); }
int sgc_batch_release(sgc_engine* e) { return sgc_pool_batch_release(e); }
int sgc_pool_batch_release(sgc_pool* p) {
    if (!p) return SGC_ERR_INVALID;
    p->batch_drop();
    return SGC_OK;
}
int sgc_pool_kernel_timing_enable(sgc_pool* p, int enable) {
    if (!p) return SGC_ERR_INVALID;
    return timing_enable_impl(p, enable);
}
int sgc_pool_kernel_times(sgc_pool* p, char* names, uint32_t stride, double* total_ms, uint64_t* launches,
                          uint32_t cap, uint32_t* count) {
    if (!p || !count) return SGC_ERR_INVALID;
    return kernel_times_impl(p, names, stride, total_ms, launches, cap, count);
}

}  // extern "C"
