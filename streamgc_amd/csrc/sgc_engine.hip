// sgc_engine.cpp -- the single-graph StreamGC engine behind the C ABI.
//
// Layout in HBM (one graph): word[n] u32, deg0[n] u32, row_ptr[n+1] u32,
// col[S] u32 (slack rows), bitmap[n] u64, hd_map[n] u32, hist[nhd*1024] u32,
// overflow 2*cap u32, counters, snapshot 2 x n u16 (double buffer).
// All work is enqueued on one HIP stream; the host synchronises only where a
// data-dependent decision is needed (Tier-1 convergence, the Tier-2 check).
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
#include "streamgc.h"

using namespace sgc;

namespace {

std::string g_last_error;
std::mutex g_err_mu;

struct KTime {
    double ms = 0;
    uint64_t launches = 0;
};

struct PendingTiming {
    const char* name;
    hipEvent_t a, b;
};

}  // namespace

struct sgc_engine {
    sgc_config cfg{};
    uint32_t n = 0;
    uint64_t S = 0;
    uint32_t nhd = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t qstream = nullptr;

    uint32_t *word = nullptr, *deg0 = nullptr, *row_ptr = nullptr, *col = nullptr, *hd_map = nullptr;
    uint32_t* hubs = nullptr;
    uint32_t *hist = nullptr, *ovf = nullptr, *ctr = nullptr;
    uint64_t* bitmap = nullptr;
    uint16_t* snap = nullptr;            // 2 * n
    uint16_t** snap_dst = nullptr;       // device: {snap, snap + n}
    GraphDev hg{};
    GraphDev* dg = nullptr;
    uint64_t* totals = nullptr;          // device T_COUNT
    sgc_edge_update* batch = nullptr;    // device staging, batch_size records
    uint32_t* seg = nullptr;             // device: one segment [0, count)
    EdgeGroups eg;                       // per-edge grouping of parallel batches
    uint32_t* h_seg = nullptr;           // pinned
    uint32_t* h_ctr = nullptr;           // pinned C_COUNT
    uint64_t* h_totals = nullptr;        // pinned T_COUNT
    uint16_t* d_tmp16 = nullptr;
    uint32_t* d_qids = nullptr;          // query staging
    uint16_t* d_qout = nullptr;
    uint32_t qcap = 0;

    int published = -1;                  // half holding the published snapshot
    hipEvent_t pub_ev[2] = {nullptr, nullptr};
    hipEvent_t read_ev[2] = {nullptr, nullptr};
    std::mutex qmu;

    uint16_t* history = nullptr;
    uint32_t hist_cap = 0, hist_count = 0;

    uint32_t epochs = 0;
    uint64_t t1_sweeps = 0, t2_runs = 0;
    std::string err;

    bool ktiming = false;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, KTime> ktimes;
    double last_ms[4] = {0, 0, 0, 0};

    bool deterministic() const { return (cfg.flags & SGC_DETERMINISTIC) != 0; }
    bool exact_hist() const { return (cfg.flags & SGC_EXACT_HISTOGRAM) != 0; }

    // Bitmap rebuild, plus the hub histograms in exact-histogram mode.
    hipError_t rebuild(const LaunchCtx& L) {
        hipError_t e = timed("rebuild", [&] { return launch_rebuild(L); });
        if (e == hipSuccess && exact_hist() && nhd) e = timed("rebuild_hist", [&] { return launch_rebuild_hist(L); });
        return e;
    }

    int fail(int code, const std::string& msg) {
        err = msg;
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_error = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        return fail(SGC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }

    LaunchCtx ctx() const {
        LaunchCtx L;
        L.stream = stream;
        L.graphs = dg;
        L.ng = 1;
        L.vbase = nullptr;
        L.V = n;
        L.active = nullptr;
        L.totals = totals;
        L.H = nhd;
        return L;
    }

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

    hipError_t read_ctr() {
        hipError_t e = hipMemcpyAsync(h_ctr, ctr, C_COUNT * sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(stream);
    }
    hipError_t zero_ctr(uint32_t idx) {
        return hipMemsetAsync(ctr + idx, 0, sizeof(uint32_t), stream);
    }

    int apply_device(const sgc_edge_update* d_upd, uint32_t count) {
        if (count == 0) return SGC_OK;
        LaunchCtx L = ctx();
        hipError_t e;
        if (deterministic()) {
            h_seg[0] = 0;
            h_seg[1] = count;
            e = hipMemcpyAsync(seg, h_seg, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) return hip_fail(e, "segment upload");
            // the pinned segment is reused: wait until the copy has consumed it
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "segment upload");
            e = timed("apply", [&] { return launch_apply_seg(L, d_upd, seg, 1, 0); });
        } else {
            e = timed("apply", [&] { return launch_apply(L, d_upd, count, 0, &eg); });
        }
        if (e != hipSuccess) return hip_fail(e, "apply launch");
        return SGC_OK;
    }

    int snapshot_publish() {
        const int half = published == 0 ? 1 : 0;
        LaunchCtx L = ctx();
        // queries still reading this half must finish first
        hipError_t e = hipStreamWaitEvent(stream, read_ev[half], 0);
        if (e != hipSuccess) return hip_fail(e, "snapshot wait");
        e = timed("snapshot", [&] { return launch_snapshot_to(L, half); });
        if (e != hipSuccess) return hip_fail(e, "snapshot launch");
        e = hipEventRecord(pub_ev[half], stream);
        if (e != hipSuccess) return hip_fail(e, "snapshot event");
        if (history && hist_count < hist_cap) {
            e = hipMemcpyAsync(history + (size_t)hist_count * n, snap + (size_t)half * n, (size_t)n * 2,
                               hipMemcpyDeviceToDevice, stream);
            if (e != hipSuccess) return hip_fail(e, "history copy");
            hist_count++;
        }
        std::lock_guard<std::mutex> lk(qmu);
        published = half;
        return SGC_OK;
    }
    hipError_t launch_snapshot_to(const LaunchCtx& L, int half);

    int drain(sgc_epoch_stats* st) {
        LaunchCtx L = ctx();
        hipError_t e = timed("drain", [&] { return launch_drain(L, deterministic(), cfg.overflow_capacity); });
        if (e != hipSuccess) return hip_fail(e, "drain launch");
        (void)st;
        return SGC_OK;
    }

    // run_conflict_repair (A/src/streamgc.cu:1317-1393)
    int tier1(sgc_epoch_stats* st) {
        LaunchCtx L = ctx();
        uint32_t prev = 0xFFFFFFFFu, plateau = 0;
        for (uint32_t it = 0; it < cfg.max_repair_iters; it++) {
            hipError_t e = rebuild(L);
            if (e == hipSuccess) e = zero_ctr(C_CONFLICTS);
            if (e == hipSuccess) e = timed("tier1", [&] { return launch_tier1(L, deterministic()); });
            if (e == hipSuccess) e = read_ctr();
            if (e != hipSuccess) return hip_fail(e, "tier-1");
            t1_sweeps++;
            const uint32_t c = h_ctr[C_CONFLICTS];
            st->t1_iterations = it + 1;
            st->t1_conflicts_last = c;
            if (c == 0) break;
            if (c == prev) {
                if (++plateau >= cfg.plateau_iters) break;
            } else {
                plateau = 0;
            }
            prev = c;
        }
        return SGC_OK;
    }

    int t2_check(uint32_t* bad) {
        LaunchCtx L = ctx();
        hipError_t e = zero_ctr(C_T2BAD);
        if (e == hipSuccess) e = timed("t2_check", [&] { return launch_t2_check(L); });
        if (e == hipSuccess) e = read_ctr();
        if (e != hipSuccess) return hip_fail(e, "tier-2 check");
        *bad = h_ctr[C_T2BAD];
        return SGC_OK;
    }

    // cpu_serial_repair (A/src/streamgc.cu:1057-1240), run only when the
    // check finds work (otherwise it would change nothing).
    int tier2(sgc_epoch_stats* st) {
        uint32_t bad = 0;
        int r = t2_check(&bad);
        if (r) return r;
        if (bad == 0) {
            st->t2_remaining = 0;
            return SGC_OK;
        }
        auto t0 = std::chrono::steady_clock::now();
        std::vector<uint32_t> w(n), rp(n + 1), c(S);
        hipError_t e = hipMemcpyAsync(w.data(), word, (size_t)n * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rp.data(), row_ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c.data(), col, S * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "tier-2 download");
        st->t2_fixed = tier2_host(n, rp.data(), c.data(), w.data(), cfg.tier2_passes);
        st->t2_ran = 1;
        t2_runs++;
        e = hipMemcpyAsync(word, w.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "tier-2 upload");
        last_ms[2] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        r = t2_check(&bad);
        if (r) return r;
        st->t2_remaining = bad;
        return SGC_OK;
    }

    int max_color(uint32_t* out);

    int read_drain_stats(sgc_epoch_stats* st) {
        hipError_t e = read_ctr();
        if (e != hipSuccess) return hip_fail(e, "counters");
        st->overflow_count = h_ctr[C_OVF_RAW];
        st->overflow_drained = h_ctr[C_OVF_RAW] ? h_ctr[C_OVF_SNAP] : 0;
        st->overflow_failed = h_ctr[C_DRAIN_FAILED];
        return SGC_OK;
    }

    // One epoch step: reference order (A/src/streamgc.cu:949-1036) or the
    // repair-every-epoch order (README.md:170-176).
    int epoch(sgc_epoch_stats* out) {
        sgc_epoch_stats st{};
        st.epoch_index = epochs++;
        LaunchCtx L = ctx();
        hipError_t e;
        int r;
        if (cfg.flags & SGC_REPAIR_EVERY_EPOCH) {
            e = timed("compact", [&] { return launch_compact(L); });
            if (e != hipSuccess) return hip_fail(e, "compact");
            if ((r = drain(&st))) return r;
            if ((r = read_drain_stats(&st))) return r;
            if ((r = tier1(&st))) return r;
            if ((r = tier2(&st))) return r;
            if ((r = snapshot_publish())) return r;
            st.snapshot_published = 1;
        } else {
            if ((r = snapshot_publish())) return r;
            st.snapshot_published = 1;
            e = timed("compact", [&] { return launch_compact(L); });
            if (e != hipSuccess) return hip_fail(e, "compact");
            if ((r = drain(&st))) return r;
            e = rebuild(L);
            if (e != hipSuccess) return hip_fail(e, "rebuild");
            if (out && (r = read_drain_stats(&st))) return r;
        }
        if (out) {
            if ((r = max_color(&st.max_color))) return r;
            *out = st;
        }
        return SGC_OK;
    }

    // drain_and_compact + run_conflict_repair + cpu_serial_repair
    int finish(sgc_epoch_stats* out) {
        sgc_epoch_stats st{};
        st.epoch_index = epochs;
        LaunchCtx L = ctx();
        int r;
        hipError_t e = timed("compact", [&] { return launch_compact(L); });
        if (e != hipSuccess) return hip_fail(e, "compact");
        if ((r = drain(&st))) return r;
        if ((r = read_drain_stats(&st))) return r;
        e = rebuild(L);
        if (e != hipSuccess) return hip_fail(e, "rebuild");
        if ((r = tier1(&st))) return r;
        if ((r = tier2(&st))) return r;
        if ((r = snapshot_publish())) return r;
        st.snapshot_published = 1;
        if ((r = max_color(&st.max_color))) return r;
        if (out) *out = st;
        return SGC_OK;
    }

    void release();
};

// Device helpers that live in this translation unit.
namespace {
__global__ void k_snapshot_one(const GraphDev* __restrict__ g, uint16_t* __restrict__ dst) {
    const uint32_t n = g->n;
    const SGC_G uint32_t* w = gp(g->word);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        gp(dst)[i] = (uint16_t)(w[i] & 0xFFFFu);
}
__global__ void k_max_color(const GraphDev* __restrict__ g, uint32_t* out) {
    const uint32_t n = g->n;
    uint32_t mx = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = gp(g->word)[i] & 0xFFFFu;
        mx = c > mx ? c : mx;
    }
    mx = wave_max32(mx);
    if (lane() == 0 && mx) max32(out, mx);
}
__global__ void k_gather16(const uint16_t* __restrict__ src, uint32_t n, const uint32_t* __restrict__ ids,
                           uint32_t count, uint16_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t v = ids[i];
        out[i] = v < n ? src[v] : 0;
    }
}
}  // namespace

hipError_t sgc_engine::launch_snapshot_to(const LaunchCtx& L, int half) {
    uint32_t grid = (n + 1023) / 1024;
    grid = grid < 1 ? 1 : (grid > 8192 ? 8192 : grid);
    k_snapshot_one<<<grid, 256, 0, L.stream>>>(dg, snap + (size_t)half * n);
    return hipGetLastError();
}

int sgc_engine::max_color(uint32_t* out) {
    hipError_t e = zero_ctr(C_MAXC);
    uint32_t grid = (n + 1023) / 1024;
    grid = grid < 1 ? 1 : (grid > 4096 ? 4096 : grid);
    if (e == hipSuccess) {
        k_max_color<<<grid, 256, 0, stream>>>(dg, ctr + C_MAXC);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = read_ctr();
    if (e != hipSuccess) return hip_fail(e, "max colour");
    *out = h_ctr[C_MAXC];
    return SGC_OK;
}

void sgc_engine::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (qstream) (void)hipStreamSynchronize(qstream);
    for (void* p : {(void*)word, (void*)hubs, (void*)deg0, (void*)row_ptr, (void*)col, (void*)hd_map, (void*)hist, (void*)ovf,
                    (void*)ctr, (void*)bitmap, (void*)snap, (void*)snap_dst, (void*)dg, (void*)totals, (void*)batch,
                    (void*)seg, (void*)d_tmp16, (void*)d_qids, (void*)d_qout, (void*)history, (void*)eg.keys,
                    (void*)eg.head, (void*)eg.next})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)h_seg, (void*)h_ctr, (void*)h_totals})
        if (p) (void)hipHostFree(p);
    for (auto& p : pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; i++) {
        if (pub_ev[i]) (void)hipEventDestroy(pub_ev[i]);
        if (read_ev[i]) (void)hipEventDestroy(read_ev[i]);
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (qstream) (void)hipStreamDestroy(qstream);
}

// ======================================================================== C ABI
extern "C" {

const char* sgc_version(void) { return "streamgc_amd 0.1 (gfx950)"; }

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

int sgc_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

static int global_fail(int code, const std::string& m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = m;
    return code;
}

#define TRY(call, what)                                   \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) {                           \
            int _r = E->hip_fail(_e, what);               \
            std::string _m = E->err;                      \
            E->release();                                 \
            delete E;                                     \
            return global_fail(_r, _m);                   \
        }                                                 \
    } while (0)

int sgc_create(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
               const uint32_t* deg0, sgc_engine** out) {
    if (!cfg || !row_ptr || !deg0 || !out || n == 0) return global_fail(SGC_ERR_INVALID, "sgc_create: bad argument");
    if (cfg->batch_size == 0 || cfg->epoch_size == 0) return global_fail(SGC_ERR_INVALID, "sgc_create: zero batch/epoch");
    const uint64_t nnz = row_ptr[n];
    if (nnz && !col_idx) return global_fail(SGC_ERR_INVALID, "sgc_create: col_idx missing");
    for (uint32_t v = 0; v < n; v++)
        if (row_ptr[v + 1] < row_ptr[v]) return global_fail(SGC_ERR_INVALID, "sgc_create: row_ptr not monotone");
    for (uint64_t j = 0; j < nnz; j++)
        if (col_idx[j] >= n) return global_fail(SGC_ERR_INVALID, "sgc_create: column id out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return global_fail(SGC_ERR_HIP, "sgc_create: no HIP device");

    // slack capacities (A/src/streamgc.cu:373-378)
    std::vector<uint32_t> rp(n + 1, 0);
    uint64_t S = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint64_t d = row_ptr[v + 1] - row_ptr[v];
        uint64_t cap = d + d * cfg->slack_pct / 100;
        if (cap < cfg->min_row_capacity) cap = cfg->min_row_capacity;
        S += cap;
        if (S >= 0xFFFFFFFFull) return global_fail(SGC_ERR_RANGE, "sgc_create: slot count exceeds 32-bit offsets");
        rp[v + 1] = (uint32_t)S;
    }
    // histogram rows for deg0 > threshold, in vertex order (A/src/streamgc.cu:226-250)
    std::vector<uint32_t> hd(n, INVALID), hub_list;
    uint32_t nhd = 0;
    for (uint32_t v = 0; v < n; v++)
        if (deg0[v] > cfg->histogram_threshold) {
            hd[v] = nhd++;
            hub_list.push_back(v);
        }

    sgc_engine* E = new sgc_engine;
    E->cfg = *cfg;
    E->n = n;
    E->S = S;
    E->nhd = nhd;
    TRY(hipGetDevice(&E->device), "device");
    TRY(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking), "stream");
    TRY(hipStreamCreateWithFlags(&E->qstream, hipStreamNonBlocking), "query stream");
    for (int i = 0; i < 2; i++) {
        TRY(hipEventCreateWithFlags(&E->pub_ev[i], hipEventDisableTiming), "event");
        TRY(hipEventCreateWithFlags(&E->read_ev[i], hipEventDisableTiming), "event");
    }
    TRY(hipMalloc(&E->word, (size_t)n * 4), "alloc word");
    TRY(hipMalloc(&E->deg0, (size_t)n * 4), "alloc deg0");
    TRY(hipMalloc(&E->row_ptr, ((size_t)n + 1) * 4), "alloc row_ptr");
    TRY(hipMalloc(&E->col, S * 4), "alloc col");
    TRY(hipMalloc(&E->bitmap, (size_t)n * 8), "alloc bitmap");
    TRY(hipMalloc(&E->hd_map, (size_t)n * 4), "alloc hd_map");
    TRY(hipMalloc(&E->hubs, (size_t)(nhd ? nhd : 1) * 4), "alloc hubs");
    TRY(hipMalloc(&E->hist, (size_t)(nhd ? nhd : 1) * HIST_BINS * 4), "alloc hist");
    TRY(hipMalloc(&E->ovf, (size_t)(cfg->overflow_capacity ? cfg->overflow_capacity : 1) * 8), "alloc overflow");
    TRY(hipMalloc(&E->ctr, C_COUNT * 4), "alloc counters");
    TRY(hipMalloc(&E->snap, (size_t)n * 2 * 2), "alloc snapshot");
    TRY(hipMalloc(&E->dg, sizeof(GraphDev)), "alloc descriptor");
    TRY(hipMalloc(&E->totals, T_COUNT * 8), "alloc totals");
    TRY(hipMalloc(&E->batch, (size_t)cfg->batch_size * sizeof(sgc_edge_update)), "alloc batch");
    TRY(hipMalloc(&E->seg, 2 * 4), "alloc seg");
    E->eg.size = 16;
    while (E->eg.size < 2u * cfg->batch_size) E->eg.size <<= 1;
    E->eg.next_cap = cfg->batch_size;
    TRY(hipMalloc(&E->eg.keys, (size_t)E->eg.size * 8), "alloc edge groups");
    TRY(hipMalloc(&E->eg.head, (size_t)E->eg.size * 4), "alloc edge groups");
    TRY(hipMalloc(&E->eg.next, (size_t)cfg->batch_size * 4), "alloc edge groups");
    TRY(hipMalloc(&E->d_tmp16, 64), "alloc tmp");
    TRY(hipHostMalloc(&E->h_seg, 2 * 4), "pinned seg");
    TRY(hipHostMalloc(&E->h_ctr, C_COUNT * 4), "pinned counters");
    TRY(hipHostMalloc(&E->h_totals, T_COUNT * 8), "pinned totals");

    hipStream_t s = E->stream;
    TRY(hipMemsetAsync(E->word, 0, (size_t)n * 4, s), "init");
    TRY(hipMemsetAsync(E->bitmap, 0, (size_t)n * 8, s), "init");
    TRY(hipMemsetAsync(E->hist, 0, (size_t)(nhd ? nhd : 1) * HIST_BINS * 4, s), "init");
    TRY(hipMemsetAsync(E->ctr, 0, C_COUNT * 4, s), "init");
    TRY(hipMemsetAsync(E->snap, 0, (size_t)n * 4, s), "init");
    TRY(hipMemsetAsync(E->totals, 0, T_COUNT * 8, s), "init");
    TRY(hipMemcpyAsync(E->deg0, deg0, (size_t)n * 4, hipMemcpyHostToDevice, s), "upload deg0");
    TRY(hipMemcpyAsync(E->row_ptr, rp.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, s), "upload row_ptr");
    TRY(hipMemcpyAsync(E->hd_map, hd.data(), (size_t)n * 4, hipMemcpyHostToDevice, s), "upload hd_map");
    if (nhd) TRY(hipMemcpyAsync(E->hubs, hub_list.data(), (size_t)nhd * 4, hipMemcpyHostToDevice, s), "upload hubs");
    {
        // prefill the slack rows on the device from the loaded CSR
        uint32_t *rp0 = nullptr, *col0 = nullptr;
        TRY(hipMalloc(&rp0, ((size_t)n + 1) * 4), "alloc staging");
        hipError_t e2 = hipMalloc(&col0, (nnz ? nnz : 1) * 4);
        if (e2 != hipSuccess) {
            (void)hipFree(rp0);
            TRY(e2, "alloc staging");
        }
        e2 = hipMemcpyAsync(rp0, row_ptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, s);
        if (e2 == hipSuccess && nnz) e2 = hipMemcpyAsync(col0, col_idx, nnz * 4, hipMemcpyHostToDevice, s);
        if (e2 == hipSuccess) e2 = launch_fill_rows(s, n, rp0, col0, E->row_ptr, E->col);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(s);
        (void)hipFree(rp0);
        (void)hipFree(col0);
        TRY(e2, "slack rows");
    }
    GraphDev& g = E->hg;
    g.word = E->word;
    g.deg0 = E->deg0;
    g.row_ptr = E->row_ptr;
    g.col = E->col;
    g.bitmap = E->bitmap;
    g.hd_map = E->hd_map;
    g.hubs = E->hubs;
    g.hist = E->hist;
    g.ovf = E->ovf;
    g.ctr = E->ctr;
    g.snap_write = E->snap;
    g.n = n;
    g.ovf_cap = cfg->overflow_capacity;
    g.nhd = nhd;
    TRY(hipMemcpyAsync(E->dg, &g, sizeof(GraphDev), hipMemcpyHostToDevice, s), "upload descriptor");
    TRY(hipStreamSynchronize(s), "create sync");
    *out = E;
    return SGC_OK;
}

int sgc_destroy(sgc_engine* e) {
    if (!e) return SGC_ERR_INVALID;
    e->release();
    delete e;
    return SGC_OK;
}

int sgc_apply(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device) {
    if (!e || (!updates && count)) return SGC_ERR_INVALID;
    uint64_t done = 0;
    while (done < count) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(e->cfg.batch_size, count - done);
        const sgc_edge_update* src = updates + done;
        if (!on_device) {
            hipError_t r = hipMemcpyAsync(e->batch, src, (size_t)b * sizeof(sgc_edge_update), hipMemcpyHostToDevice,
                                          e->stream);
            if (r != hipSuccess) return e->hip_fail(r, "batch upload");
            src = e->batch;
        }
        int rc = e->apply_device(src, b);
        if (rc) return rc;
        if (!on_device) {
            // the staging buffer is reused by the next batch
            hipError_t r = hipStreamSynchronize(e->stream);
            if (r != hipSuccess) return e->hip_fail(r, "apply");
        }
        done += b;
    }
    hipError_t r = hipStreamSynchronize(e->stream);
    if (r != hipSuccess) return e->hip_fail(r, "apply");
    e->resolve_timing();
    return SGC_OK;
}

int sgc_submit(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device,
               sgc_epoch_stats* epochs_out, uint32_t max_epochs, uint32_t* epochs_done) {
    if (!e || (!updates && count)) return SGC_ERR_INVALID;
    auto t0 = std::chrono::steady_clock::now();
    e->last_ms[2] = 0;
    uint64_t done = 0, since = 0;
    uint32_t ne = 0;
    while (done < count) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(e->cfg.batch_size, count - done);
        const sgc_edge_update* src = updates + done;
        if (!on_device) {
            hipError_t r = hipMemcpyAsync(e->batch, src, (size_t)b * sizeof(sgc_edge_update), hipMemcpyHostToDevice,
                                          e->stream);
            if (r != hipSuccess) return e->hip_fail(r, "batch upload");
            src = e->batch;
        }
        int rc = e->apply_device(src, b);
        if (rc) return rc;
        done += b;
        since += b;
        if (since >= e->cfg.epoch_size) {
            sgc_epoch_stats st{};
            rc = e->epoch(epochs_out && ne < max_epochs ? &st : nullptr);
            if (rc) return rc;
            if (epochs_out && ne < max_epochs) epochs_out[ne] = st;
            ne++;
            since = 0;
        }
        if (!on_device) {
            hipError_t r = hipStreamSynchronize(e->stream);
            if (r != hipSuccess) return e->hip_fail(r, "submit");
        }
    }
    hipError_t r = hipStreamSynchronize(e->stream);
    if (r != hipSuccess) return e->hip_fail(r, "submit");
    e->resolve_timing();
    e->last_ms[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (epochs_done) *epochs_done = ne;
    return SGC_OK;
}

int sgc_epoch(sgc_engine* e, sgc_epoch_stats* out) {
    if (!e) return SGC_ERR_INVALID;
    sgc_epoch_stats st{};
    int r = e->epoch(&st);
    if (r) return r;
    hipError_t h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return e->hip_fail(h, "epoch");
    e->resolve_timing();
    if (out) *out = st;
    return SGC_OK;
}

int sgc_finish(sgc_engine* e, sgc_epoch_stats* out) {
    if (!e) return SGC_ERR_INVALID;
    sgc_epoch_stats st{};
    int r = e->finish(&st);
    if (r) return r;
    hipError_t h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return e->hip_fail(h, "finish");
    e->resolve_timing();
    if (out) *out = st;
    return SGC_OK;
}

int sgc_set_flags(sgc_engine* e, uint32_t flags) {
    if (!e || (flags & ~(SGC_DETERMINISTIC | SGC_REPAIR_EVERY_EPOCH | SGC_EXACT_HISTOGRAM))) return SGC_ERR_INVALID;
    e->cfg.flags = flags;
    return SGC_OK;
}

int sgc_query(sgc_engine* e, const uint32_t* ids, uint32_t count, uint16_t* out_colors) {
    if (!e || (!ids && count) || (!out_colors && count)) return SGC_ERR_INVALID;
    if (count == 0) return SGC_OK;
    std::lock_guard<std::mutex> lk(e->qmu);
    if (e->published < 0) {
        memset(out_colors, 0, (size_t)count * 2);
        return SGC_OK;
    }
    if (count > e->qcap) {
        if (e->d_qids) (void)hipFree(e->d_qids);
        if (e->d_qout) (void)hipFree(e->d_qout);
        e->d_qids = nullptr;
        e->d_qout = nullptr;
        hipError_t r = hipMalloc(&e->d_qids, (size_t)count * 4);
        if (r == hipSuccess) r = hipMalloc(&e->d_qout, (size_t)count * 2);
        if (r != hipSuccess) {
            e->qcap = 0;
            return e->hip_fail(r, "query buffers");
        }
        e->qcap = count;
    }
    const int half = e->published;
    hipError_t r = hipStreamWaitEvent(e->qstream, e->pub_ev[half], 0);
    if (r == hipSuccess) r = hipMemcpyAsync(e->d_qids, ids, (size_t)count * 4, hipMemcpyHostToDevice, e->qstream);
    if (r == hipSuccess) {
        uint32_t grid = (count + 255) / 256;
        k_gather16<<<grid > 1024 ? 1024 : grid, 256, 0, e->qstream>>>(e->snap + (size_t)half * e->n, e->n, e->d_qids,
                                                                       count, e->d_qout);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipEventRecord(e->read_ev[half], e->qstream);
    if (r == hipSuccess) r = hipMemcpyAsync(out_colors, e->d_qout, (size_t)count * 2, hipMemcpyDeviceToHost, e->qstream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->qstream);
    if (r != hipSuccess) return e->hip_fail(r, "query");
    return SGC_OK;
}

static int download(sgc_engine* e, void* dst, const void* src, size_t bytes) {
    if (!e || !dst) return SGC_ERR_INVALID;
    hipError_t r = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    if (r != hipSuccess) return e->hip_fail(r, "download");
    return SGC_OK;
}
static int upload(sgc_engine* e, void* dst, const void* src, size_t bytes) {
    if (!e || !src) return SGC_ERR_INVALID;
    hipError_t r = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    if (r != hipSuccess) return e->hip_fail(r, "upload");
    return SGC_OK;
}

int sgc_download_words(sgc_engine* e, uint32_t* w) { return download(e, w, e ? e->word : nullptr, e ? (size_t)e->n * 4 : 0); }
int sgc_download_colors(sgc_engine* e, uint16_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    std::vector<uint32_t> w(e->n);
    int r = sgc_download_words(e, w.data());
    if (r) return r;
    for (uint32_t i = 0; i < e->n; i++) c[i] = (uint16_t)(w[i] & 0xFFFF);
    return SGC_OK;
}
int sgc_download_snapshot(sgc_engine* e, uint16_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    if (e->published < 0) {
        memset(c, 0, (size_t)e->n * 2);
        return SGC_OK;
    }
    return download(e, c, e->snap + (size_t)e->published * e->n, (size_t)e->n * 2);
}
int sgc_num_slots(sgc_engine* e, uint64_t* s) {
    if (!e || !s) return SGC_ERR_INVALID;
    *s = e->S;
    return SGC_OK;
}
int sgc_download_csr(sgc_engine* e, uint32_t* rp, uint32_t* col) {
    if (!e || !rp || !col) return SGC_ERR_INVALID;
    int r = download(e, rp, e->row_ptr, ((size_t)e->n + 1) * 4);
    if (r) return r;
    return download(e, col, e->col, e->S * 4);
}
int sgc_download_bitmaps(sgc_engine* e, uint64_t* b) { return download(e, b, e ? e->bitmap : nullptr, e ? (size_t)e->n * 8 : 0); }
int sgc_num_hd(sgc_engine* e, uint32_t* k) {
    if (!e || !k) return SGC_ERR_INVALID;
    *k = e->nhd;
    return SGC_OK;
}
int sgc_download_hist(sgc_engine* e, uint32_t* h) {
    if (!e) return SGC_ERR_INVALID;
    if (e->nhd == 0) return SGC_OK;
    return download(e, h, e->hist, (size_t)e->nhd * HIST_BINS * 4);
}
int sgc_upload_words(sgc_engine* e, const uint32_t* w) { return upload(e, e ? e->word : nullptr, w, e ? (size_t)e->n * 4 : 0); }
int sgc_upload_bitmaps(sgc_engine* e, const uint64_t* b) { return upload(e, e ? e->bitmap : nullptr, b, e ? (size_t)e->n * 8 : 0); }
int sgc_upload_hist(sgc_engine* e, const uint32_t* h) {
    if (!e) return SGC_ERR_INVALID;
    if (e->nhd == 0) return SGC_OK;
    return upload(e, e->hist, h, (size_t)e->nhd * HIST_BINS * 4);
}
int sgc_overflow_count(sgc_engine* e, uint32_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    return download(e, c, e->ctr + C_OVF, 4);
}
int sgc_totals_get(sgc_engine* e, sgc_totals* out) {
    if (!e || !out) return SGC_ERR_INVALID;
    int r = download(e, e->h_totals, e->totals, T_COUNT * 8);
    if (r) return r;
    memset(out, 0, sizeof *out);
    out->updates = e->h_totals[T_UPDATES];
    out->scan_slots = e->h_totals[T_SCAN];
    out->hd_endpoints = e->h_totals[T_HD];
    out->recolors = e->h_totals[T_RECOLOR];
    out->recolor_cap_slots = e->h_totals[T_RC_CAP];
    out->recolor_live = e->h_totals[T_RC_LIVE];
    out->overflow_events = e->h_totals[T_OVF_EVENTS];
    out->live_slots = e->h_totals[T_LIVE_SLOTS];
    out->grouped = e->h_totals[T_GROUPED];
    out->epochs = e->epochs;
    out->t1_sweeps = e->t1_sweeps;
    out->t2_runs = e->t2_runs;
    return SGC_OK;
}

int sgc_find_color(sgc_engine* e, uint32_t v, uint16_t* out) {
    if (!e || !out || v >= e->n) return SGC_ERR_INVALID;
    LaunchCtx L = e->ctx();
    hipError_t r = launch_find_color(L, 0, v, e->d_tmp16);
    if (r != hipSuccess) return e->hip_fail(r, "find colour");
    return download(e, out, e->d_tmp16, 2);
}

int sgc_validate(sgc_engine* e, int use_snapshot, uint64_t* violations, uint64_t* uncolored, uint32_t* max_color) {
    if (!e) return SGC_ERR_INVALID;
    if (use_snapshot && e->published < 0) return e->fail(SGC_ERR_STATE, "no snapshot published");
    hipError_t r = hipMemsetAsync(e->ctr + C_VIOL, 0, 3 * 4, e->stream);
    LaunchCtx L = e->ctx();
    const uint16_t* colors = use_snapshot ? e->snap + (size_t)e->published * e->n : nullptr;
    if (r == hipSuccess) r = launch_validate(L, 0, e->n, colors);
    if (r == hipSuccess) r = e->read_ctr();
    if (r != hipSuccess) return e->hip_fail(r, "validate");
    if (violations) *violations = e->h_ctr[C_VIOL];
    if (uncolored) *uncolored = e->h_ctr[C_UNCOL];
    if (max_color) *max_color = e->h_ctr[C_MAXC];
    return SGC_OK;
}

int sgc_history_reserve(sgc_engine* e, uint32_t capacity) {
    if (!e) return SGC_ERR_INVALID;
    (void)hipStreamSynchronize(e->stream);
    if (e->history) (void)hipFree(e->history);
    e->history = nullptr;
    e->hist_cap = e->hist_count = 0;
    if (capacity == 0) return SGC_OK;
    hipError_t r = hipMalloc(&e->history, (size_t)capacity * e->n * 2);
    if (r != hipSuccess) return e->hip_fail(r, "history alloc");
    e->hist_cap = capacity;
    return SGC_OK;
}
int sgc_history_count(sgc_engine* e, uint32_t* c) {
    if (!e || !c) return SGC_ERR_INVALID;
    *c = e->hist_count;
    return SGC_OK;
}

int sgc_validate_history(sgc_engine* e, const uint32_t* eu, const uint32_t* ev, const uint64_t* born,
                         const uint64_t* died, uint64_t m, const uint64_t* pos, uint32_t count, uint64_t* bad_out) {
    if (!e || !pos || !bad_out || count > e->hist_count) return SGC_ERR_INVALID;
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
    if (r == hipSuccess && m) r = hipMemcpyAsync(deu, eu, m * 4, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess && m) r = hipMemcpyAsync(dev, ev, m * 4, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess && m) r = hipMemcpyAsync(dborn, born, m * 8, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess && m) r = hipMemcpyAsync(ddied, died, m * 8, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(dpos, pos, (size_t)count * 8, hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess) r = hipMemsetAsync(dbad, 0, (size_t)count * 8, e->stream);
    if (r == hipSuccess)
        r = launch_validate_history(e->stream, e->history, e->n, deu, dev, dborn, ddied, m, dpos, count, dbad);
    if (r == hipSuccess) r = hipMemcpyAsync(bad_out, dbad, (size_t)count * 8, hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    for (void* p : {(void*)deu, (void*)dev, (void*)dborn, (void*)ddied, (void*)dpos, (void*)dbad})
        if (p) (void)hipFree(p);
    if (r != hipSuccess) return e->hip_fail(r, "validate history");
    return SGC_OK;
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

int sgc_last_timing(sgc_engine* e, double* ms4) {
    if (!e || !ms4) return SGC_ERR_INVALID;
    double upd = 0, ep = 0;
    for (auto& kv : e->ktimes) {
        if (kv.first == "apply") upd += kv.second.ms;
        else ep += kv.second.ms;
    }
    ms4[0] = upd;
    ms4[1] = ep;
    ms4[2] = e->last_ms[2];
    ms4[3] = e->last_ms[3];
    return SGC_OK;
}
int sgc_kernel_timing_enable(sgc_engine* e, int enable) {
    if (!e) return SGC_ERR_INVALID;
    e->resolve_timing();
    e->ktiming = enable != 0;
    if (enable) e->ktimes.clear();
    return SGC_OK;
}
int sgc_kernel_times(sgc_engine* e, char* names, uint32_t stride, double* total_ms, uint64_t* launches, uint32_t cap,
                     uint32_t* count) {
    if (!e || !count) return SGC_ERR_INVALID;
    e->resolve_timing();
    uint32_t k = 0;
    for (auto& kv : e->ktimes) {
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

}  // extern "C"
