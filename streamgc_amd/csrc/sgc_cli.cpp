// sgc_cli.cpp -- streaming_gc: the reference's command line over the C ABI.
//
// Same flags, phases and "[StreamGC] ..." report lines as the reference's
// main (A/src/streamgc.cu:653-1581), so its benchmark scripts
// (A/test/benchmark/bench_throughput.sh, bench_color_count.sh) parse this
// binary's output unchanged:
//   Phase 1: the init stream, then drain, Tier-1 and Tier-2 ("Init").
//   Phase 2: the churn stream, then drain, Tier-1 and Tier-2 ("Bench");
//            bench time = stream + that repair.
//   --validate checks the live colouring against the rows; --output-dir
//   writes .coloring, .coloring.txt and .coloring.meta.
// Extra flags: --repair-every-epoch, --exact-histogram, --deterministic, and
// --ablation force-bitmap|force-histogram|no-overflow|live-deg (the
// reference builds each ablation as its own binary).
// Unknown options are ignored, as the reference ignores them; --log-dir,
// --profile-hw and --init-only are accepted and, as there, change nothing.
// One GPU: --num-gpus other than 1 is an error.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "streamgc.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

int die(const char* what, int rc, sgc_engine* e = nullptr) {
    fprintf(stderr, "[StreamGC] ERROR: %s (%d): %s\n", what, rc, sgc_last_error(e));
    return 2;
}

void ensure_dir(const std::string& path) {
    for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1))
        mkdir(path.substr(0, p).c_str(), 0755);
    mkdir(path.c_str(), 0755);
}

// The end-of-phase repair lines (A/src/streamgc.cu:1237,1310,1371,1381).
void report_repair(const char* phase, const sgc_epoch_stats& st) {
    printf("[StreamGC] %s overflow drain: %u local + 0 ghost attempted, %u failed\n", phase, st.overflow_drained,
           st.overflow_failed);
    if (st.t1_conflicts_last == 0)
        printf("[StreamGC] %s repair: converged after %u iteration(s)\n", phase, st.t1_iterations);
    else
        printf("[StreamGC] %s GPU repair: plateau at %u conflicts, deferring to CPU\n", phase, st.t1_conflicts_last);
    if (st.t2_ran) printf("[StreamGC] %s CPU repair: fixed %u vertices\n", phase, st.t2_fixed);
    fflush(stdout);
}

struct Args {
    const char* graph = nullptr;
    uint64_t stream = 1000000;
    uint32_t insert_pct = 50;
    uint64_t seed = 42;
    uint32_t num_gpus = 1;
    const char* hw_tag = nullptr;
    const char* output_dir = nullptr;
    uint32_t epoch_size = 0, histogram_threshold = 0;
    bool validate = false;
    uint32_t flags = 0;
};

bool parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; i++) {
        const char* s = argv[i];
        const bool val = i + 1 < argc;
        if (!strcmp(s, "--graph") && val) a.graph = argv[++i];
        else if (!strcmp(s, "--stream") && val) a.stream = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(s, "--insert-pct") && val) a.insert_pct = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(s, "--seed") && val) a.seed = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(s, "--num-gpus") && val) a.num_gpus = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(s, "--hw-tag") && val) a.hw_tag = argv[++i];
        else if (!strcmp(s, "--output-dir") && val) a.output_dir = argv[++i];
        else if (!strcmp(s, "--log-dir") && val) ++i;   // accepted; nothing is logged separately
        else if (!strcmp(s, "--epoch-size") && val) a.epoch_size = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(s, "--histogram-threshold") && val) a.histogram_threshold = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(s, "--validate")) a.validate = true;
        else if (!strcmp(s, "--repair-every-epoch")) a.flags |= SGC_REPAIR_EVERY_EPOCH;
        else if (!strcmp(s, "--exact-histogram")) a.flags |= SGC_EXACT_HISTOGRAM;
        else if (!strcmp(s, "--deterministic")) a.flags |= SGC_DETERMINISTIC;
        else if (!strcmp(s, "--ablation") && val) {   // the reference's abl-* builds (Makefile:58-78)
            const char* k = argv[++i];
            if (!strcmp(k, "force-bitmap")) a.flags |= SGC_ABL_FORCE_BITMAP;
            else if (!strcmp(k, "force-histogram")) a.flags |= SGC_ABL_FORCE_HISTOGRAM;
            else if (!strcmp(k, "no-overflow")) a.flags |= SGC_ABL_NO_OVERFLOW;
            else if (!strcmp(k, "live-deg")) a.flags |= SGC_ABL_LIVE_DEGREE;
            else {
                fprintf(stderr, "unknown ablation: %s\n", k);
                return false;
            }
        }
    }
    return a.graph != nullptr;
}

// Violations over the distinct live edges of the rows and uncoloured
// vertices with a live edge, with the reference's report lines
// (A/src/streamgc.cu:448-556).  Returns true when there is no violation.
bool validate_rows(sgc_engine* e, uint32_t n, const std::vector<uint16_t>& color) {
    uint64_t slots = 0;
    sgc_num_slots(e, &slots);
    std::vector<uint32_t> rp(n + 1), col(slots);
    sgc_download_csr(e, rp.data(), col.data());
    std::unordered_set<uint64_t> edges;
    std::vector<uint8_t> has(n, 0);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = rp[u]; j < rp[u + 1]; j++) {
            const uint32_t v = col[j];
            if (v == 0xFFFFFFFFu || v >= n) continue;
            edges.insert(((uint64_t)std::min(u, v) << 32) | std::max(u, v));
            has[u] = 1;
        }
    uint32_t violations = 0, uncolored = 0;
    for (uint64_t ed : edges) {
        const uint32_t u = (uint32_t)(ed >> 32), v = (uint32_t)ed;
        if (color[u] == color[v] && color[u] != 0) violations++;
    }
    for (uint32_t u = 0; u < n; u++) uncolored += color[u] == 0 && has[u];
    if (violations)
        fprintf(stderr, "[StreamGC] COLORING INVALID: %u violations, %u uncolored-with-edges (edges: %zu)\n",
                violations, uncolored, edges.size());
    else
        printf("[StreamGC] Coloring valid: 0 violations, %u uncolored-with-edges (edges: %zu)\n", uncolored,
               edges.size());
    const uint16_t maxc = n ? *std::max_element(color.begin(), color.end()) : 0;
    printf("[StreamGC] Colors used: %u\n", (unsigned)maxc);
    return violations == 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a)) {
        fprintf(stderr, "Usage: %s --graph <path.mtx> [options]\n", argv[0]);
        return 1;
    }
    if (a.num_gpus != 1) {
        fprintf(stderr, "[StreamGC] this build drives one GPU (asked for %u)\n", a.num_gpus);
        return 1;
    }
    if (sgc_device_count() < 1) return die("no HIP device", SGC_ERR_HIP);
    std::string hw;
    if (a.hw_tag) {
        hw = a.hw_tag;
    } else {   // the device name with spaces as underscores, as the reference tags runs
        char name[256];
        hw = sgc_device_name(0, name, sizeof name) == SGC_OK ? name : "unknown";
        std::replace(hw.begin(), hw.end(), ' ', '_');
    }
    printf("[StreamGC] Hardware: %s, GPUs: 1\n", hw.c_str());
    sgc_csr* csr = nullptr;
    int rc = sgc_csr_load_mtx(a.graph, &csr);
    if (rc) return die("load graph", rc);
    uint32_t n = 0, maxdeg = 0;
    uint64_t nnz = 0;
    sgc_csr_info(csr, &n, &nnz, &maxdeg);
    std::vector<uint32_t> rp(n + 1), col(std::max<uint64_t>(nnz, 1)), deg(std::max<uint32_t>(n, 1));
    sgc_csr_export(csr, rp.data(), col.data(), deg.data());
    sgc_csr_free(csr);

    sgc_config cfg;
    sgc_default_config(&cfg);
    if (a.epoch_size) cfg.epoch_size = a.epoch_size;
    if (a.histogram_threshold) cfg.histogram_threshold = a.histogram_threshold;
    cfg.flags = a.flags;
    printf("[StreamGC] Config: epoch_size=%u, histogram_threshold=%u\n", cfg.epoch_size, cfg.histogram_threshold);
    sgc_engine* e = nullptr;
    rc = sgc_create(&cfg, n, rp.data(), col.data(), deg.data(), &e);
    if (rc) return die("create engine", rc);
    uint32_t nhd = 0;
    sgc_num_hd(e, &nhd);
    printf("[StreamGC] GPU 0 partition allocated: %u owned, 0 ghost, %llu local edges, %u high-degree\n", n,
           (unsigned long long)nnz, nhd);

    // Phase 1: init stream
    std::vector<sgc_edge_update> init(std::max<uint64_t>(nnz / 2, 1));
    uint64_t ninit = 0;
    rc = sgc_init_stream(n, rp.data(), col.data(), deg.data(), init.data(), &ninit);
    if (rc) return die("init stream", rc, e);
    printf("[StreamGC] Starting Phase 1: Init stream (%llu edges)...\n", (unsigned long long)ninit);
    fflush(stdout);
    auto t0 = Clock::now();
    rc = sgc_submit(e, init.data(), ninit, 0, nullptr, 0, nullptr);
    if (rc) return die("init submit", rc, e);
    const double init_ms = ms_since(t0);
    printf("[StreamGC] Init complete: %.2f ms\n", init_ms);
    sgc_epoch_stats st{};
    rc = sgc_finish(e, &st);
    if (rc) return die("init repair", rc, e);
    report_repair("Init", st);

    double stream_ms = 0, total_ms = 0, repair_ms = 0, kernel_ms = 0, epoch_ms = 0;
    {
        std::vector<sgc_edge_update> bench(std::max<uint64_t>(a.stream, 1));
        rc = sgc_churn_stream(n, rp.data(), col.data(), a.stream, a.insert_pct, a.seed, bench.data());
        if (rc) return die("bench stream", rc, e);
        printf("[StreamGC] Starting Phase 2: Bench stream (%llu events)...\n", (unsigned long long)a.stream);
        fflush(stdout);
        t0 = Clock::now();
        rc = sgc_submit(e, bench.data(), a.stream, 0, nullptr, 0, nullptr);
        if (rc) return die("bench submit", rc, e);
        stream_ms = ms_since(t0);
        double ms4[4] = {0, 0, 0, 0};
        sgc_last_timing(e, ms4);
        kernel_ms = ms4[0];
        epoch_ms = ms4[1];
        printf("[StreamGC] Bench stream complete: %.2f ms\n", stream_ms);
        rc = sgc_finish(e, &st);
        if (rc) return die("bench repair", rc, e);
        total_ms = ms_since(t0);
        report_repair("Bench", st);
        repair_ms = total_ms - stream_ms;
        printf("[StreamGC] Init time: %.2f ms\n", init_ms);
        printf("[StreamGC] Bench time: %.2f ms (stream: %.2f ms, repair: %.2f ms)\n", total_ms, stream_ms, repair_ms);
        if (total_ms > 0)
            printf("[StreamGC] Throughput: %.2f M updates/sec (effective: %.2f M updates/sec)\n",
                   a.stream / stream_ms / 1000.0, a.stream / total_ms / 1000.0);
    }
    fflush(stdout);

    std::vector<uint16_t> colors(std::max<uint32_t>(n, 1), 0);
    rc = sgc_download_colors(e, colors.data());
    if (rc) return die("download colours", rc, e);
    colors.resize(n);
    const uint32_t maxc = n ? *std::max_element(colors.begin(), colors.end()) : 0;
    if (a.validate && !validate_rows(e, n, colors)) fprintf(stderr, "[StreamGC] VALIDATION FAILED\n");

    if (a.output_dir) {
        ensure_dir(a.output_dir);
        std::string gp(a.graph);
        const size_t slash = gp.rfind('/'), dot = gp.rfind('.');
        const size_t from = slash == std::string::npos ? 0 : slash + 1;
        const std::string name = gp.substr(from, dot == std::string::npos || dot < from ? std::string::npos : dot - from);
        char path[1024];
        snprintf(path, sizeof path, "%s/streamgc_1gpu_stream%llu_%upct.coloring", a.output_dir,
                 (unsigned long long)a.stream, a.insert_pct);
        rc = sgc_write_coloring(path, colors.data(), n);
        if (rc) return die("write coloring", rc, e);
        printf("[StreamGC] Coloring written: %s (%u vertices)\n", path, n);
        printf("[StreamGC] Running static greedy coloring (CPU reference)...\n");
        uint32_t greedy = 0;
        sgc_static_greedy(n, rp.data(), col.data(), &greedy);
        printf("[StreamGC] Static greedy: %u colors\n", greedy);
        sgc_run_meta m{};
        m.graph = name.c_str();
        m.hw_tag = hw.c_str();
        m.num_vertices = n;
        m.num_colors_used = maxc;
        m.num_edges = nnz;
        m.static_greedy_colors = greedy;
        m.insert_pct = a.insert_pct;
        m.stream_size = a.stream;
        m.seed = a.seed;
        m.init_time_ms = init_ms;
        m.bench_time_ms = total_ms;
        m.bench_stream_ms = stream_ms;
        m.repair_time_ms = repair_ms;
        m.local_kernel_ms = kernel_ms;
        m.epoch_maintenance_ms = epoch_ms;
        const std::string meta = std::string(path) + ".meta";
        rc = sgc_write_meta(meta.c_str(), &m);
        if (rc) return die("write meta", rc, e);
        printf("[StreamGC] Meta written: %s (ratio: %.2f)\n", meta.c_str(), greedy ? (double)maxc / greedy : 0.0);
    }
    sgc_destroy(e);
    printf("[StreamGC] Done.\n");
    return 0;
}
