// sgc_output.cpp -- the reference's on-disk result formats.
//
//   <name>.coloring       raw little-endian u16 colour per vertex
//   <name>.coloring.txt   "# vertices: N", "# colors: C", "# algorithm: streamgc",
//                         then "id colour" per line            (A/src/streamgc.cu:595-642)
//   <name>.coloring.meta  JSON with the reference's keys, in its order
//                                                              (A/src/streamgc.cu:1488-1573)
// Fields this engine has no counterpart for (the multi-GPU classify /
// boundary / ghost stages) are written as 0, and partition_info lists the
// single partition.
#include <algorithm>
#include <cstdio>
#include <string>

#include "streamgc.h"

namespace {

std::string json_escape(const char* s) {
    std::string o;
    for (const char* p = s ? s : ""; *p; p++) {
        const unsigned char c = (unsigned char)*p;
        if (c == '"' || c == '\\') {
            o += '\\';
            o += (char)c;
        } else if (c < 0x20) {
            char buf[8];
            snprintf(buf, sizeof buf, "\\u%04x", c);
            o += buf;
        } else {
            o += (char)c;
        }
    }
    return o;
}

}  // namespace

extern "C" {

int sgc_write_coloring(const char* path, const uint16_t* colors, uint32_t n) {
    if (!path || (!colors && n)) return SGC_ERR_INVALID;
    FILE* f = fopen(path, "wb");
    if (!f) return SGC_ERR_IO;
    const bool ok = fwrite(colors, sizeof(uint16_t), n, f) == n;
    if (fclose(f) != 0 || !ok) return SGC_ERR_IO;
    const std::string txt = std::string(path) + ".txt";
    FILE* t = fopen(txt.c_str(), "w");
    if (!t) return SGC_ERR_IO;
    const uint16_t maxc = n ? *std::max_element(colors, colors + n) : 0;
    int r = fprintf(t, "# vertices: %u\n# colors: %u\n# algorithm: streamgc\n", n, (unsigned)maxc);
    for (uint32_t i = 0; i < n && r >= 0; i++) r = fprintf(t, "%u %u\n", i, (unsigned)colors[i]);
    if (fclose(t) != 0 || r < 0) return SGC_ERR_IO;
    return SGC_OK;
}

int sgc_write_meta(const char* path, const sgc_run_meta* m) {
    if (!path || !m) return SGC_ERR_INVALID;
    FILE* f = fopen(path, "w");
    if (!f) return SGC_ERR_IO;
    const double ratio = m->static_greedy_colors ? (double)m->num_colors_used / m->static_greedy_colors : 0.0;
    const double raw = m->bench_stream_ms > 0 ? m->stream_size / m->bench_stream_ms / 1000.0 : 0.0;
    const double eff = m->bench_time_ms > 0 ? m->stream_size / m->bench_time_ms / 1000.0 : 0.0;
    fprintf(f, "{\n");
    fprintf(f, "  \"algorithm\": \"streamgc\",\n");
    fprintf(f, "  \"graph\": \"%s\",\n", json_escape(m->graph).c_str());
    fprintf(f, "  \"num_vertices\": %u,\n", m->num_vertices);
    fprintf(f, "  \"num_edges\": %llu,\n", (unsigned long long)m->num_edges);
    fprintf(f, "  \"num_colors_used\": %u,\n", m->num_colors_used);
    fprintf(f, "  \"static_greedy_colors\": %u,\n", m->static_greedy_colors);
    fprintf(f, "  \"color_ratio\": %.2f,\n", ratio);
    fprintf(f, "  \"init_time_ms\": %.2f,\n", m->init_time_ms);
    fprintf(f, "  \"bench_time_ms\": %.2f,\n", m->bench_time_ms);
    fprintf(f, "  \"bench_stream_ms\": %.2f,\n", m->bench_stream_ms);
    fprintf(f, "  \"repair_time_ms\": %.2f,\n", m->repair_time_ms);
    fprintf(f, "  \"throughput_mups\": %.2f,\n", raw);
    fprintf(f, "  \"effective_throughput_mups\": %.2f,\n", eff);
    fprintf(f, "  \"classify_ms\": %.2f,\n", 0.0);
    fprintf(f, "  \"local_kernel_ms\": %.2f,\n", m->local_kernel_ms);
    fprintf(f, "  \"boundary_kernel_ms\": %.2f,\n", 0.0);
    fprintf(f, "  \"gpu_sync_ms\": %.2f,\n", m->gpu_sync_ms);
    fprintf(f, "  \"ghost_sync_ms\": %.2f,\n", 0.0);
    fprintf(f, "  \"epoch_maintenance_ms\": %.2f,\n", m->epoch_maintenance_ms);
    fprintf(f, "  \"memcpy_ms\": %.2f,\n", m->memcpy_ms);
    fprintf(f, "  \"query_latency_us\": %.2f,\n", m->query_latency_us);
    fprintf(f, "  \"query_count\": %llu,\n", (unsigned long long)m->query_count);
    fprintf(f, "  \"partition_time_ms\": %.2f,\n", 0.0);
    fprintf(f, "  \"partition_info\": [\n");
    fprintf(f, "    {\"gpu\": 0, \"owned\": %u, \"ghost\": 0, \"edges\": %llu}\n", m->num_vertices,
            (unsigned long long)m->num_edges);
    fprintf(f, "  ],\n");
    fprintf(f, "  \"stream_size\": %llu,\n", (unsigned long long)m->stream_size);
    fprintf(f, "  \"insert_pct\": %u,\n", m->insert_pct);
    fprintf(f, "  \"seed\": %llu,\n", (unsigned long long)m->seed);
    fprintf(f, "  \"num_gpus\": 1,\n");
    const int r = fprintf(f, "  \"hw_tag\": \"%s\"\n}\n", json_escape(m->hw_tag).c_str());
    if (fclose(f) != 0 || r < 0) return SGC_ERR_IO;
    return SGC_OK;
}

}  // extern "C"
