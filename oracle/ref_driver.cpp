// ref_driver.cpp -- command-line driver around the reference's own host C++
// (A/src/host/stream_gen.cpp, A/src/host/graph_loader.cpp, compiled from
// where they lie by oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE:
// it writes golden vectors for tests/golden; nothing in the product uses it.
//
//   ref_streams mtx   <graph.mtx> <out.csr>
//   ref_streams init  <in.csr> <out.pairs>
//   ref_streams bench <in.csr> <size> <pct> <seed> <out.updates>
//
// .csr    : u32 'SGCR', u32 n, u64 nnz, u32 row_ptr[n+1], u32 col[nnz], u32 deg[n]
// .pairs  : u64 count, then count x (u32 u, u32 v)
// .updates: u64 count, then count x 16-byte records laid out like the
//           reference EdgeUpdate (A/include/types.h:16-21), padding zeroed as
//           the reference batch fill does (A/src/streamgc.cu:832-837).

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

// Interface of the reference host sources (declared as A/src/streamgc.cu:22-58 does).
struct CSRGraph {
    uint32_t num_vertices;
    uint64_t num_edges;
    uint32_t* row_ptr;
    uint32_t* col_idx;
    uint32_t* degree;
};
CSRGraph load_graph_csr(const char* filename);
struct StreamConfig {
    uint64_t stream_size;
    uint32_t insert_pct;
    uint64_t random_seed;
};
std::vector<std::pair<uint32_t, uint32_t>> generate_init_stream(const uint32_t*, const uint32_t*, const uint32_t*,
                                                                 uint32_t);
std::vector<std::pair<uint32_t, uint32_t>> generate_bench_stream_edges(const uint32_t*, const uint32_t*,
                                                                        const uint32_t*, uint32_t,
                                                                        const StreamConfig&, std::vector<uint8_t>&);

static void die(const char* m) {
    fprintf(stderr, "ref_streams: %s\n", m);
    exit(2);
}

static void write_csr(const char* path, uint32_t n, uint64_t nnz, const uint32_t* rp, const uint32_t* col,
                      const uint32_t* deg) {
    FILE* f = fopen(path, "wb");
    if (!f) die("cannot write csr");
    uint32_t magic = 0x52434753u;  // "SGCR"
    fwrite(&magic, 4, 1, f);
    fwrite(&n, 4, 1, f);
    fwrite(&nnz, 8, 1, f);
    fwrite(rp, 4, (size_t)n + 1, f);
    fwrite(col, 4, nnz, f);
    fwrite(deg, 4, n, f);
    fclose(f);
}

struct Csr {
    uint32_t n = 0;
    uint64_t nnz = 0;
    std::vector<uint32_t> rp, col, deg;
};

static Csr read_csr(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) die("cannot read csr");
    Csr g;
    uint32_t magic = 0;
    if (fread(&magic, 4, 1, f) != 1 || magic != 0x52434753u) die("bad csr magic");
    if (fread(&g.n, 4, 1, f) != 1 || fread(&g.nnz, 8, 1, f) != 1) die("short csr");
    g.rp.resize((size_t)g.n + 1);
    g.col.resize(g.nnz);
    g.deg.resize(g.n);
    if (fread(g.rp.data(), 4, g.rp.size(), f) != g.rp.size()) die("short csr");
    if (g.nnz && fread(g.col.data(), 4, g.nnz, f) != g.nnz) die("short csr");
    if (fread(g.deg.data(), 4, g.n, f) != g.n) die("short csr");
    fclose(f);
    return g;
}

int main(int argc, char** argv) {
    if (argc < 2) die("usage");
    if (!strcmp(argv[1], "mtx") && argc == 4) {
        CSRGraph g = load_graph_csr(argv[2]);
        write_csr(argv[3], g.num_vertices, g.num_edges, g.row_ptr, g.col_idx, g.degree);
        return 0;
    }
    if (!strcmp(argv[1], "init") && argc == 4) {
        Csr g = read_csr(argv[2]);
        auto e = generate_init_stream(g.rp.data(), g.col.data(), g.deg.data(), g.n);
        FILE* f = fopen(argv[3], "wb");
        if (!f) die("cannot write pairs");
        uint64_t c = e.size();
        fwrite(&c, 8, 1, f);
        for (auto& p : e) {
            fwrite(&p.first, 4, 1, f);
            fwrite(&p.second, 4, 1, f);
        }
        fclose(f);
        return 0;
    }
    if (!strcmp(argv[1], "bench") && argc == 7) {
        Csr g = read_csr(argv[2]);
        StreamConfig sc;
        sc.stream_size = strtoull(argv[3], nullptr, 10);
        sc.insert_pct = (uint32_t)strtoul(argv[4], nullptr, 10);
        sc.random_seed = strtoull(argv[5], nullptr, 10);
        std::vector<uint8_t> types;
        auto e = generate_bench_stream_edges(g.rp.data(), g.col.data(), g.deg.data(), g.n, sc, types);
        FILE* f = fopen(argv[6], "wb");
        if (!f) die("cannot write updates");
        uint64_t c = e.size();
        fwrite(&c, 8, 1, f);
        for (size_t i = 0; i < e.size(); i++) {
            unsigned char rec[16];
            memset(rec, 0, sizeof rec);
            memcpy(rec, &e[i].first, 4);
            memcpy(rec + 4, &e[i].second, 4);
            rec[8] = types[i];
            fwrite(rec, 16, 1, f);
        }
        fclose(f);
        return 0;
    }
    die("usage");
    return 2;
}
