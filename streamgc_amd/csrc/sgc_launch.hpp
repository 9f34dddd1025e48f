// sgc_launch.hpp -- host-side launch interface of the StreamGC kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgc {

struct GraphDev;

// One launch context: the graph descriptors (device array), per-graph
// active flags (graph-indexed kernels), the vertex space of the graphs taking
// part (vertex-indexed kernels) and totals.
struct LaunchCtx {
    hipStream_t stream = nullptr;
    const GraphDev* graphs = nullptr;  // device array of ng descriptors
    uint32_t ng = 1;
    const uint32_t* active = nullptr;  // device, ng flags (null = all)
    const uint32_t* sbase = nullptr;   // device, ns + 1 prefix of vertex counts of the graphs taking part
    const uint32_t* sgid = nullptr;    // device, their graph ids (null = 0..ns-1)
    uint32_t ns = 1;
    uint32_t V = 0;                    // vertices of the graphs taking part
    const uint32_t* stile = nullptr;   // device, ns + 1 prefix of their 4096-vertex tile counts (ns > 1)
    uint32_t tiles = 0;                // their tiles
    const uint32_t* gbase = nullptr;   // device, ng global-vertex offsets (pool edge keys)
    uint64_t* totals = nullptr;        // device u64 counters
    uint32_t max_blocks = 16384;       // grid cap of per-update and full-sweep launches
    uint32_t sweep_blocks = 2048;      // grid cap of worklist sweeps
    uint32_t long_row = 1024;          // worklist rows with more filled slots go to a whole block
    const uint32_t* list = nullptr;    // device worklist of virtual vertices (null = sweep all)
    const uint32_t* list_count = nullptr;  // device [rows at the front, long rows at the back]
    uint32_t list_cap = 0;
    // Device-resident launch parameters (captured launches of a batched run):
    const uint32_t* dyn = nullptr;     // [offset, count, segment offset] of the batch in the resident stream;
                                       // the launchers' count then only sizes the grid
    const uint32_t* gate = nullptr;    // worklist builds do nothing while *gate == 0
    const uint32_t* dspace = nullptr;  // device vertex space of the graphs taking part ([ns, total, -, -, base[ng + 1],
                                       // gid[ng]], written by a control kernel; replaces sbase/sgid/ns/V)
};
// Words of a device vertex space over up to ng graphs.
inline uint32_t dspace_words(uint32_t ng) { return 4 + 2 * ng + 1; }

// Per-edge grouping of a parallel batch: hash table (size: power of two >=
// 2 x batch) of edge keys with list heads, a next link per update, and the
// list of the batch's edges.
struct EdgeGroups {
    uint64_t* keys = nullptr;
    uint32_t* head = nullptr;
    uint32_t* next = nullptr;
    uint32_t* heads = nullptr;    // next_cap entries: one per edge of the batch (slot, or position | 1 << 31),
                                  // long scans from the front, the others from the back
    uint32_t* nheads = nullptr;   // device words: entries at the front, at the back
    uint4* pack = nullptr;        // next_cap x 4 records: each listed edge's updates in stream order (or null)
    uint32_t size = 0;
    uint32_t next_cap = 0;
};

// Per-edge grouping of a parallel batch (before launch_apply with the same eg).
hipError_t launch_group(const LaunchCtx& L, const void* upd, uint32_t count, int multi, const EdgeGroups* eg);
hipError_t launch_apply(const LaunchCtx& L, const void* upd, uint32_t count, int multi, const EdgeGroups* eg);
hipError_t launch_apply_seg(const LaunchCtx& L, const void* upd, const uint32_t* seg, uint32_t nseg,
                            const uint32_t* perm, int multi);
hipError_t launch_zero_ctr(const LaunchCtx& L, uint32_t idx);
hipError_t launch_max_color(const LaunchCtx& L);
struct HistDev;
// The snapshot of the active graphs into their back buffers; with hist and
// fl_round (a captured round), also into the history slot of each graph of
// the round that keeps a history.
hipError_t launch_snapshot(const LaunchCtx& L, const uint32_t* fl_round = nullptr, const HistDev* hist = nullptr);
// drain_prep: the compaction launch also takes the drain's counter snapshot
// (the launch_drain that follows then passes prepped).
hipError_t launch_compact(const LaunchCtx& L, bool drain_prep = false);
hipError_t launch_drain(const LaunchCtx& L, bool serial, uint32_t max_count, bool prepped = false);
// Worklist of the vertices marked `kind` (or sticky).  `list` holds
// list_words(cap, graphs) u32: the list, then the builder's scratch of
// 4096-vertex tiles (each inside one graph, so up to one more per graph).
hipError_t launch_build_list(const LaunchCtx& L, uint32_t kind, uint32_t* list, uint32_t cap, uint32_t* count);
inline size_t list_tiles(uint32_t cap, uint32_t graphs) { return ((size_t)cap + 4095) / 4096 + graphs; }
inline size_t list_words(uint32_t cap, uint32_t graphs) {
    const size_t tiles = list_tiles(cap, graphs);
    return (size_t)cap + tiles * 4096 + 2 * tiles;
}
hipError_t launch_rebuild(const LaunchCtx& L, bool exact_hist);
// serial: vertex order, one wavefront per graph; fresh: parallel-mode colour
// selection from the exact neighbour colours (0 none, 1 bitmap-mode rows,
// 2 all rows; see tier1_vertex).
hipError_t launch_tier1(const LaunchCtx& L, bool serial, int fresh);
hipError_t launch_t2_check(const LaunchCtx& L);
// Tier-2 on the device for the active graphs: begin clears the pass flags and
// lists the reverse-only neighbours of sticky vertices (rev: cap 12-byte
// entries, rev_count: 2 device words); each pass is the uncoloured pass or a
// conflict pass (flag scan, then the ordered recolouring).
hipError_t launch_t2_begin(const LaunchCtx& L, void* rev, uint32_t* rev_count, uint32_t cap);
hipError_t launch_t2_pass(const LaunchCtx& L, const void* rev, const uint32_t* rev_count, uint32_t cap,
                          bool uncoloured);
// Marks for the vertices a host Tier-2 recoloured: (graph, vertex, old colour) triples.
hipError_t launch_mark_recolored(hipStream_t s, const GraphDev* graphs, const uint32_t* items, uint32_t count);
hipError_t launch_validate(const LaunchCtx& L, uint32_t gid, uint32_t n, const uint16_t* colors);
hipError_t launch_find_color(const LaunchCtx& L, uint32_t gid, uint32_t v, uint16_t* out);
hipError_t launch_validate_history(hipStream_t s, const uint16_t* snaps, uint32_t n, const uint32_t* eu,
                                   const uint32_t* ev, const uint64_t* born, const uint64_t* died, uint64_t m,
                                   const uint64_t* pos, uint32_t count, unsigned long long* bad_out);
// ---------------------------------------------------------------- batched runs
// Control block of a batched run (device u32 words).
enum : uint32_t {
    RC_ROUND = 0,     // next round to run
    RC_ROUNDS,        // rounds in the plan
    RC_STOP,          // 1: a round needs the host; later replays do nothing
    RC_STOP_ROUND,    // the round that stopped
    RC_T1_ANY,        // some graph runs another Tier-1 sweep
    RC_OFF,           // current round: first record in the resident stream
    RC_COUNT,         //   its record count
    RC_SEG,           //   offset of its segment table (deterministic replay)
    RC_T2_ANY,        // some graph runs another device Tier-2 pass
    RC_T2_RAN,        // some graph ran the device Tier-2 this round
    RC_T2_MISSING,    // the round stopped for a Tier-2 check the capture leaves out
    RC_T1_SHORT,      // the round stopped with a graph still iterating Tier-1
    RC_WORDS = 16
};
// Per-graph flag arrays (ng u32 each): taking part in the round, Tier-1
// iterating, parallel sweep next, serial sweep next, Tier-2 check, device
// Tier-2 passes still running, device Tier-2 ran this round.
enum : uint32_t { F_ROUND = 0, F_ITER, F_PAR, F_SER, F_T2, F_T2RUN, F_T2RAN, F_KINDS };
struct RoundState {   // Tier-1 loop state of one graph in the current round
    uint32_t iterating, prev, plateau, iters, last, clean;
};
struct RoundGraph {   // constant per graph during a run
    uint32_t max_iters, plateau_iters, n, t2_passes;
};
struct HistDev {      // a graph's snapshot history
    uint16_t* base;
    uint32_t cap, count;
};
hipError_t launch_round_begin(hipStream_t s, uint32_t* ctl, const uint32_t* tab, const uint8_t* inround, uint32_t ng,
                              uint32_t* fl, RoundState* st, const RoundGraph* rg);
hipError_t launch_t1_control(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                             RoundState* st, const RoundGraph* rg, bool after, bool before, bool deterministic,
                             uint32_t serial_repair, uint32_t serial_conflicts);
// lean: the capture has no Tier-2 section, so a graph that needs the check
// stops the round too (RC_T2_MISSING).
hipError_t launch_t2_prep(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                          const RoundState* st, bool lean);
hipError_t launch_round_end(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl);
// launch_t1_control(after, !before), launch_t2_prep and, when lean, launch_round_end in one launch.
hipError_t launch_t1_finish(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                            RoundState* st, const RoundGraph* rg, bool deterministic, uint32_t serial_repair,
                            uint32_t serial_conflicts, bool lean);
// Device Tier-2 control of a round: 0 after the check (graphs with bad slots
// run it), 1 after the reverse-only list (a graph whose list overflowed stops
// the round), 2 before pass `pass`, 3 after it, 4 at the end (graphs still
// running stop the round; the graphs that ran are checked again).
// Steps 0, 1 and 3 also write into dspace the vertex space of the graphs
// still running Tier-2 (the Tier-2 row kernels of the round launch over it).
hipError_t launch_t2_control(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                             const RoundGraph* rg, int step, uint32_t pass, uint32_t* dspace);
// After the round's snapshot (which also filled the history slots): each
// round graph's snapshot becomes the published one and its history count
// moves on.
hipError_t launch_round_publish(hipStream_t s, GraphDev* graphs, uint32_t ng, const uint32_t* fl_round,
                                uint16_t* const* snap_base, uint32_t* pub_half, HistDev* hist);

hipError_t launch_fill_rows(hipStream_t s, uint32_t n, const uint32_t* rp0, const uint32_t* col0, const uint32_t* rp,
                            uint32_t* col, uint32_t* fill);

// T_rc microbenchmark: `waves` wavefronts, lane 0 of each running `iters`
// dependent version-bumping CAS operations on words[wave x stride].
hipError_t launch_trc(hipStream_t s, uint32_t* words, uint32_t waves, uint32_t stride, uint32_t iters);

}  // namespace sgc
