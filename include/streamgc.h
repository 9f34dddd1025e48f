/* streamgc.h -- C ABI of the MI355X StreamGC engine (libstreamgc_amd.so).
 *
 * Plain pointers and sizes only.  Every call returns SGC_OK (0) or a negative
 * error code and never exits the process; sgc_last_error() gives the text.
 * Calls on one engine handle are serialised by the caller (one HIP stream per
 * engine); sgc_query may run from another thread because it only reads the
 * published snapshot.
 *
 * Reference interfaces replaced (paths relative to the reference checkout,
 * A/ = algorithms/streamgc/):
 *   EdgeUpdate / UpdateType ........ A/include/types.h:8-21
 *   StreamGCConfig ................. A/include/streamgc.cuh:113-135
 *   GPUPartition + allocate ........ A/include/streamgc.cuh:26-110, A/src/streamgc.cu:144-444
 *   local_update_kernel launch ..... A/src/streamgc.cu:847-854
 *   process_stream ................. A/src/streamgc.cu:809-1053
 *   epoch step ..................... A/src/streamgc.cu:949-1036
 *   drain_and_compact + repair ..... A/src/streamgc.cu:1243-1314, 1317-1393, 1057-1240
 *   check_coloring_valid ........... A/src/streamgc.cu:448-556
 *   point_query .................... A/src/host/query_server.cpp:23-28
 *   load_graph_csr ................. A/src/host/graph_loader.cpp:29-169
 *   generate_init_stream ........... A/src/host/stream_gen.cpp:30-61
 *   generate_bench_stream_edges .... A/src/host/stream_gen.cpp:65-177
 *   static_greedy_coloring ......... A/src/streamgc.cu:559-592
 */
#ifndef STREAMGC_AMD_H
#define STREAMGC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGC_OK 0
#define SGC_ERR_INVALID (-1)  /* bad argument */
#define SGC_ERR_HIP (-2)      /* HIP runtime error (no device, launch failure, ...) */
#define SGC_ERR_NOMEM (-3)    /* device or host allocation failed, or over budget */
#define SGC_ERR_RANGE (-4)    /* a size exceeds what the layout supports */
#define SGC_ERR_IO (-5)       /* file could not be read or written */
#define SGC_ERR_STATE (-6)    /* call not valid in the engine's current state */

/* 16-byte update record, byte-identical to the reference EdgeUpdate.
 * type: 0 = ADD, 1 = DEL (2, 3 are accepted and ignored, as in the reference).
 * pad bytes 0..2 are zero; bytes 3..6 (little-endian u32 at offset 12) carry
 * the graph index in a pool and must be zero for a single-graph engine. */
typedef struct sgc_edge_update {
    uint32_t u;
    uint32_t v;
    uint8_t type;
    uint8_t pad[7];
} sgc_edge_update;

/* Engine flags. */
#define SGC_DETERMINISTIC (1u << 0)      /* one wavefront, updates in stream order, repair in vertex order */
#define SGC_REPAIR_EVERY_EPOCH (1u << 1) /* epoch = compact, drain, Tier-1, Tier-2, publish (proper snapshots) */
#define SGC_EXACT_HISTOGRAM (1u << 2)    /* hub histograms recounted wherever bitmaps are rebuilt */
#define SGC_FULL_SWEEPS (1u << 3)        /* maintenance sweeps visit every vertex (no dirty tracking); same results */
#define SGC_HOST_TIER2 (1u << 4)         /* Tier-2 repair on the host (the reference's placement); same results */
/* The reference's compile-time ablations (Makefile:58-78), here runtime flags: */
#define SGC_ABL_FORCE_BITMAP (1u << 5)    /* STREAMGC_ABLATION_FORCE_BITMAP: every vertex selects from its bitmap
                                             (A/include/color_select.cuh:98-100) */
#define SGC_ABL_FORCE_HISTOGRAM (1u << 6) /* STREAMGC_ABLATION_FORCE_HISTOGRAM: the default dispatch, as in the
                                             reference (:101-107) */
#define SGC_ABL_NO_OVERFLOW (1u << 7)     /* STREAMGC_ABLATION_NO_OVERFLOW: a full row drops the edge, nothing is
                                             buffered or counted (A/src/kernels/local_update.cu:28-38) */
#define SGC_ABL_LIVE_DEGREE (1u << 8)     /* STREAMGC_ABLATION_LIVE_DEG_PRIORITY: Tier-1 takes a vertex's own
                                             priority from its live row count, its neighbours' from deg0
                                             (A/src/kernels/conflict_repair.cu:46-58) */
#define SGC_ALL_FLAGS 0x1FFu

/* Engine configuration; the defaults (sgc_default_config) are the reference's. */
typedef struct sgc_config {
    uint32_t epoch_size;          /* E: epoch step after a batch once >= E updates (reference 32768 effective) */
    uint32_t batch_size;          /* host batch (reference BATCH_SIZE 32768, A/src/streamgc.cu:785) */
    uint32_t histogram_threshold; /* deg0 > threshold -> histogram colour selection (64) */
    uint32_t slack_pct;           /* row capacity d + d*slack/100 (300) */
    uint32_t min_row_capacity;    /* row capacity floor (8) */
    uint32_t overflow_capacity;   /* overflow entries kept (1 << 18) */
    uint32_t max_repair_iters;    /* Tier-1 iteration cap (20) */
    uint32_t plateau_iters;       /* Tier-1 stops after this many equal counts (3) */
    uint32_t tier2_passes;        /* Tier-2 pass cap (30) */
    uint32_t flags;               /* SGC_* flags */
} sgc_config;

/* What one epoch step (or the end-of-phase repair) did. */
typedef struct sgc_epoch_stats {
    uint32_t epoch_index;
    uint32_t overflow_count;     /* raw overflow counter at the drain (may exceed the capacity) */
    uint32_t overflow_drained;   /* min(count, capacity) entries retried */
    uint32_t overflow_failed;    /* retried entries that still found no slot */
    uint32_t t1_iterations;
    uint32_t t1_conflicts_last;  /* conflicts counted by the last Tier-1 sweep (0 = converged) */
    uint32_t t2_fixed;           /* vertices recoloured by Tier-2 */
    uint32_t t2_remaining;       /* row slots still conflicting or uncoloured after repair (0 = proper) */
    uint32_t max_color;
    uint32_t snapshot_published;
    uint32_t t2_ran;             /* Tier-2 found work and ran */
    uint32_t reserved;
} sgc_epoch_stats;

/* Running totals over the engine's life (algorithmic counters for the roofline). */
typedef struct sgc_totals {
    uint64_t updates;            /* updates applied (both endpoints in range) */
    uint64_t scan_slots;         /* slots scanned by inserts and deletes (p_u + p_v) */
    uint64_t hd_endpoints;       /* endpoints in histogram mode */
    uint64_t recolors;           /* vertex recolours during updates */
    uint64_t recolor_cap_slots;  /* row slots scanned by those recolours (the filled prefix of the row) */
    uint64_t recolor_live;       /* live neighbours updated by those recolours */
    uint64_t epochs;
    uint64_t t1_sweeps;
    uint64_t t2_runs;
    uint64_t overflow_events;
    uint64_t live_slots;         /* live row slots summed over the rows compacted */
    uint64_t grouped;            /* parallel-batch updates applied in stream order by the wave owning their edge */
    uint64_t sweep_rows[5];      /* rows visited by: compaction, bitmap/histogram rebuild, Tier-1 (parallel
                                    sweeps), Tier-2 check, Tier-1 (sweeps in vertex order) */
    uint64_t sweep_slots[5];     /* row slots read by those sweeps */
} sgc_totals;

typedef struct sgc_engine sgc_engine;

void sgc_default_config(sgc_config* cfg);

/* Launch-shape knobs (no reference counterpart: the autotuner's search space).
 * They change only how work is spread over the GPU, never a result. */
typedef struct sgc_tuning {
    uint32_t long_row;      /* worklist rows with more filled slots are swept by a 256-thread block (default 1024) */
    uint32_t sweep_blocks;  /* grid cap of worklist sweeps, in 256-thread blocks (default 2048) */
    uint32_t apply_blocks;  /* grid cap of the per-update launch and of full sweeps (default 16384) */
    uint32_t serial_repair; /* a graph of at most this many vertices whose last Tier-1 sweep found at most
                               serial_conflicts conflicts, after two parallel sweeps, sweeps next in vertex
                               order on one wavefront (default 65536; deterministic mode: every sweep) */
    uint32_t serial_conflicts; /* default 64 */
    uint32_t fresh_repair;  /* parallel-mode Tier-1 selects from the exact colours of the loser's neighbours
                               gathered by its conflict scan: 0 never (the reference's bitmap/histogram),
                               1 bitmap-mode rows, 2 all rows (default 2) */
} sgc_tuning;
void sgc_default_tuning(sgc_tuning* t);
const char* sgc_last_error(const sgc_engine* e); /* NULL engine: last global error */
const char* sgc_version(void);

/* Device probe: number of HIP devices (0 when none); never fails. */
int sgc_device_count(void);
/* The device's marketing name (cudaGetDeviceProperties().name, which the
 * reference uses as its hardware tag, A/src/streamgc.cu:691-699); the ISA
 * name ("gfx950") when the runtime reports no marketing name. */
int sgc_device_name(int device, char* name, uint32_t cap);

/* T_rc (A/benchmark/microbench/t_rc_atomic.cu:18-62): `waves` wavefronts,
 * lane 0 of each running `iters` dependent CAS operations that bump a
 * vertex word's version and toggle its colour's low bit (a failed CAS takes
 * the value it found).  shared = 0: each wavefront on its own word, a
 * 128-byte line apart; 1: all on one word.  *ns = the median over `runs`
 * (at most 101) of the kernel time per iteration of one wavefront (with one
 * wavefront: the latency of a dependent CAS).  final_words (optional, 2):
 * the first and the last wavefront's word after the last run. */
int sgc_microbench_trc(int device, uint32_t waves, uint32_t shared, uint32_t iters, uint32_t runs, double* ns,
                       uint32_t* final_words);

/* = allocate_gpu_partition (A/src/streamgc.cu:144-444) + the priority fill
 * (:755-767).  Input: symmetric CSR with sorted rows (as the loader makes),
 * deg0 = row lengths (fixes the priority pi(v) = (deg0 << 32) | v). */
int sgc_create(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
               const uint32_t* deg0, sgc_engine** out);
int sgc_destroy(sgc_engine* e);

/* = process_stream (A/src/streamgc.cu:809-1053): batches of batch_size, the
 * per-update kernel, an epoch step after each batch once >= epoch_size.
 * updates on_device = 0: host memory; 1: device memory (already in HBM).
 * epochs_out (optional) receives up to max_epochs per-epoch records. */
int sgc_submit(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device,
               sgc_epoch_stats* epochs_out, uint32_t max_epochs, uint32_t* epochs_done);

/* = one local_update_kernel launch over count updates (no epoch step). */
int sgc_apply(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, int on_device);

/* = one epoch step (A/src/streamgc.cu:949-1036, or the repair-every-epoch order). */
int sgc_epoch(sgc_engine* e, sgc_epoch_stats* out);

/* = drain_and_compact + run_conflict_repair + cpu_serial_repair at the end of
 * a phase (A/src/streamgc.cu:1405-1408, 1420-1423); publishes a snapshot. */
int sgc_finish(sgc_engine* e, sgc_epoch_stats* out);

/* Change the engine flags between phases (e.g. the reference's epoch order
 * during the init stream, repair every epoch during the measured stream).
 * Any SGC_* flag may change; the layout does not. */
int sgc_set_flags(sgc_engine* e, uint32_t flags);
int sgc_set_tuning(sgc_engine* e, const sgc_tuning* t);

/* = point_query (A/src/host/query_server.cpp:23-28) over the published snapshot. */
int sgc_query(sgc_engine* e, const uint32_t* ids, uint32_t count, uint16_t* out_colors);

/* Downloads for validation and output. */
int sgc_download_words(sgc_engine* e, uint32_t* words);          /* n packed (version << 16 | color) */
int sgc_download_colors(sgc_engine* e, uint16_t* colors);        /* n live colours */
int sgc_download_snapshot(sgc_engine* e, uint16_t* colors);      /* n published colours */
int sgc_num_slots(sgc_engine* e, uint64_t* slots);
int sgc_download_csr(sgc_engine* e, uint32_t* row_ptr, uint32_t* col_idx); /* n+1, slots */
int sgc_download_bitmaps(sgc_engine* e, uint64_t* bitmaps);
int sgc_num_hd(sgc_engine* e, uint32_t* num_hd);
int sgc_download_hist(sgc_engine* e, uint32_t* hist);            /* num_hd * 1024 */
int sgc_upload_words(sgc_engine* e, const uint32_t* words);
int sgc_upload_bitmaps(sgc_engine* e, const uint64_t* bitmaps);
int sgc_upload_hist(sgc_engine* e, const uint32_t* hist);
int sgc_totals_get(sgc_engine* e, sgc_totals* out);
int sgc_overflow_count(sgc_engine* e, uint32_t* count);

/* One colour selection for vertex v on the device (FindAvailableColor,
 * A/include/color_select.cuh:93-115), for known-answer tests. */
int sgc_find_color(sgc_engine* e, uint32_t v, uint16_t* out);

/* = check_coloring_valid (A/src/streamgc.cu:448-556) on the engine's own rows:
 * conflicting live edges (counted once per stored slot) and uncoloured
 * vertices with a live slot, from the live colours (snapshot = 0) or the
 * published snapshot (snapshot = 1).  max_color as the reference counts it. */
int sgc_validate(sgc_engine* e, int snapshot, uint64_t* violations, uint64_t* uncolored_with_edges,
                 uint32_t* max_color);

/* Keep a copy of every published snapshot in a device-side history of up to
 * `capacity` snapshots (0 disables).  Copies are device-to-device. */
int sgc_history_reserve(sgc_engine* e, uint32_t capacity);
int sgc_history_count(sgc_engine* e, uint32_t* count);

/* Device-resident update buffers (stream staging in HBM). */
int sgc_device_upload(const void* host, uint64_t bytes, void** dev_out);
int sgc_device_free(void* dev);

/* Independent validator: an edge list with life intervals [born, died) in
 * stream positions (died = UINT64_MAX: alive at the end), checked against each
 * snapshot in the history; snapshot k is checked at stream position pos[k].
 * bad_out[k] = conflicting live edges + uncoloured endpoints of live edges. */
int sgc_validate_history(sgc_engine* e, const uint32_t* eu, const uint32_t* ev, const uint64_t* born,
                         const uint64_t* died, uint64_t m, const uint64_t* pos, uint32_t count,
                         uint64_t* bad_out);

/* Timing of the last sgc_submit / sgc_epoch calls in milliseconds, measured
 * with HIP events on the engine's stream: [0] update kernels, [1] epoch
 * kernels, [2] host-side Tier-2, [3] wall.  Kernel-level per-name totals:
 * sgc_kernel_times fills up to cap (name, total_ms, launches) triples. */
int sgc_last_timing(sgc_engine* e, double* ms4);
int sgc_kernel_timing_enable(sgc_engine* e, int enable);
int sgc_kernel_times(sgc_engine* e, char* names, uint32_t name_stride, double* total_ms, uint64_t* launches,
                     uint32_t cap, uint32_t* count);

/* ---- pools: many graphs resident on one GPU (batched and served modes) ----
 * No reference counterpart: the reference runs one graph per process.  Every
 * phase is one launch over all graphs taking part.  Records carry the graph
 * id in bytes 12..15 (sgc_edge_update.pad[3..6]).  Flags are pool-wide (the
 * sgc_config flags of sgc_pool_add are ignored).  Deterministic pools replay
 * each graph's updates in stream order on one wavefront per graph, so every
 * graph's result is the same alone and in the mix. */
typedef struct sgc_pool sgc_pool;
typedef struct sgc_pool_config {
    uint32_t max_graphs;       /* graph slots */
    uint32_t batch_capacity;   /* largest batch one launch takes (all graphs together) */
    uint32_t flags;            /* SGC_* flags for every graph */
    uint32_t reserved;
    uint64_t hbm_budget;       /* bytes the pool's graphs may occupy (0 = unlimited) */
} sgc_pool_config;

const char* sgc_pool_last_error(const sgc_pool* p);
int sgc_pool_create(const sgc_pool_config* cfg, sgc_pool** out);
int sgc_pool_destroy(sgc_pool* p);
/* Device bytes a graph would occupy (what admission charges against the budget). */
int sgc_pool_graph_bytes(const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* deg0,
                         uint64_t* bytes);
/* Admit a graph: SGC_ERR_NOMEM when the pool is full or the budget would be exceeded. */
int sgc_pool_add(sgc_pool* p, const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                 const uint32_t* deg0, uint32_t* gid);
int sgc_pool_remove(sgc_pool* p, uint32_t gid);
int sgc_pool_usage(sgc_pool* p, uint64_t* used_bytes, uint64_t* budget, uint32_t* graphs);
int sgc_pool_set_flags(sgc_pool* p, uint32_t flags);
int sgc_pool_set_tuning(sgc_pool* p, const sgc_tuning* t);
/* One launch per batch of up to batch_capacity records, no epoch step. */
int sgc_pool_apply(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, int on_device);
/* Graphs whose updates since their last epoch reached their epoch_size. */
int sgc_pool_due(sgc_pool* p, uint32_t* gids, uint32_t cap, uint32_t* count);
/* One epoch step for a set of graphs, in shared launches. */
int sgc_pool_epoch(sgc_pool* p, const uint32_t* gids, uint32_t count, sgc_epoch_stats* out);
/* End-of-phase drain, repair and publish for a set of graphs. */
int sgc_pool_finish(sgc_pool* p, const uint32_t* gids, uint32_t count, sgc_epoch_stats* out);
/* process_stream over a mixed stream: batches, then one shared epoch for the due graphs. */
int sgc_pool_submit(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, int on_device,
                    sgc_epoch_stats* stats_out, uint32_t max_stats, uint32_t* nstats);
int sgc_pool_query(sgc_pool* p, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out_colors);
/* sgc_pool_query without any device allocation, for callers that cannot
 * allocate (tight serving loops, code that forbids hipMalloc).  No counterpart
 * in the reference.  Same arguments, results and errors as sgc_pool_query;
 * the ids and colours are staged in a static device array of
 * SGC_QUERY_NOALLOC_MAX entries (384 KiB, held by each GPU the library runs
 * on until the process ends) instead of the pool's growable buffers.  More
 * than SGC_QUERY_NOALLOC_MAX ids: SGC_ERR_RANGE.  Every call on one GPU
 * shares that array; the library serialises the calls of one process with a
 * lock, and the array keeps the last call's contents. */
#define SGC_QUERY_NOALLOC_MAX 65536u
int sgc_pool_query_noalloc(sgc_pool* p, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out_colors);
int sgc_pool_download_words(sgc_pool* p, uint32_t gid, uint32_t* words);
int sgc_pool_upload_words(sgc_pool* p, uint32_t gid, const uint32_t* words);  /* every sweep then visits every vertex */
/* One resident graph (no reference counterpart). */
typedef struct sgc_graph_info {
    uint32_t n;
    uint32_t num_hd;             /* vertices in histogram mode */
    uint64_t slots;              /* slack CSR capacity */
    uint64_t bytes;              /* device bytes charged to the pool's budget */
    uint32_t epoch_size;
    uint32_t epochs;             /* epoch steps taken */
    uint32_t overflow_events;    /* row-full events ever: edges whose two rows may disagree, or that were dropped */
    int32_t published;           /* published snapshot buffer (-1: none yet) */
    double tier2_ms;             /* host Tier-2 time, all runs */
    uint32_t tier2_runs;
    uint32_t lost;               /* half-edges lost for good (the reference drops them too): overflow entries past
                                    the buffer's capacity, drains that found no room, full rows under the
                                    no-overflow ablation.  An edge lost from both rows is invisible to repair. */
} sgc_graph_info;
int sgc_pool_graph_info(sgc_pool* p, uint32_t gid, sgc_graph_info* out);
/* Per-graph state downloads (as sgc_download_csr / _bitmaps / _hist). */
int sgc_pool_download_csr(sgc_pool* p, uint32_t gid, uint32_t* row_ptr, uint32_t* col_idx);  /* n+1, slots */
int sgc_pool_download_bitmaps(sgc_pool* p, uint32_t gid, uint64_t* bitmaps);
int sgc_pool_download_hist(sgc_pool* p, uint32_t gid, uint32_t* hist);                     /* num_hd * 1024 */
int sgc_pool_download_snapshot(sgc_pool* p, uint32_t gid, uint16_t* colors);
int sgc_pool_validate(sgc_pool* p, uint32_t gid, int snapshot, uint64_t* violations, uint64_t* uncolored_with_edges,
                      uint32_t* max_color);
int sgc_pool_history_reserve(sgc_pool* p, uint32_t gid, uint32_t capacity);
int sgc_pool_history_count(sgc_pool* p, uint32_t gid, uint32_t* count);
int sgc_pool_validate_history(sgc_pool* p, uint32_t gid, const uint32_t* eu, const uint32_t* ev, const uint64_t* born,
                              const uint64_t* died, uint64_t m, const uint64_t* pos, uint32_t count,
                              uint64_t* bad_out);
int sgc_pool_totals(sgc_pool* p, sgc_totals* out);
int sgc_pool_kernel_timing_enable(sgc_pool* p, int enable);
int sgc_pool_kernel_times(sgc_pool* p, char* names, uint32_t name_stride, double* total_ms, uint64_t* launches,
                          uint32_t cap, uint32_t* count);
/* Per-graph epoch size E_g (no reference counterpart; the reference has one
 * --epoch-size for its single graph).  Takes effect at the graph's next batch. */
int sgc_pool_set_epoch_size(sgc_pool* p, uint32_t gid, uint32_t epoch_size);

/* ---- batched pattern: many graphs, one resident stream, captured rounds ----
 * No reference counterpart.  sgc_pool_batch_prepare copies a mixed stream
 * (records tagged with their graph id, stream order kept per graph) to the
 * device and cuts it into rounds: round r gives every graph with records left
 * its next min(E_g - pending, left) records, then every graph of the round
 * takes an epoch step; a graph's last partial epoch is repaired and published
 * like a full one.  Each graph thus sees exactly the batches, epochs and
 * final repair it sees alone (batch = epoch = E_g, then sgc_finish).
 * sgc_pool_batch_run runs the rounds: either one captured hipGraph replayed per
 * round (every per-round parameter, the Tier-1 loop control and the snapshot
 * flip live on the device; the host steps in only for a round that needs a
 * host Tier-2 repair or more Tier-1 sweeps than were captured), or host-driven
 * launches.  The autotuner times each launch-shape candidate on interleaved
 * rounds and keeps the fastest per update (the pool keeps it afterwards).
 * Needs SGC_REPAIR_EVERY_EPOCH and dirty tracking.  Queries of the pool wait
 * while a run is in progress. */
typedef struct sgc_batch_config {
    uint32_t use_graph;        /* 1: replay a captured hipGraph per round; 0: host-driven launches */
    uint32_t t1_sweeps;        /* Tier-1 sweeps captured per round (at most max_repair_iters); a graph still
                                  repairing after them makes the host finish the round.  0: start with 2 and
                                  double them after each round the host finished for that reason */
    uint32_t autotune_rounds;  /* > 0: captured rounds each launch-shape candidate is timed on (hipGraph mode);
                                  rounds the host finished do not count, and a recapture restarts the timing */
    uint32_t t2_passes;        /* device Tier-2 passes captured per round (0: 4); a graph needing more, or a
                                  Tier-2 on the host (SGC_HOST_TIER2), makes the host finish the round */
} sgc_batch_config;
#define SGC_BATCH_CANDIDATES 4
typedef struct sgc_batch_stats {
    uint32_t rounds;           /* rounds in the plan */
    uint32_t host_rounds;      /* rounds the host finished */
    uint32_t candidates;       /* launch-shape candidates timed (0: no tuning) */
    uint32_t chosen;           /* the candidate kept */
    uint64_t updates;
    double ms;                 /* wall time of the run, first launch to the last snapshot published */
    double candidate_ns_per_update[SGC_BATCH_CANDIDATES];   /* over the candidate's samples (0: none) */
    sgc_tuning candidate[SGC_BATCH_CANDIDATES];
    uint32_t candidate_samples[SGC_BATCH_CANDIDATES];         /* captured rounds each candidate was timed on */
    uint32_t tune_rounds;      /* rounds [0, tune_rounds) ran while tuning (included in ms) */
    uint64_t tune_updates;     /* their updates */
    double tune_ms;            /* wall time from the start of the run to the end of tuning */
} sgc_batch_stats;
void sgc_default_batch_config(sgc_batch_config* c);
int sgc_pool_batch_prepare(sgc_pool* p, const sgc_edge_update* updates, uint64_t count, const sgc_batch_config* cfg,
                           uint32_t* rounds);
int sgc_pool_batch_run(sgc_pool* p, sgc_batch_stats* out);
int sgc_pool_batch_release(sgc_pool* p);
/* The same for a single-graph engine: its epochs of E updates replayed as
 * captured rounds (records need no graph id). */
int sgc_batch_prepare(sgc_engine* e, const sgc_edge_update* updates, uint64_t count, const sgc_batch_config* cfg,
                      uint32_t* rounds);
int sgc_batch_run(sgc_engine* e, sgc_batch_stats* out);
int sgc_batch_release(sgc_engine* e);

/* ---- served pattern: a persistent request loop over a pool ----
 * No reference counterpart.  Any thread submits update batches and queries
 * for resident graphs as they arrive; one loop thread packs every arrived
 * request into one batch per launch (each graph filled at most to its own
 * epoch boundary, so each graph sees the epoch boundaries it would see
 * alone), then runs one shared epoch step for the graphs that are due (epoch
 * full, or oldest unpublished update older than max_delay_us).  A request
 * completes when the snapshot holding its last update is published; its
 * update-to-snapshot latency is kept raw.  Admissions wait in arrival order
 * until the pool's HBM budget has room.  While a server runs, its pool is
 * driven only by the server.  SGC_SERVE_PLAIN runs the same loop as the plain
 * per-launch engine would serve the tenants (the baseline of the served
 * pattern). */
typedef struct sgc_server sgc_server;
#define SGC_SERVE_BATCHED 0u  /* continuous batching: every arrived request in one launch, shared epoch steps */
#define SGC_SERVE_PLAIN   1u  /* the plain per-launch engine: one launch per request (oldest arrival first),
                                 one epoch step per due graph; same epoch and deadline rules */
typedef struct sgc_server_config {
    uint32_t max_batch;     /* updates per launch over all tenants (default 32768) */
    uint32_t max_delay_us;  /* epoch for a graph whose oldest unpublished update waited this long (0: full epochs only) */
    uint32_t mode;          /* SGC_SERVE_BATCHED or SGC_SERVE_PLAIN */
    uint32_t reserved;
} sgc_server_config;
typedef struct sgc_server_stats {
    uint64_t updates_applied;
    uint64_t batches;         /* apply launches (each over all tenants with work) */
    uint64_t epoch_steps;     /* shared epoch steps */
    uint64_t graph_epochs;    /* graph epochs inside them */
    uint64_t requests_done;
    uint64_t queries;
    uint64_t queued_updates;  /* submitted, not yet applied */
    double span_us;           /* first submit to last publish */
    double busy_us;           /* loop time in turns that did work, split into: */
    double intake_us;         /*   admissions, removals, routing arrived requests */
    double pack_us;           /*   packing launches on the host */
    double apply_us;          /*   sgc_pool_apply (upload, grouping and per-update kernels) */
    double epoch_us;          /*   sgc_pool_epoch (epoch steps up to the published snapshot) */
} sgc_server_stats;
void sgc_default_server_config(sgc_server_config* c);
int sgc_server_create(sgc_pool* pool, const sgc_server_config* cfg, sgc_server** out);
int sgc_server_destroy(sgc_server* s);   /* publishes everything pending, then stops */
const char* sgc_server_last_error(const sgc_server* s);
/* Queue a graph for admission; sgc_server_admitted waits for its pool id. */
int sgc_server_admit(sgc_server* s, const sgc_config* cfg, uint32_t n, const uint32_t* row_ptr,
                     const uint32_t* col_idx, const uint32_t* deg0, uint64_t* ticket);
int sgc_server_admitted(sgc_server* s, uint64_t ticket, uint32_t timeout_ms, uint32_t* gid);
/* Serve a graph already in the pool (its epochs every epoch_size updates); wait with sgc_server_admitted. */
int sgc_server_attach(sgc_server* s, uint32_t gid, uint32_t epoch_size, uint64_t* ticket);
/* Remove a graph once its pending updates are published. */
int sgc_server_remove(sgc_server* s, uint32_t gid);
/* Copy and queue `count` updates for graph gid (graph id bytes are stamped). */
int sgc_server_submit(sgc_server* s, uint32_t gid, const sgc_edge_update* updates, uint32_t count, uint64_t* ticket);
/* Wait until the ticket's updates are in a published snapshot (SGC_ERR_STATE on timeout). */
int sgc_server_wait(sgc_server* s, uint64_t ticket, uint32_t timeout_ms);
/* Publish every pending update now (partial epochs) and wait for it. */
int sgc_server_flush(sgc_server* s, uint32_t timeout_ms);
/* Colours of the published snapshot; never waits for the loop. */
int sgc_server_query(sgc_server* s, uint32_t gid, const uint32_t* ids, uint32_t count, uint16_t* out_colors);
int sgc_server_get_stats(sgc_server* s, sgc_server_stats* out);
/* Raw latencies in microseconds: which 0 = update-to-snapshot per request, 1 = query. */
int sgc_server_latencies(sgc_server* s, int which, double* out, uint64_t cap, uint64_t* count);
int sgc_server_reset_stats(sgc_server* s);
/* Per published snapshot of graph gid, in publish order: how many of the
 * graph's updates submitted through the server it contains (the updates are
 * applied in submission order, so each watermark is a prefix of them). */
int sgc_server_watermarks(sgc_server* s, uint32_t gid, uint64_t* out, uint64_t cap, uint64_t* count);
/* Open-loop trace replay (no reference counterpart; the reference's feeder,
 * A/src/host/feeder_thread.cpp:28-138, is never started).  Each item arrives
 * at_ns after the call starts: update items are submitted from the calling
 * thread, queries from a second thread, each at its own time, and latencies
 * run from that scheduled arrival (not from when the submission happened).
 * Items must be in non-decreasing at_ns per kind.  Returns after everything is
 * published (a flush), or SGC_ERR_STATE after timeout_ms. */
typedef struct sgc_trace_item {
    uint64_t at_ns;         /* arrival, ns after the replay starts */
    uint64_t offset;        /* first record in `updates` (updates) or first id in `query_ids` (query) */
    uint32_t gid;
    uint32_t count;
    uint32_t kind;          /* 0: updates, 1: query */
    uint32_t reserved;
} sgc_trace_item;
typedef struct sgc_replay_stats {
    double span_us;         /* replay start to the last publish */
    double max_late_us;     /* largest delay of a submission behind its schedule */
    double mean_late_us;
    uint64_t requests;
    uint64_t updates;
    uint64_t queries;
} sgc_replay_stats;
int sgc_server_replay(sgc_server* s, const sgc_trace_item* items, uint64_t count, const sgc_edge_update* updates,
                      const uint32_t* query_ids, uint32_t timeout_ms, sgc_replay_stats* out);

/* ---- the reference's result files (A/src/streamgc.cu:595-642,1469-1573) ----
 * sgc_write_coloring writes <path> (raw u16 per vertex) and <path>.txt (the
 * "# vertices / # colors / # algorithm" header, then "id colour" lines).
 * sgc_write_meta writes the .coloring.meta JSON with the reference's keys in
 * its order; throughput and colour ratio are derived as the reference derives
 * them, multi-GPU stage timings are 0. */
typedef struct sgc_run_meta {
    const char* graph;          /* graph name (file name without directory and extension) */
    const char* hw_tag;
    uint32_t num_vertices;
    uint32_t num_colors_used;
    uint64_t num_edges;         /* CSR entries, as the reference's loader counts them */
    uint32_t static_greedy_colors;
    uint32_t insert_pct;
    uint64_t stream_size;
    uint64_t seed;
    double init_time_ms;
    double bench_time_ms;       /* stream + final repair */
    double bench_stream_ms;
    double repair_time_ms;
    double local_kernel_ms;     /* per-update kernels */
    double gpu_sync_ms;
    double epoch_maintenance_ms;
    double memcpy_ms;
    double query_latency_us;
    uint64_t query_count;
} sgc_run_meta;
int sgc_write_coloring(const char* path, const uint16_t* colors, uint32_t n);
int sgc_write_meta(const char* path, const sgc_run_meta* m);

/* ---- host pipeline (C++17) ---- */

/* Graph container (host CSR, symmetric, sorted rows, no self-loops). */
typedef struct sgc_csr sgc_csr;

/* = load_graph_csr (A/src/host/graph_loader.cpp:29-169). */
int sgc_csr_load_mtx(const char* path, sgc_csr** out);
/* Dedup, drop self-loops, symmetrise, sort rows (the loader's rules). */
int sgc_csr_from_edges(uint32_t n, const uint32_t* u, const uint32_t* v, uint64_t m, sgc_csr** out);
int sgc_csr_info(const sgc_csr* g, uint32_t* n, uint64_t* nnz, uint32_t* max_deg);
int sgc_csr_export(const sgc_csr* g, uint32_t* row_ptr, uint32_t* col_idx, uint32_t* deg);
int sgc_csr_free(sgc_csr* g);

/* Synthetic inputs (no reference counterpart; seeded, thread-count independent).
 * Chung-Lu: power-law expected degrees (exponent gamma, given max degree),
 * exactly m distinct edges, optional random vertex permutation.
 * R-MAT: edge_factor * 2^scale draws with quadrant probabilities a, b, c. */
int sgc_gen_chung_lu(uint32_t n, uint64_t m, double gamma, double max_degree, uint64_t seed, int permute,
                     sgc_csr** out);
int sgc_gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c, uint64_t seed, sgc_csr** out);
/* Core numbers (Batagelj-Zaversnik). */
int sgc_core_numbers(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint32_t* core);
/* Dense-growth stream inside {v : core(v) >= core_min}: ins_per_del inserts
 * (new pairs) per delete (live edges inside the region), on average. */
int sgc_dense_growth_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const uint32_t* core,
                            uint32_t core_min, uint64_t size, uint32_t ins_per_del, uint64_t seed,
                            sgc_edge_update* out);

/* The other update streams of the reference's README.md:83-91 (no code in the
 * reference; specified by that table): "Uniform" inserts `size` uniform
 * non-adjacent vertex pairs, each new; "Hold-out" removes round(m * pct / 100)
 * uniformly drawn edges from the graph (*base_out, same n) and returns them
 * as inserts in random order (two calls: with out = base_out = NULL it only
 * sets *count).  The "Dense" stream is sgc_dense_growth_stream with
 * ins_per_del = 1. */
int sgc_uniform_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint64_t size, uint64_t seed,
                       sgc_edge_update* out);
int sgc_holdout_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint32_t pct, uint64_t seed,
                       sgc_csr** base_out, sgc_edge_update* out, uint64_t* count);

/* = generate_init_stream (A/src/host/stream_gen.cpp:30-61): row_ptr[n]/2 records. */
int sgc_init_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, const uint32_t* deg,
                    sgc_edge_update* out, uint64_t* count);

/* = generate_bench_stream_edges (A/src/host/stream_gen.cpp:65-177); byte-exact. */
int sgc_churn_stream(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint64_t size,
                     uint32_t insert_pct, uint64_t seed, sgc_edge_update* out);

/* = static_greedy_coloring (A/src/streamgc.cu:559-592): max colour of first fit in id order. */
int sgc_static_greedy(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx, uint32_t* max_color);

/* Edge life intervals of a stream over a base graph, for sgc_validate_history.
 * Two calls: with null outputs it reports m; then fills eu, ev, born, died. */
int sgc_edge_intervals(uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                       const sgc_edge_update* stream, uint64_t count, uint64_t* m, uint32_t* eu, uint32_t* ev,
                       uint64_t* born, uint64_t* died);

#ifdef __cplusplus
}
#endif
#endif
