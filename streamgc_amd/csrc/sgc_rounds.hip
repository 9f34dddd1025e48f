// sgc_rounds.hip -- device-side control of a batched run.
//
// A batched run replays one captured round body per round.  Everything that
// changes from round to round lives in device memory and is read by the
// kernels themselves: the round index, the round's slice of the resident
// stream, which graphs take part, the Tier-1 loop state of every graph
// (iterations, plateau, serial or parallel sweep) and the snapshot buffer each
// graph writes next.  The host replays the body and synchronises only when a
// round needs it (a Tier-2 repair, or Tier-1 still running after the captured
// iterations); then the control block says which round stopped, and every
// kernel of the replays after it does nothing.
//
// The Tier-1 control follows run_conflict_repair (A/src/streamgc.cu:1317-1393):
// a graph stops at zero conflicts, at its iteration cap, or after
// plateau_iters sweeps with an unchanged count.
#include <hip/hip_runtime.h>

#include "sgc_device.hpp"
#include "sgc_launch.hpp"

namespace sgc {

constexpr uint32_t CTL_THREADS = 1024;

// Block-wide OR of a flag (single-block control kernels).
__device__ __forceinline__ uint32_t block_any(uint32_t x) {
    __shared__ uint32_t part[CTL_THREADS / WAVE];
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t any = ballot(x != 0) != 0;
    if (lane() == 0) part[w] = any;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t k = 0; k < blockDim.x / WAVE; k++) t |= part[k];
    __syncthreads();
    return t;
}

__device__ __forceinline__ uint32_t* flag(uint32_t* fl, uint32_t ng, uint32_t kind, uint32_t g) {
    return fl + (size_t)kind * ng + g;
}

// Block-wide exclusive prefix sum (single-block control kernels); *total
// gets the sum.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_xscan(uint32_t x, uint32_t* total) {
    __shared__ uint32_t part[CTL_THREADS / WAVE];
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t inc = x;
    for (uint32_t d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, d);
        if (lane() >= d) inc += y;
    }
    if (lane() == WAVE - 1) part[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < blockDim.x / WAVE; k++) {
        if (k < w) before += part[k];
        all += part[k];
    }
    __syncthreads();
    *total = all;
    return before + inc - x;
}

// The vertex space of the graphs whose `kind` flag is set, as the host's
// active sets lay it out: [ns, total, -, -, base[0..ns], gid[0..ns-1]] with
// base the prefix of their vertex counts, graphs in id order.
__device__ __forceinline__ void write_space(const GraphDev* __restrict__ graphs, uint32_t ng, uint32_t* fl,
                                            uint32_t kind, uint32_t* dspace) {
    const uint32_t per = (ng + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = threadIdx.x * per < ng ? threadIdx.x * per : ng;
    const uint32_t g1 = g0 + per < ng ? g0 + per : ng;
    uint32_t cg = 0, cv = 0;
    for (uint32_t g = g0; g < g1; g++)
        if (*flag(fl, ng, kind, g)) {
            cg++;
            cv += graphs[g].n;
        }
    uint32_t tg, tv;
    uint32_t pg = block_xscan(cg, &tg);
    uint32_t pv = block_xscan(cv, &tv);
    uint32_t* base = dspace + 4;
    uint32_t* gid = dspace + 4 + ng + 1;
    for (uint32_t g = g0; g < g1; g++)
        if (*flag(fl, ng, kind, g)) {
            base[pg] = pv;
            gid[pg] = g;
            pg++;
            pv += graphs[g].n;
        }
    if (threadIdx.x == 0) {
        dspace[0] = tg;
        dspace[1] = tv;
        base[tg] = tv;
    }
}

// Start of a round: the graphs taking part, their Tier-1 state, the slice.
// After a stop (or past the last round) every flag is cleared and the slice
// is empty, and the stopped round's Tier-1 state is left for the host.
__global__ __launch_bounds__(CTL_THREADS) void k_round_begin(uint32_t* __restrict__ ctl,
                                                             const uint32_t* __restrict__ tab,
                                                             const uint8_t* __restrict__ inround, uint32_t ng,
                                                             uint32_t* __restrict__ fl, RoundState* __restrict__ st,
                                                             const RoundGraph* __restrict__ rg) {
    const uint32_t r = ctl[RC_ROUND];
    const bool live = ctl[RC_STOP] == 0 && r < ctl[RC_ROUNDS];
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
        const uint32_t in = live ? inround[(size_t)r * ng + g] : 0u;
        const uint32_t it = in && rg[g].max_iters > 0;
        if (live) {
            RoundState s;
            s.prev = 0xFFFFFFFFu;
            s.plateau = 0;
            s.iters = 0;
            s.last = 0;
            s.clean = 0;
            s.iterating = it;
            st[g] = s;
        }
        *flag(fl, ng, F_ROUND, g) = in;
        *flag(fl, ng, F_ITER, g) = it;
        *flag(fl, ng, F_PAR, g) = 0;
        *flag(fl, ng, F_SER, g) = 0;
        *flag(fl, ng, F_T2, g) = 0;
        *flag(fl, ng, F_T2RUN, g) = 0;
        *flag(fl, ng, F_T2RAN, g) = 0;
    }
    if (threadIdx.x == 0) {
        ctl[RC_T2_ANY] = 0;
        ctl[RC_T2_RAN] = 0;
        ctl[RC_OFF] = live ? tab[2 * (size_t)r] : 0u;
        ctl[RC_COUNT] = live ? tab[2 * (size_t)r + 1] : 0u;
        ctl[RC_SEG] = live ? r * 2u * ng : 0u;
        ctl[RC_T1_ANY] = 0;
    }
}

// Between Tier-1 sweeps.  after: fold the last sweep's conflict count into
// each iterating graph's state (stop on zero, cap or plateau).  before: pick
// the next sweep of each graph still iterating -- serial (vertex order, one
// wavefront) when deterministic or when the graph is small and two parallel
// sweeps have left few conflicts, else parallel -- and clear its counter.
__device__ __forceinline__ void t1_control(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs, uint32_t ng,
                                           uint32_t* __restrict__ fl, RoundState* __restrict__ st,
                                           const RoundGraph* __restrict__ rg, int after, int before, int deterministic,
                                           uint32_t serial_repair, uint32_t serial_conflicts) {
    uint32_t any = 0;
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
        uint32_t it = *flag(fl, ng, F_ITER, g);
        if (!it) {
            *flag(fl, ng, F_PAR, g) = 0;
            *flag(fl, ng, F_SER, g) = 0;
            continue;
        }
        RoundState s = st[g];
        SGC_G uint32_t* c = gp(graphs[g].ctr);
        if (after) {
            const uint32_t k = c[C_CONFLICTS];
            s.iters++;
            s.last = k;
            s.clean = k == 0;
            if (k == 0 || s.iters >= rg[g].max_iters) {
                it = 0;
            } else if (k == s.prev) {
                if (++s.plateau >= rg[g].plateau_iters) it = 0;
            } else {
                s.plateau = 0;
            }
            s.prev = k;
            s.iterating = it;
            st[g] = s;
            *flag(fl, ng, F_ITER, g) = it;
        }
        const bool run = before && it;
        const bool serial = deterministic || (rg[g].n <= serial_repair && s.prev <= serial_conflicts && s.iters >= 2);
        *flag(fl, ng, F_PAR, g) = run && !serial;
        *flag(fl, ng, F_SER, g) = run && serial;
        if (run) c[C_CONFLICTS] = 0;
        any |= run;
    }
    any = block_any(any);
    if (threadIdx.x == 0) ctl[RC_T1_ANY] = any;
}
__global__ __launch_bounds__(CTL_THREADS) void k_t1_control(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs,
                                                            uint32_t ng, uint32_t* __restrict__ fl,
                                                            RoundState* __restrict__ st,
                                                            const RoundGraph* __restrict__ rg, int after, int before,
                                                            int deterministic, uint32_t serial_repair,
                                                            uint32_t serial_conflicts) {
    t1_control(ctl, graphs, ng, fl, st, rg, after, before, deterministic, serial_repair, serial_conflicts);
}

// After the captured Tier-1 sweeps: a graph still iterating stops the round
// (the host finishes its repair).  The others need the Tier-2 check unless
// their last sweep was clean and no overflow ever left an edge in one row; in
// a capture without the Tier-2 section (lean) such a graph stops the round.
__device__ __forceinline__ void t2_prep(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs, uint32_t ng,
                                        uint32_t* __restrict__ fl, const RoundState* __restrict__ st, int lean) {
    uint32_t stop = 0, missing = 0;
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
        uint32_t need = 0;
        if (*flag(fl, ng, F_ROUND, g)) {
            if (*flag(fl, ng, F_ITER, g)) {
                stop = 1;
            } else {
                SGC_G uint32_t* c = gp(graphs[g].ctr);
                need = !(st[g].clean && c[C_ASYM] == 0);
                if (need) c[C_T2BAD] = 0;
                missing |= need && lean;
            }
        }
        *flag(fl, ng, F_T2, g) = need;
    }
    stop = block_any(stop);
    missing = block_any(missing);
    if (threadIdx.x == 0 && (stop || missing) && ctl[RC_STOP] == 0) {
        ctl[RC_STOP] = 1;
        ctl[RC_STOP_ROUND] = ctl[RC_ROUND];
        ctl[RC_T2_MISSING] = missing;
        ctl[RC_T1_SHORT] = stop;
    }
}
__global__ __launch_bounds__(CTL_THREADS) void k_t2_prep(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs,
                                                         uint32_t ng, uint32_t* __restrict__ fl,
                                                         const RoundState* __restrict__ st, int lean) {
    t2_prep(ctl, graphs, ng, fl, st, lean);
}

// Device Tier-2 inside a round (k_t2_* kernels), controlled per graph as the
// host's device_tier2 controls it.
__global__ __launch_bounds__(CTL_THREADS) void k_t2_control(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs,
                                                            uint32_t ng, uint32_t* __restrict__ fl,
                                                            const RoundGraph* __restrict__ rg, int step, uint32_t pass,
                                                            uint32_t* __restrict__ dspace) {
    const bool stopped = ctl[RC_STOP] != 0;
    uint32_t any = 0, halt = 0;
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
        uint32_t* run = flag(fl, ng, F_T2RUN, g);
        if (step == 0) {   // the graphs whose check found bad slots
            const uint32_t r = !stopped && *flag(fl, ng, F_T2, g) && gp(graphs[g].ctr)[C_T2BAD] != 0;
            *run = r;
            *flag(fl, ng, F_T2RAN, g) = r;
            if (r) {
                SGC_G uint32_t* c = gp(graphs[g].ctr);
                c[C_T2TOTAL] = 0;
                c[C_REVOVF] = 0;
            }
            any |= r;
        } else if (step == 1) {   // a reverse-only list that did not fit: the host repairs
            if (*run && gp(graphs[g].ctr)[C_REVOVF] != 0) halt = 1;
            any |= *run;
        } else if (step == 2) {
            if (*run) {
                SGC_G uint32_t* c = gp(graphs[g].ctr);
                c[C_T2TODO] = 0;
                c[C_T2FIXED] = 0;
            }
            any |= *run;
        } else if (step == 3) {
            if (*run) {
                SGC_G uint32_t* c = gp(graphs[g].ctr);
                *run = pass + 1 < rg[g].t2_passes && c[C_T2TODO] != 0 && c[C_T2FIXED] != 0;
            }
            any |= *run;
        } else {   // step 4
            if (*run) halt = 1;
            if (*flag(fl, ng, F_T2RAN, g)) gp(graphs[g].ctr)[C_T2BAD] = 0;
        }
    }
    any = block_any(any);
    halt = block_any(halt);
    if (step == 1 && halt)
        for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
            *flag(fl, ng, F_T2RUN, g) = 0;
            *flag(fl, ng, F_T2RAN, g) = 0;
        }
    if (dspace && (step == 0 || step == 1 || step == 3)) {
        __syncthreads();   // this block's flag writes above, read by other threads below
        write_space(graphs, ng, fl, F_T2RUN, dspace);
    }
    if (threadIdx.x == 0) {
        if (halt && !stopped) {
            ctl[RC_STOP] = 1;
            ctl[RC_STOP_ROUND] = ctl[RC_ROUND];
        }
        if (step == 0) ctl[RC_T2_RAN] = any;
        if (step == 1 && halt) ctl[RC_T2_RAN] = 0;
        if (step <= 3) ctl[RC_T2_ANY] = halt ? 0u : any;
    }
}

// End of the repair: a stopped round publishes nothing (its graphs leave the
// round); otherwise the round index advances.
__device__ __forceinline__ void round_end(uint32_t* __restrict__ ctl, uint32_t ng, uint32_t* __restrict__ fl) {
    const uint32_t stopped = ctl[RC_STOP];
    const uint32_t r = ctl[RC_ROUND];
    __syncthreads();
    if (stopped)
        for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) *flag(fl, ng, F_ROUND, g) = 0;
    if (threadIdx.x == 0 && !stopped && r < ctl[RC_ROUNDS]) ctl[RC_ROUND] = r + 1;
}
__global__ __launch_bounds__(CTL_THREADS) void k_round_end(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs,
                                                           uint32_t ng, uint32_t* __restrict__ fl) {
    round_end(ctl, ng, fl);
}

// The three steps after the last captured Tier-1 sweep in one launch: fold
// its counts (k_t1_control after), prepare the Tier-2 check (k_t2_prep) and,
// in a capture without the Tier-2 section, end the round (k_round_end).  Each
// step reads what the one before wrote, after a barrier of the one block.
__global__ __launch_bounds__(CTL_THREADS) void k_t1_finish(uint32_t* __restrict__ ctl, const GraphDev* __restrict__ graphs,
                                                           uint32_t ng, uint32_t* __restrict__ fl,
                                                           RoundState* __restrict__ st,
                                                           const RoundGraph* __restrict__ rg, int deterministic,
                                                           uint32_t serial_repair, uint32_t serial_conflicts,
                                                           int lean) {
    t1_control(ctl, graphs, ng, fl, st, rg, 1, 0, deterministic, serial_repair, serial_conflicts);
    __syncthreads();
    t2_prep(ctl, graphs, ng, fl, st, lean);
    __syncthreads();
    if (lean) round_end(ctl, ng, fl);
}

// The snapshot just written becomes the published one; the next publish
// writes the other buffer.
__global__ void k_round_flip(GraphDev* __restrict__ graphs, uint32_t ng, const uint32_t* __restrict__ fl_round,
                             uint16_t* const* __restrict__ snap_base, uint32_t* __restrict__ pub_half,
                             HistDev* __restrict__ hist) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng || !fl_round[g]) return;
    uint16_t* base = snap_base[g];
    const uint32_t n = graphs[g].n;
    const uint32_t half = graphs[g].snap_write == base ? 0u : 1u;
    pub_half[g] = half;
    graphs[g].snap_write = base + (half ? 0u : (size_t)n);
    if (hist[g].base && hist[g].count < hist[g].cap) hist[g].count++;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_round_begin(hipStream_t s, uint32_t* ctl, const uint32_t* tab, const uint8_t* inround, uint32_t ng,
                              uint32_t* fl, RoundState* st, const RoundGraph* rg) {
    k_round_begin<<<1, CTL_THREADS, 0, s>>>(ctl, tab, inround, ng, fl, st, rg);
    return hipGetLastError();
}
hipError_t launch_t1_control(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                             RoundState* st, const RoundGraph* rg, bool after, bool before, bool deterministic,
                             uint32_t serial_repair, uint32_t serial_conflicts) {
    k_t1_control<<<1, CTL_THREADS, 0, s>>>(ctl, graphs, ng, fl, st, rg, after ? 1 : 0, before ? 1 : 0,
                                           deterministic ? 1 : 0, serial_repair, serial_conflicts);
    return hipGetLastError();
}
hipError_t launch_t1_finish(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                            RoundState* st, const RoundGraph* rg, bool deterministic, uint32_t serial_repair,
                            uint32_t serial_conflicts, bool lean) {
    k_t1_finish<<<1, CTL_THREADS, 0, s>>>(ctl, graphs, ng, fl, st, rg, deterministic ? 1 : 0, serial_repair,
                                          serial_conflicts, lean ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_t2_prep(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                          const RoundState* st, bool lean) {
    k_t2_prep<<<1, CTL_THREADS, 0, s>>>(ctl, graphs, ng, fl, st, lean ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_round_end(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl) {
    k_round_end<<<1, CTL_THREADS, 0, s>>>(ctl, graphs, ng, fl);
    return hipGetLastError();
}
hipError_t launch_t2_control(hipStream_t s, uint32_t* ctl, const GraphDev* graphs, uint32_t ng, uint32_t* fl,
                             const RoundGraph* rg, int step, uint32_t pass, uint32_t* dspace) {
    k_t2_control<<<1, CTL_THREADS, 0, s>>>(ctl, graphs, ng, fl, rg, step, pass, dspace);
    return hipGetLastError();
}
hipError_t launch_round_publish(hipStream_t s, GraphDev* graphs, uint32_t ng, const uint32_t* fl_round,
                                uint16_t* const* snap_base, uint32_t* pub_half, HistDev* hist) {
    k_round_flip<<<(ng + 255) / 256, 256, 0, s>>>(graphs, ng, fl_round, snap_base, pub_half, hist);
    return hipGetLastError();
}

}  // namespace sgc
