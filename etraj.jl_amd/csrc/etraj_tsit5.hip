// etraj_tsit5.hip -- the adaptive Tsit5 integrator kernel (the reference's
// solve(EnsembleProblem, Tsit5(), adaptive = true, reltol = traj_rtol),
// TrajectorySimulation.jl:332-342) and its launchers.
//
// A file of its own because it is built with machine-level loop-invariant code
// motion off (-mllvm -disable-machine-licm): hoisting the FP64 constants of the
// laser's sin/cos/tanh and of pow out of the step loop pins them in about 80
// VGPRs for the whole kernel. Rematerialised inside the loop they cost scalar
// moves only, and the kernel needs 170 VGPRs instead of 251 (three waves per
// SIMD instead of two).
#include "etraj_engine.hpp"
#include "etraj_device.hpp"
#include "etraj_ts5.hpp"

#include <hip/hip_runtime.h>

#include <atomic>

namespace etraj {

// Per-lane output of one finished electron for kernels whose lanes finish at
// different times (no wave-wide operations inside). Accumulates the scalar
// counters into `acc` for one reduction per wave at the end.
struct LaneAcc {
    double ip = 0.0, iu = 0.0;
    Fix128 ipq{0, 0}, iuq{0, 0};   // ETRAJ_PMD_FIXED
    int64_t failed = 0, bound = 0, steps = 0, attempts = 0, range = 0, restep = 0;
};

__device__ inline void fix_add(Fix128& a, Fix128 b) {
    const uint64_t lo = a.lo + b.lo;
    a.hi = a.hi + b.hi + (lo < a.lo ? 1 : 0);
    a.lo = lo;
}

template <int DIM, int PHASE, int OUT>
__device__ inline void emit_lane(const IntegP& P, const OutP& O, int64_t col_idx, bool failed, const double* r,
                                 const double* p, double phase, double t0, double prob, LaneAcc& acc) {
    Final f = kepler_final<DIM, PHASE>(P.T, P.quirks, r, p, phase, t0);
    if (failed) f.ionized = false;
    int64_t lin = -1;
    bool inb = false;
    double cre = 0.0, cim = 0.0;
    if (f.ionized) {
        int64_t idx[3] = {1, 1, 1};
        inb = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            idx[d] = bin_index(f.pinf[d], P.pmax[d], P.pnum[d]);
            if (idx[d] < 1 || idx[d] > P.pnum[d]) inb = false;
        }
        if (inb) {
            lin = (idx[0] - 1) + (int64_t)P.pnum[0] * ((idx[1] - 1) + (DIM == 3 ? (int64_t)P.pnum[1] * (idx[2] - 1) : 0));
            if (PHASE == ETRAJ_PHASE_CTMC) {
                cre = prob;
            } else {
                double sq = sqrt(prob), sp, cp;
                sincos(f.phase, &sp, &cp);
                cre = sq * cp;
                cim = sq * sp;
            }
        }
    }
    if (failed) acc.failed++;
    else if (!f.ionized) acc.bound++;
    else if (OUT == OUT_PMD && O.pmdq) {
        bool ok;
        const Fix128 q = fix_from_double(prob, ok);
        acc.range += ok ? 0 : 1;
        fix_add(acc.ipq, q);
        if (!inb) fix_add(acc.iuq, q);
    } else {
        acc.ip += prob;
        if (!inb) acc.iu += prob;
    }
    if (OUT == OUT_PMD) {
        if (O.bin_cell) {   // records for the tile binning
            O.bin_cell[col_idx] = inb ? (uint32_t)(O.cell0 + lin) : kNoCell;
            O.bin_amp[col_idx] = make_double2(cre, cim);
        } else if (inb) {
            if (O.pmdq) {
                bool ok1 = true, ok2 = true;
                if (PHASE == ETRAJ_PHASE_CTMC) {
                    fix_atomic_add(O.pmdq + 2 * lin, fix_from_double(cre, ok1));
                } else {
                    fix_atomic_add(O.pmdq + 4 * lin, fix_from_double(cre, ok1));
                    fix_atomic_add(O.pmdq + 4 * lin + 2, fix_from_double(cim, ok2));
                }
                acc.range += (ok1 && ok2) ? 0 : 1;
            } else if (PHASE == ETRAJ_PHASE_CTMC) {
                unsafeAtomicAdd(O.pmd + lin, cre);
            } else {
                unsafeAtomicAdd(O.pmd + 2 * lin, cre);
                unsafeAtomicAdd(O.pmd + 2 * lin + 1, cim);
            }
            if (O.hits) atomicAdd((unsigned long long*)(O.hits + lin), 1ull);
        }
    } else if (OUT == OUT_RECORDS) {
        double* rec = O.records + 12 * col_idx;
        rec[0] = r[0]; rec[1] = r[1]; rec[2] = DIM == 3 ? r[2] : 0.0;
        rec[3] = p[0]; rec[4] = p[1]; rec[5] = DIM == 3 ? p[2] : 0.0;
        rec[6] = phase; rec[7] = f.E;
        rec[8] = f.pinf[0]; rec[9] = f.pinf[1]; rec[10] = f.pinf[2];
        rec[11] = f.phase;
        O.status[col_idx] = failed ? 1 : (f.ionized ? 0 : 2);
    } else {
        Contrib c;
        c.lin = lin;
        c.re = cre;
        c.im = cim;
        c.prob = prob;
        c.status = failed ? 1 : (f.ionized ? (inb ? 0 : 3) : 2);
        c.pad = 0;
        O.contrib[col_idx] = c;
    }
}

// steps64: the launch integrates in FP64 (its steps also count as FP64 steps)
__device__ inline void flush_lane_acc(const OutP& O, const LaneAcc& a, bool steps64) {
    const int64_t nf = wave_sum_i64(a.failed), nb = wave_sum_i64(a.bound), ns = wave_sum_i64(a.steps);
    const int64_t nr = wave_sum_i64(a.range), nq = wave_sum_i64(a.restep), na = wave_sum_i64(a.attempts);
    if (O.pmdq) {
        const Fix128 ip = fix_wave_sum(a.ipq), iu = fix_wave_sum(a.iuq);
        if (lane_id() == 0) {
            fix_atomic_add(O.scalq + 4 * acc_slot() + 0, ip);
            fix_atomic_add(O.scalq + 4 * acc_slot() + 2, iu);
        }
    } else {
        const double ip = wave_sum_f64(a.ip), iu = wave_sum_f64(a.iu);
        if (lane_id() == 0) {
            if (ip != 0.0) unsafeAtomicAdd(O.scal + 2 * acc_slot() + 0, ip);
            if (iu != 0.0) unsafeAtomicAdd(O.scal + 2 * acc_slot() + 1, iu);
        }
    }
    if (lane_id() == 0) {
        int64_t* cs = O.cnt + CNT_N * acc_slot();
        if (nf) atomicAdd((unsigned long long*)(cs + CNT_FAILED), (unsigned long long)nf);
        if (nb) atomicAdd((unsigned long long*)(cs + CNT_BOUND), (unsigned long long)nb);
        if (ns) atomicAdd((unsigned long long*)(cs + CNT_STEPS), (unsigned long long)ns);
        if (ns && steps64) atomicAdd((unsigned long long*)(cs + CNT_STEPS64), (unsigned long long)ns);
        if (nr) atomicAdd((unsigned long long*)(cs + CNT_FIXRANGE), (unsigned long long)nr);
        if (nq) atomicAdd((unsigned long long*)(cs + CNT_RESTEP), (unsigned long long)nq);
        if (na) atomicAdd((unsigned long long*)(cs + CNT_ATTEMPTS), (unsigned long long)na);
    }
}

// P.ts5_claim: electrons a wave claims from the global queue at a time.
// P.ts5_refill: idle lanes (finished or empty) a wave collects before it writes
// their results and refills them: one pass of the divergent start-up (column
// look-up, loads, initial step size) and finish (Kepler, binning) code serves
// many lanes.

// Adaptive Tsit5, one electron per lane, each lane advancing its electron one
// attempted step per loop trip. Two ways to hand out the electrons (items,
// 0-based over all electrons of all segments):
//  - QUEUE: a grid of resident waves; once P.ts5_refill lanes of a wave are idle,
//    their results are written and they take the next items of the wave's
//    pool, and a wave whose pool runs dry claims P.ts5_claim more from the global
//    counter *queue (zero before the launch). Trajectories of very uneven cost
//    (close encounters take tens of times the steps of the rest) then no longer
//    hold 63 idle lanes for the slowest one. A wave leaves when the queue is
//    drained and all its lanes are idle; the queue only grows, so every wave
//    gets there.
//  - static: wave w takes items [64 w, 64 w + 64) and runs until its slowest
//    lane is done.
// MIXED: stage derivatives in FP32; an electron with a stage position within
// restep_radius of the core stops at once, is not binned and gets flags[col]
// = 1 (every other column of the launch gets 0), for the FP64 launch over the
// compacted flagged columns. orig (may be null): the output index of a
// compacted column (records / contributions of that FP64 launch).
// OUT_STATE: an electron starts from its column and the step size and first
// stage derivative k_ts5_prep put in prep, and its final state goes back into
// its column (state = the init buffer) for k_ts5_finish: the start-up and
// finish code that runs with only the refilled lanes active is a few loads and
// stores, so the wave refills often and keeps its lanes busy.
// SL: the laser is simple_laser (one cos^n colour of N >= 2 cycles, azimuth 0; the launcher checks).
template <int DIM, int PHASE, int OUT, int TGT, bool MIXED, bool QUEUE, bool SL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tsit5(
    IntegP P, OutP O, const double* __restrict__ init, const int64_t* __restrict__ seg_start,
    const int64_t* __restrict__ pre, int64_t n_seg, unsigned long long* __restrict__ queue,
    uint8_t* __restrict__ flags, const int64_t* __restrict__ orig, double* __restrict__ state,
    const double* __restrict__ prep) {
    const int lane = lane_id();
    const int64_t total = pre[n_seg];
    const int64_t gw = uniform_i64((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (!QUEUE && gw * 64 >= total) return;
    const bool saeq = (DIM == 2 && PHASE == ETRAJ_PHASE_SCTS && (P.quirks & ETRAJ_QUIRK_SAE2D_SCTS) != 0);
    const double atol = P.atol;
    // the FP32 hydrogen-like EOM reports |r|^2 + soft core
    const float rc2 = (float)(P.restep_radius * P.restep_radius + (TGT != ETRAJ_TARGET_SAE ? P.T.sc : 0.0));
    LaneAcc acc;
    Ts5Lane<DIM, PHASE> st;
    // valid: integrating; pending: finished, result not yet written
    bool valid = false, fresh = false, pending = false, failed = false, close = false;
    int64_t item = gw * 64 + lane, col = 0;
    if (!QUEUE && item < total) {
        valid = true;
        fresh = true;
        const int64_t sg = item_segment(item, pre, n_seg);
        col = seg_start[sg] + (item - pre[sg]);
    }
    // QUEUE, wave-uniform: the claimed, not yet started items [pool, pool_end), the
    // segment holding item `pool`, and whether the counter has passed the last item
    int64_t pool = 0, pool_end = 0, pool_seg = 0;
    bool drained = false;
    const int refill = P.ts5_refill;
    const int64_t claim = P.ts5_claim;
    double t0 = 0.0, prob = 0.0;
    while (true) {
        const uint64_t idle = __ballot(!valid);
        if (idle == ~0ull || (QUEUE && __popcll(idle) >= refill)) {
            if (pending) {
                pending = false;
                if constexpr (OUT == OUT_STATE) {
                    double* c = state + col * P.rows;
#pragma unroll
                    for (int i = 0; i < DIM; ++i) {
                        c[i] = st.r[i];
                        c[DIM + i] = st.p[i];
                    }
                    if (failed) c[0] = __builtin_nan("");
                    if (PHASE != ETRAJ_PHASE_CTMC) c[2 * DIM + 2] = st.ph;
                } else {
                    emit_lane<DIM, PHASE, OUT>(P, O, orig ? orig[col] : col, failed, st.r, st.p, st.ph, t0, prob,
                                               acc);
                }
            }
            if (QUEUE) {
                uint64_t want = idle;
                for (int pass = 0; pass < 2 && want != 0; ++pass) {
                    if (pool >= pool_end) {
                        if (drained) break;
                        // lane 0 claims for the wave; the other lanes add 0, so the claim
                        // is right whether or not the compiler merges the lanes' atomics
                        const unsigned long long old = atomicAdd(queue, lane == 0 ? (unsigned long long)claim : 0ull);
                        const int64_t got = (int64_t)((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)old, 0) |
                                                      ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(old >> 32), 0) << 32));
                        drained = got + claim >= total;
                        if (got >= total) break;
                        pool = got;
                        pool_end = drained ? total : got + claim;
                        pool_seg = item_segment(pool, pre, n_seg);
                    }
                    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0));
                    const int64_t n_want = __popcll(want), left = pool_end - pool;
                    const int64_t take = n_want < left ? n_want : left;
                    if (((want >> lane) & 1) && rank < take) {
                        item = pool + rank;
                        int64_t sg = pool_seg;
                        while (pre[sg + 1] <= item) ++sg;
                        col = seg_start[sg] + (item - pre[sg]);
                        valid = true;
                        fresh = true;
                    }
                    pool += take;
                    while (pool_seg < n_seg - 1 && pre[pool_seg + 1] <= pool) ++pool_seg;
                    want = __ballot(!valid);
                }
            }
            if (__ballot(valid) == 0) break;
        }
        if (fresh) {
            fresh = false;
            const double* c = init + col * P.rows;
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                st.r[i] = c[i];
                st.p[i] = c[DIM + i];
            }
            t0 = c[2 * DIM];
            prob = c[2 * DIM + 1];
            st.ph = (PHASE != ETRAJ_PHASE_CTMC) ? c[2 * DIM + 2] : 0.0;
            close = false;
            if constexpr (OUT == OUT_STATE) {
                const double* q = prep + col * ts5_prep_doubles(DIM);
                st.t = t0;
                st.dt = q[0];
#pragma unroll
                for (int i = 0; i < DIM; ++i) st.k1p[i] = q[1 + i];
                st.k1ph = q[1 + DIM];
                st.lqold = kTs5LogQmin;   // as ts5_init
                if constexpr (SL) {   // the laser's phases at t0
                    st.bsw = q[2 + DIM];
                    st.bcw = q[3 + DIM];
                    st.bsh = q[4 + DIM];
                    st.bch = q[5 + DIM];
                    st.base_ok = 1;
                } else {
                    st.base_ok = 0;
                }
                st.accepted = 0;
                st.attempted = 0;
            } else {
                ts5_init<DIM, PHASE, TGT>(P.L, P.T, saeq, t0, P.tf, P.rtol, atol, st);
            }
        }
        if (valid) {
            const int rc = ts5_attempt<DIM, PHASE, TGT, MIXED, SL>(P.L, P.T, saeq, t0, P.tf, P.rtol, atol,
                                                                   P.maxiters, st, rc2, &close);
            if (MIXED && close) {   // re-integrated in FP64 by the next launch
                valid = false;
                acc.steps += st.accepted;
                acc.attempts += st.attempted;
                acc.restep++;
                flags[col] = 1;
            } else if (rc != 0) {
                valid = false;
                pending = true;
                failed = rc < 0;
#pragma unroll
                for (int i = 0; i < DIM; ++i) failed = failed || isnan(st.r[i]) || isnan(st.p[i]);
                acc.steps += st.accepted;
                acc.attempts += st.attempted;
                if (MIXED) flags[col] = 0;
            }
        }
    }
    if (OUT == OUT_PMD || OUT == OUT_STATE) flush_lane_acc(O, acc, !MIXED);
}

// The start values of the adaptive step (ts5_init) of every electron into prep
// (ts5_prep_doubles per column): the first step size and the first stage
// derivative. Wavefront w takes chunk w % cps (64 columns) of segment w / cps.
template <int DIM, int PHASE, int TGT>
__global__ __launch_bounds__(256) void k_ts5_prep(IntegP P, const double* __restrict__ init,
                                                  const int64_t* __restrict__ seg_start,
                                                  const int32_t* __restrict__ seg_count, int64_t n_seg, int cps,
                                                  double* __restrict__ prep) {
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t seg = gw / cps;
    if (seg >= n_seg) return;
    const int k = (int)(gw - seg * cps) * 64 + lane_id();
    if (k >= seg_count[seg]) return;
    const int64_t col = seg_start[seg] + k;
    const double* c = init + col * P.rows;
    Ts5Lane<DIM, PHASE> st;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        st.r[i] = c[i];
        st.p[i] = c[DIM + i];
    }
    st.ph = (PHASE != ETRAJ_PHASE_CTMC) ? c[2 * DIM + 2] : 0.0;
    const bool saeq = (DIM == 2 && PHASE == ETRAJ_PHASE_SCTS && (P.quirks & ETRAJ_QUIRK_SAE2D_SCTS) != 0);
    double* q = prep + col * ts5_prep_doubles(DIM);
    if (simple_laser(P.L)) {   // the laser's phases at t0, for the start values and the simple-laser kernel
        const MonoBase b = mono_base(P.L.c[0], c[2 * DIM]);
        ts5_init<DIM, PHASE, TGT, true>(P.L, P.T, saeq, c[2 * DIM], P.tf, P.rtol, P.atol, st, &b);
        q[2 + DIM] = b.sw;
        q[3 + DIM] = b.cw;
        q[4 + DIM] = b.sh;
        q[5 + DIM] = b.ch;
    } else {
        ts5_init<DIM, PHASE, TGT>(P.L, P.T, saeq, c[2 * DIM], P.tf, P.rtol, P.atol, st);
    }
    q[0] = st.dt;
#pragma unroll
    for (int i = 0; i < DIM; ++i) q[1 + i] = st.k1p[i];
    q[1 + DIM] = st.k1ph;
}

__global__ void k_zero_queue(unsigned long long* q) { *q = 0ull; }

struct Ts5Launch {
    const double* init;
    const int64_t* seg_start;
    const int64_t* pre;
    int64_t n_seg, max_items;
    unsigned long long* queue;
    uint8_t* flags;
    const int64_t* orig;
    double* state;        // OUT_STATE: the init buffer
    const double* prep;   // OUT_STATE: k_ts5_prep's start values
};

// Blocks of the resident grid of one k_tsit5 instance: the blocks the chip
// holds at once, looked up once per device, and no more than the items need.
template <typename K>
static hipError_t ts5_grid(K kernel, int64_t max_items, unsigned& blocks) {
    constexpr int kMaxDev = 64;
    static std::atomic<int> resident[kMaxDev] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int r = dev < kMaxDev ? resident[dev].load(std::memory_order_relaxed) : 0;
    if (r == 0) {
        int cus = 0, per_cu = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0);
        if (e != hipSuccess) return e;
        r = cus * (per_cu > 0 ? per_cu : 1);
        if (dev < kMaxDev) resident[dev].store(r, std::memory_order_relaxed);
    }
    const int64_t need = (max_items + 255) / 256;
    blocks = (unsigned)(need < r ? need : r);
    return hipSuccess;
}

template <int DIM, int PHASE, int OUT, int TGT, bool MIXED, bool SL>
static hipError_t launch_tsit5_s(hipStream_t st, const IntegP& P, const OutP& O, const Ts5Launch& a) {
    auto kernel = k_tsit5<DIM, PHASE, OUT, TGT, MIXED, true, SL>;
    unsigned blocks = 0;
    hipError_t e = ts5_grid(kernel, a.max_items, blocks);
    if (e != hipSuccess || blocks == 0) return e;
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(1), 0, st, a.queue);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, st, P, O, a.init, a.seg_start, a.pre, a.n_seg, a.queue,
                       a.flags, a.orig, a.state, a.prep);
    return hipGetLastError();
}

// The run_job launches (OUT_STATE) of a simple laser take the kernel built for it.
template <int DIM, int PHASE, int OUT, int TGT, bool MIXED>
static hipError_t launch_tsit5_k(hipStream_t st, const IntegP& P, const OutP& O, const Ts5Launch& a) {
    if constexpr (OUT == OUT_STATE) {
        if (simple_laser(P.L)) return launch_tsit5_s<DIM, PHASE, OUT, TGT, MIXED, true>(st, P, O, a);
    }
    return launch_tsit5_s<DIM, PHASE, OUT, TGT, MIXED, false>(st, P, O, a);
}

template <int DIM, int PHASE, int OUT, bool MIXED>
static hipError_t launch_tsit5_t(hipStream_t st, const IntegP& P, const OutP& O, const Ts5Launch& a) {
    // The FP32 hydrogen-like EOM drops the soft-core term -Z a / r2^(3/2) of the SCTS phase rate;
    // a soft core that FP32 resolves (the molecular a = 1) takes the kTargetHSoft form, which keeps it.
    if constexpr (MIXED) {
        if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE && P.T.sc > 1e-6)
            return launch_tsit5_k<DIM, PHASE, OUT, kTargetHSoft, MIXED>(st, P, O, a);
    }
    if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        return launch_tsit5_k<DIM, PHASE, OUT, ETRAJ_TARGET_HYDROGEN_LIKE, MIXED>(st, P, O, a);
    return launch_tsit5_k<DIM, PHASE, OUT, ETRAJ_TARGET_SAE, MIXED>(st, P, O, a);
}

template <int DIM, int PHASE, bool MIXED>
static hipError_t launch_tsit5_o(hipStream_t st, const IntegP& P, const OutP& O, int out, const Ts5Launch& a) {
    switch (out) {
        case OUT_PMD: return launch_tsit5_t<DIM, PHASE, OUT_PMD, MIXED>(st, P, O, a);
        case OUT_RECORDS: return launch_tsit5_t<DIM, PHASE, OUT_RECORDS, MIXED>(st, P, O, a);
        case OUT_STATE: return launch_tsit5_t<DIM, PHASE, OUT_STATE, MIXED>(st, P, O, a);
        default: return launch_tsit5_t<DIM, PHASE, OUT_CONTRIB, MIXED>(st, P, O, a);
    }
}

template <int DIM, bool MIXED>
static hipError_t launch_tsit5_p(hipStream_t st, const IntegP& P, const OutP& O, int out, const Ts5Launch& a) {
    switch (P.phase) {
        case ETRAJ_PHASE_CTMC: return launch_tsit5_o<DIM, ETRAJ_PHASE_CTMC, MIXED>(st, P, O, out, a);
        case ETRAJ_PHASE_QTMC: return launch_tsit5_o<DIM, ETRAJ_PHASE_QTMC, MIXED>(st, P, O, out, a);
        default: return launch_tsit5_o<DIM, ETRAJ_PHASE_SCTS, MIXED>(st, P, O, out, a);
    }
}

hipError_t launch_tsit5(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                        const int64_t* seg_start, const int64_t* pre, int64_t n_seg, int64_t max_items,
                        unsigned long long* queue, uint8_t* flags, const int64_t* orig) {
    if (n_seg <= 0 || max_items <= 0) return hipSuccess;
    if (!queue || out == OUT_STATE) return hipErrorInvalidValue;
    const Ts5Launch a{init, seg_start, pre, n_seg, max_items, queue, flags, orig, nullptr, nullptr};
    if (P.mixed) {
        if (!flags) return hipErrorInvalidValue;
        return P.dim == 3 ? launch_tsit5_p<3, true>(st, P, O, out, a) : launch_tsit5_p<2, true>(st, P, O, out, a);
    }
    return P.dim == 3 ? launch_tsit5_p<3, false>(st, P, O, out, a) : launch_tsit5_p<2, false>(st, P, O, out, a);
}

template <int DIM, int PHASE>
static hipError_t launch_ts5_prep_p(hipStream_t st, const IntegP& P, const double* init, const int64_t* seg_start,
                                    const int32_t* seg_count, int64_t n_seg, int cps, double* prep) {
    const dim3 grid((unsigned)((n_seg * (int64_t)cps + 3) / 4));
    if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        hipLaunchKernelGGL((k_ts5_prep<DIM, PHASE, ETRAJ_TARGET_HYDROGEN_LIKE>), grid, dim3(256), 0, st, P, init,
                           seg_start, seg_count, n_seg, cps, prep);
    else
        hipLaunchKernelGGL((k_ts5_prep<DIM, PHASE, ETRAJ_TARGET_SAE>), grid, dim3(256), 0, st, P, init, seg_start,
                           seg_count, n_seg, cps, prep);
    return hipGetLastError();
}

template <int DIM>
static hipError_t launch_ts5_prep(hipStream_t st, const IntegP& P, const double* init, const int64_t* seg_start,
                                  const int32_t* seg_count, int64_t n_seg, int cps, double* prep) {
    switch (P.phase) {
        case ETRAJ_PHASE_CTMC:
            return launch_ts5_prep_p<DIM, ETRAJ_PHASE_CTMC>(st, P, init, seg_start, seg_count, n_seg, cps, prep);
        case ETRAJ_PHASE_QTMC:
            return launch_ts5_prep_p<DIM, ETRAJ_PHASE_QTMC>(st, P, init, seg_start, seg_count, n_seg, cps, prep);
        default: return launch_ts5_prep_p<DIM, ETRAJ_PHASE_SCTS>(st, P, init, seg_start, seg_count, n_seg, cps, prep);
    }
}

hipError_t launch_tsit5_staged(hipStream_t st, const IntegP& P, const OutP& O, double* init, const int64_t* seg_start,
                               const int32_t* seg_count, const int64_t* pre, int64_t n_seg, int cps,
                               int64_t max_items, unsigned long long* queue, double* prep, uint8_t* flags,
                               hipEvent_t main_start, hipEvent_t main_done) {
    if (n_seg <= 0 || max_items <= 0 || cps <= 0) return hipSuccess;
    if (!queue || !prep || (P.mixed && !flags)) return hipErrorInvalidValue;
    hipError_t e = P.dim == 3 ? launch_ts5_prep<3>(st, P, init, seg_start, seg_count, n_seg, cps, prep)
                              : launch_ts5_prep<2>(st, P, init, seg_start, seg_count, n_seg, cps, prep);
    if (e == hipSuccess && main_start) e = hipEventRecord(main_start, st);
    if (e != hipSuccess) return e;
    const Ts5Launch a{init, seg_start, pre, n_seg, max_items, queue, flags, nullptr, init, prep};
    if (P.mixed)
        e = P.dim == 3 ? launch_tsit5_p<3, true>(st, P, O, OUT_STATE, a) : launch_tsit5_p<2, true>(st, P, O, OUT_STATE, a);
    else
        e = P.dim == 3 ? launch_tsit5_p<3, false>(st, P, O, OUT_STATE, a)
                       : launch_tsit5_p<2, false>(st, P, O, OUT_STATE, a);
    if (e == hipSuccess && main_done) e = hipEventRecord(main_done, st);
    if (e != hipSuccess) return e;
    return launch_ts5_finish(st, P, O, init, seg_start, seg_count, n_seg, cps, P.mixed ? flags : nullptr);
}

}  // namespace etraj
