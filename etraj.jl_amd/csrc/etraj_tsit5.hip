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

#include <hip/hip_runtime.h>

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
        if (inb) {
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
            fix_atomic_add(O.scalq + 0, ip);
            fix_atomic_add(O.scalq + 2, iu);
        }
    } else {
        const double ip = wave_sum_f64(a.ip), iu = wave_sum_f64(a.iu);
        if (lane_id() == 0) {
            if (ip != 0.0) unsafeAtomicAdd(O.scal + 0, ip);
            if (iu != 0.0) unsafeAtomicAdd(O.scal + 1, iu);
        }
    }
    if (lane_id() == 0) {
        if (nf) atomicAdd((unsigned long long*)(O.cnt + CNT_FAILED), (unsigned long long)nf);
        if (nb) atomicAdd((unsigned long long*)(O.cnt + CNT_BOUND), (unsigned long long)nb);
        if (ns) atomicAdd((unsigned long long*)(O.cnt + CNT_STEPS), (unsigned long long)ns);
        if (ns && steps64) atomicAdd((unsigned long long*)(O.cnt + CNT_STEPS64), (unsigned long long)ns);
        if (nr) atomicAdd((unsigned long long*)(O.cnt + CNT_FIXRANGE), (unsigned long long)nr);
        if (nq) atomicAdd((unsigned long long*)(O.cnt + CNT_RESTEP), (unsigned long long)nq);
        if (na) atomicAdd((unsigned long long*)(O.cnt + CNT_ATTEMPTS), (unsigned long long)na);
    }
}

// item (0-based over all electrons of all segments) -> column index
__device__ inline int64_t item_column(int64_t item, const int64_t* pre, int64_t n_seg, const int64_t* seg_start) {
    int64_t lo = 0, hi = n_seg;  // pre[lo] <= item < pre[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (pre[mid] <= item) lo = mid; else hi = mid;
    }
    return seg_start[lo] + (item - pre[lo]);
}

// Adaptive Tsit5, one electron per lane: wavefront w integrates items
// [64 w, 64 w + 64) (0-based over all electrons of all segments), each lane
// advancing its electron one attempted step per loop trip until every lane of
// the wave is done. (A form in which a finished lane took the next electron
// of a longer per-wave range measured slower at every range length: 40.4 ms
// at 128 electrons per wave, 44.1 at 512, 53.9 at 2048, against 39.7 for one
// electron per lane, on the Ar SPA 3D grid of 9.6M electrons.)
// MIXED: stage derivatives in FP32; an electron with a stage position within
// restep_radius of the core stops at once, is not binned and gets flags[col]
// = 1 (every other column of the launch gets 0), for the FP64 launch over the
// compacted flagged columns. orig (may be null): the output index of a
// compacted column (records / contributions of that FP64 launch).
template <int DIM, int PHASE, int OUT, int TGT, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tsit5(
    IntegP P, OutP O, const double* __restrict__ init, const int64_t* __restrict__ seg_start,
    const int64_t* __restrict__ pre, int64_t n_seg, uint8_t* __restrict__ flags, const int64_t* __restrict__ orig) {
    const int lane = lane_id();
    const int64_t total = pre[n_seg];
    const int64_t gw = uniform_i64((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int64_t item = gw * 64 + lane;
    if (gw * 64 >= total) return;
    const bool saeq = (DIM == 2 && PHASE == ETRAJ_PHASE_SCTS && (P.quirks & ETRAJ_QUIRK_SAE2D_SCTS) != 0);
    const double atol = P.atol;
    // the FP32 hydrogen-like EOM reports |r|^2 + soft core
    const float rc2 = (float)(P.restep_radius * P.restep_radius + (TGT != ETRAJ_TARGET_SAE ? P.T.sc : 0.0));
    LaneAcc acc;
    Ts5Lane<DIM, PHASE> st;
    bool valid = item < total, close = false;
    int64_t col = 0;
    double t0 = 0.0, prob = 0.0;
    if (valid) {
        col = item_column(item, pre, n_seg, seg_start);
        const double* c = init + col * P.rows;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            st.r[i] = c[i];
            st.p[i] = c[DIM + i];
        }
        t0 = c[2 * DIM];
        prob = c[2 * DIM + 1];
        st.ph = (PHASE != ETRAJ_PHASE_CTMC) ? c[2 * DIM + 2] : 0.0;
        ts5_init<DIM, PHASE, TGT>(P.L, P.T, saeq, t0, P.tf, P.rtol, atol, st);
    }
    while (__ballot(valid) != 0) {
        if (valid) {
            const int rc = ts5_attempt<DIM, PHASE, TGT, MIXED>(P.L, P.T, saeq, t0, P.tf, P.rtol, atol, P.maxiters,
                                                               st, rc2, &close);
            if (MIXED && close) {   // re-integrated in FP64 by the next launch
                valid = false;
                acc.steps += st.accepted;
                acc.attempts += st.attempted;
                acc.restep++;
                flags[col] = 1;
            } else if (rc != 0) {
                valid = false;
                bool failed = rc < 0;
#pragma unroll
                for (int i = 0; i < DIM; ++i) failed = failed || isnan(st.r[i]) || isnan(st.p[i]);
                acc.steps += st.accepted;
                acc.attempts += st.attempted;
                if (MIXED) flags[col] = 0;
                emit_lane<DIM, PHASE, OUT>(P, O, orig ? orig[col] : col, failed, st.r, st.p, st.ph, t0, prob, acc);
            }
        }
    }
    if (OUT == OUT_PMD) flush_lane_acc(O, acc, !MIXED);
}

struct Ts5Launch {
    const double* init;
    const int64_t* seg_start;
    const int64_t* pre;
    int64_t n_seg, max_items;
    uint8_t* flags;
    const int64_t* orig;
};

template <int DIM, int PHASE, int OUT, bool MIXED>
static hipError_t launch_tsit5_t(hipStream_t st, const IntegP& P, const OutP& O, const Ts5Launch& a) {
    const int64_t waves = (a.max_items + 63) / 64;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    if (blocks == 0) return hipSuccess;
    // The FP32 hydrogen-like EOM drops the soft-core term -Z a / r2^(3/2) of the SCTS phase rate;
    // a soft core that FP32 resolves (the molecular a = 1) takes the kTargetHSoft form, which keeps it.
    if constexpr (MIXED) {
        if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE && P.T.sc > 1e-6) {
            hipLaunchKernelGGL((k_tsit5<DIM, PHASE, OUT, kTargetHSoft, MIXED>), dim3(blocks), dim3(256), 0, st, P, O,
                               a.init, a.seg_start, a.pre, a.n_seg, a.flags, a.orig);
            return hipGetLastError();
        }
    }
    if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        hipLaunchKernelGGL((k_tsit5<DIM, PHASE, OUT, ETRAJ_TARGET_HYDROGEN_LIKE, MIXED>), dim3(blocks), dim3(256), 0, st,
                           P, O, a.init, a.seg_start, a.pre, a.n_seg, a.flags, a.orig);
    else
        hipLaunchKernelGGL((k_tsit5<DIM, PHASE, OUT, ETRAJ_TARGET_SAE, MIXED>), dim3(blocks), dim3(256), 0, st, P, O,
                           a.init, a.seg_start, a.pre, a.n_seg, a.flags, a.orig);
    return hipGetLastError();
}

template <int DIM, int PHASE, bool MIXED>
static hipError_t launch_tsit5_o(hipStream_t st, const IntegP& P, const OutP& O, int out, const Ts5Launch& a) {
    switch (out) {
        case OUT_PMD: return launch_tsit5_t<DIM, PHASE, OUT_PMD, MIXED>(st, P, O, a);
        case OUT_RECORDS: return launch_tsit5_t<DIM, PHASE, OUT_RECORDS, MIXED>(st, P, O, a);
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
                        uint8_t* flags, const int64_t* orig) {
    if (n_seg <= 0 || max_items <= 0) return hipSuccess;
    const Ts5Launch a{init, seg_start, pre, n_seg, max_items, flags, orig};
    if (P.mixed) {
        if (!flags) return hipErrorInvalidValue;
        return P.dim == 3 ? launch_tsit5_p<3, true>(st, P, O, out, a) : launch_tsit5_p<2, true>(st, P, O, out, a);
    }
    return P.dim == 3 ? launch_tsit5_p<3, false>(st, P, O, out, a) : launch_tsit5_p<2, false>(st, P, O, out, a);
}

}  // namespace etraj
