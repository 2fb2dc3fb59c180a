// etraj_kernels.hip -- gfx950 kernels of the trajectory engine.
//
// Pipeline (one call of etraj_run_job):
//   k_adk_sample      (etraj_sample.hip) one workgroup per t0 batch: tunnel
//                     exit, weight, cutoff and order-preserving compaction into
//                     the init buffer (Julia column layout, one column per electron).
//   k_rk4             one electron per lane, 64 electrons of ONE batch per
//                     wavefront: they share t0 and the step size, so the laser
//                     field of 32 steps is evaluated once per wave (one time
//                     point per lane) and broadcast with v_readlane; fixed-step
//                     RK4 in FP64 or FP32+compensation with FP64 re-steps near
//                     the core; FP64 Kepler asymptote; FP64 atomics into the PMD.
//   k_finalize        |amplitude|^2 for coherent sums.
// The adaptive Tsit5 kernel (k_tsit5) is in etraj_tsit5.hip.
// The work is VALU-bound ODE arithmetic (no dense contraction): no MFMA.
#include "etraj_engine.hpp"
#include "etraj_device.hpp"

#include <hip/hip_runtime.h>

namespace etraj {

// ---------------------------------------------------------------------------
// field / target evaluation (LaserFx/Fy/Ax/Ay, TargetPotential/TargetForce)
// ---------------------------------------------------------------------------
__global__ void k_laser_field(LaserP L, const double* __restrict__ t, int64_t n, double* Fx, double* Fy, double* Ax,
                              double* Ay) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double fx, fy, ax, ay;
    laser_field(L, t[i], fx, fy);
    laser_potential(L, t[i], ax, ay);
    if (Fx) Fx[i] = fx;
    if (Fy) Fy[i] = fy;
    if (Ax) Ax[i] = ax;
    if (Ay) Ay[i] = ay;
}

__global__ void k_target_eval(TargetP T, const double* __restrict__ xyz, int64_t n, double* V, double* F) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double r2 = x * x + y * y + z * z;
    double v, s;
    target_vs<double>(T, r2, true, false, v, s);
    if (V) V[i] = v;
    if (F) {
        F[3 * i] = s * x;
        F[3 * i + 1] = s * y;
        F[3 * i + 2] = s * z;
    }
}

// ---------------------------------------------------------------------------
// Fixed-step RK4 integrator + Kepler asymptote + binning.
// ---------------------------------------------------------------------------
template <int DIM, int PHASE>
struct Lane {
    double t0, prob, phase0;
    double r[DIM], p[DIM];
};

template <int DIM, int PHASE>
__device__ inline void load_column(const double* col, Lane<DIM, PHASE>& e) {
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        e.r[i] = col[i];
        e.p[i] = col[DIM + i];
    }
    e.t0 = col[2 * DIM];
    e.prob = col[2 * DIM + 1];
    e.phase0 = (PHASE != ETRAJ_PHASE_CTMC) ? col[2 * DIM + 2] : 0.0;
}

// Outputs of one electron after traj_t_final.
template <int DIM, int PHASE, int OUT>
__device__ inline void emit(const IntegP& P, const OutP& O, int64_t col_idx, bool active, bool failed,
                            const double* r, const double* p, double phase, double t0, double prob) {
    Final f = kepler_final<DIM, PHASE>(P.T, P.quirks, r, p, phase, t0);
    if (failed) f.ionized = false;
    bool bound = active && !failed && !f.ionized;
    int64_t lin = -1;
    bool inb = false;
    double cre = 0.0, cim = 0.0;
    if (active && f.ionized) {
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
                double s = sqrt(prob), sp, cp;
                sincos(f.phase, &sp, &cp);
                cre = s * cp;
                cim = s * sp;
            }
        }
    }
    if (OUT == OUT_PMD) {
        int64_t nr = 0;
        const bool rec = O.bin_cell != nullptr;   // records for the tile binning (wave-uniform)
        if (rec && active) {
            O.bin_cell[col_idx] = inb ? (uint32_t)(O.cell0 + lin) : kNoCell;
            O.bin_amp[col_idx] = make_double2(cre, cim);
        }
        if (O.pmdq) {   // fixed point: exact integer sums (wave-uniform branch)
            bool ok1 = true, ok2 = true, ok3, ok4;
            if (inb && !rec) {
                if (PHASE == ETRAJ_PHASE_CTMC) {
                    fix_atomic_add(O.pmdq + 2 * lin, fix_from_double(cre, ok1));
                } else {
                    fix_atomic_add(O.pmdq + 4 * lin, fix_from_double(cre, ok1));
                    fix_atomic_add(O.pmdq + 4 * lin + 2, fix_from_double(cim, ok2));
                }
            }
            const Fix128 ip = fix_wave_sum(fix_from_double((active && f.ionized) ? prob : 0.0, ok3));
            const Fix128 iu = fix_wave_sum(fix_from_double((active && f.ionized && !inb) ? prob : 0.0, ok4));
            nr = wave_sum_i64((ok1 && ok2 && ok3 && ok4) ? 0 : 1);
            if (lane_id() == 0) {
                fix_atomic_add(O.scalq + 4 * acc_slot() + 0, ip);
                fix_atomic_add(O.scalq + 4 * acc_slot() + 2, iu);
            }
        } else {
            if (inb && !rec) {
                if (PHASE == ETRAJ_PHASE_CTMC) {
                    unsafeAtomicAdd(O.pmd + lin, cre);
                } else {
                    unsafeAtomicAdd(O.pmd + 2 * lin, cre);
                    unsafeAtomicAdd(O.pmd + 2 * lin + 1, cim);
                }
            }
            double ip = wave_sum_f64((active && f.ionized) ? prob : 0.0);
            double iu = wave_sum_f64((active && f.ionized && !inb) ? prob : 0.0);
            if (lane_id() == 0) {
                if (ip != 0.0) unsafeAtomicAdd(O.scal + 2 * acc_slot() + 0, ip);
                if (iu != 0.0) unsafeAtomicAdd(O.scal + 2 * acc_slot() + 1, iu);
            }
        }
        if (O.hits && inb && !rec) atomicAdd((unsigned long long*)(O.hits + lin), 1ull);
        int64_t nf = wave_sum_i64((active && failed) ? 1 : 0);
        int64_t nb = wave_sum_i64(bound ? 1 : 0);
        if (lane_id() == 0) {
            int64_t* cs = O.cnt + CNT_N * acc_slot();
            if (nf) atomicAdd((unsigned long long*)(cs + CNT_FAILED), (unsigned long long)nf);
            if (nb) atomicAdd((unsigned long long*)(cs + CNT_BOUND), (unsigned long long)nb);
            if (nr) atomicAdd((unsigned long long*)(cs + CNT_FIXRANGE), (unsigned long long)nr);
        }
    } else if (OUT == OUT_RECORDS) {
        if (active) {
            double* rec = O.records + 12 * col_idx;
            rec[0] = r[0]; rec[1] = r[1]; rec[2] = DIM == 3 ? r[2] : 0.0;
            rec[3] = p[0]; rec[4] = p[1]; rec[5] = DIM == 3 ? p[2] : 0.0;
            rec[6] = phase; rec[7] = f.E;
            rec[8] = f.pinf[0]; rec[9] = f.pinf[1]; rec[10] = f.pinf[2];
            rec[11] = f.phase;
            O.status[col_idx] = failed ? 1 : (f.ionized ? 0 : 2);
        }
    } else {  // OUT_CONTRIB: the caller bins on the host, in column order
        if (active) {
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
}

// Field table of one segment (equal t0): the segment's step is
// h = (t_f - t0)/n, n = ceil((t_f - t0)/dt), and step s needs the field at
// half-steps 2s+1 and 2s+2 (t0 + j h/2). Entry s holds both, as FP32
// (Fx, Fy, Fx, Fy) and FP64 (two double2); f0 holds F(t0).
__global__ __launch_bounds__(256) void k_field_table(const LaserP* __restrict__ lps, const int32_t* __restrict__ seg_job,
                                                     const double* __restrict__ t0s, int64_t nseg, double tf, double dt,
                                                     int64_t stride, float4* __restrict__ f32,
                                                     double2* __restrict__ f64, double2* __restrict__ f0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg * stride) return;
    const int64_t seg = i / stride, s = i - seg * stride;
    const LaserP& L = lps[seg_job ? seg_job[seg] : 0];
    const double t0 = t0s[seg];
    const double span = tf - t0;
    int64_t n = (int64_t)ceil(span / dt);
    if (n < 1) n = 1;
    if (s >= n) return;
    const double hh = (span / (double)n) * 0.5;
    double mx, my, ex, ey;
    laser_field(L, t0 + (double)(2 * s + 1) * hh, mx, my);
    laser_field(L, t0 + (double)(2 * s + 2) * hh, ex, ey);
    f32[i] = make_float4((float)mx, (float)my, (float)ex, (float)ey);
    f64[2 * i] = make_double2(mx, my);
    f64[2 * i + 1] = make_double2(ex, ey);
    if (s == 0) {
        double ax, ay;
        laser_field(L, t0, ax, ay);
        f0[seg] = make_double2(ax, ay);
    }
}

// The FP32 pass of one electron over n steps from t0 (see k_rk4). TGT2 is
// the target kind, kTargetHSoft for a hydrogen-like target whose soft core is
// not negligible in FP32 (decided once per wave, outside the step loop);
// UNITZ: the charge Z is 1. T32 is the segment's row of the FP32 field table.
//
// The state is an FP64 base plus an FP32 increment: r = rb + dr, rb the base
// rounded to FP32 and dr the sum of the step increments h/6 Sr since the base
// was last updated; every 32 steps dr is added to the FP64 base. The EOM sees
// the FP32 sum rb + dr, and the increments are summed at the resolution of
// |dr| (at most 32 steps' worth) rather than of |r| -- two FP32 operations
// per component and step, where a Kahan-compensated FP32 state takes four.
// The phase increments are summed the same way over the 32 steps and the
// partial sums Kahan-summed in FP64 (the SCTS phase reaches 1e3 rad).
template <int DIM, int PHASE, int TGT2, bool UNITZ>
__device__ inline void rk4_fp32_pass(const IntegP& P, bool saeq, const Lane<DIM, PHASE>& e, int64_t n, double h,
                                     __amdgpu_buffer_rsrc_t T32, double2 F0, bool& close, double* rf, double* pf,
                                     double& phase) {
    constexpr bool HL = is_hydrogen_like(TGT2);
    constexpr bool SOFT = TGT2 == kTargetHSoft;
    double r64[DIM], p64[DIM];
    float rb[DIM], pb[DIM], dr[DIM], dp[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        r64[i] = e.r[i];
        p64[i] = e.p[i];
        rb[i] = (float)r64[i];
        pb[i] = (float)p64[i];
        dr[i] = 0.0f;
        dp[i] = 0.0f;
    }
    phase = e.phase0;
    double phc = 0.0;
    const float hf = (float)h, h6 = (float)(h / 6);
    const HlPhase<float> pc = hl_phase<float, PHASE, SOFT>(h, P.T.Ip);
    const double c0 = (HL && PHASE == ETRAJ_PHASE_QTMC) ? -h * P.T.Ip : 0.0;   // QTMC's -Ip per step (HL form)
    const float Zf = (float)P.T.Zd, af = (float)P.T.sc;
    // the hydrogen-like EOM reports |r|^2 + soft core
    const float rc2 = (float)(P.restep_radius * P.restep_radius + (HL ? P.T.sc : 0.0));
    // the smallest r2 over the step starts, kept as the bit pattern of a positive float (the
    // integer order of such patterns is their numeric order: one v_min_u32 per step)
    uint32_t r2min = 0x7f7fffffu;
    float Fpx = (float)F0.x, Fpy = (float)F0.y;
    for (int64_t s0 = 0; s0 < n; s0 += 32) {
        const int cnt = (int)(s0 + 32 < n ? 32 : n - s0);   // wave-uniform trip count
        float qs = 0.0f;
        // one RK4 step with the table entry f = (F at t + h/2, F at t + h)
        auto step = [&](const float4 f) {
            float r[DIM], p[DIM], Sr[DIM], Sp[DIM], r2;
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                r[i] = rb[i] + dr[i];
                p[i] = pb[i] + dp[i];
            }
            if constexpr (HL) {
                float SA, SQ;
                hl_rk4_sums<float, DIM, PHASE, SOFT, UNITZ>(Zf, af, r, p, hf, 0.5f * hf, Fpx, Fpy, f.x, f.y, f.z, f.w, Sr,
                                                            Sp, SA, SQ, r2);
                if (PHASE != ETRAJ_PHASE_CTMC) qs = fma(pc.cA, SA, fma(pc.cQ, SQ, qs));
            } else {
                float SPh;
                gen_rk4_sums<float, DIM, PHASE, TGT2>(P.T, saeq, r, p, hf, Fpx, Fpy, f.x, f.y, f.z, f.w, Sr, Sp, SPh,
                                                      r2);
                if (PHASE != ETRAJ_PHASE_CTMC) qs = fma(h6, SPh, qs);
            }
            r2min = min(r2min, __float_as_uint(r2));
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                dr[i] = fma(h6, Sr[i], dr[i]);
                dp[i] = fma(h6, Sp[i], dp[i]);
            }
            Fpx = f.z;
            Fpy = f.w;
        };
        // steps in pairs: the entry of the next step is loaded while this one computes
        // (a read past the row's end returns zeros and is not used)
        float4 fa = row_load_f4(T32, (int)(s0 * 16));
        for (int k = 0; k < cnt; k += 2) {
            const float4 fb = row_load_f4(T32, (int)((s0 + k + 1) * 16));
            step(fa);
            if (k + 1 < cnt) {
                fa = row_load_f4(T32, (int)((s0 + k + 2) * 16));
                step(fb);
            }
        }
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            r64[i] += (double)dr[i];
            p64[i] += (double)dp[i];
            rb[i] = (float)r64[i];
            pb[i] = (float)p64[i];
            dr[i] = 0.0f;
            dp[i] = 0.0f;
        }
        if (PHASE != ETRAJ_PHASE_CTMC) {
            const double y = ((double)qs + c0 * (double)cnt) - phc;
            const double t = phase + y;
            phc = (t - phase) - y;
            phase = t;
        }
    }
    close = __uint_as_float(r2min) < rc2;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        rf[i] = r64[i];
        pf[i] = p64[i];
    }
}

// The FP64 pass of one electron over n steps from t0: the same RK4 step in
// FP64, with the exact hydrogen-like phase rate (the soft-core term kept).
// T64 is the segment's row of the FP64 field table (32 bytes per step).
template <int DIM, int PHASE, int TGT, bool UNITZ>
__device__ inline void rk4_fp64_pass(const IntegP& P, bool saeq, const Lane<DIM, PHASE>& e, int64_t n, double h,
                                     __amdgpu_buffer_rsrc_t T64, double2 F0, double* r, double* p, double& phase) {
    constexpr bool HL = is_hydrogen_like(TGT);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        r[i] = e.r[i];
        p[i] = e.p[i];
    }
    phase = e.phase0;
    h = uniform_f64(h);
    const double hh = uniform_f64(0.5 * h), h6 = uniform_f64(h / 6);
    const HlPhase<double> pc0 = hl_phase<double, PHASE, true>(h, P.T.Ip);
    const HlPhase<double> pc{uniform_f64(pc0.cA), uniform_f64(pc0.cQ), uniform_f64(pc0.c0)};
    double Fpx = F0.x, Fpy = F0.y;
    auto step = [&](const double2 Fm, const double2 F1) {
        double Sr[DIM], Sp[DIM], r2;
        if constexpr (HL) {
            double SA, SQ;
            hl_rk4_sums<double, DIM, PHASE, true, UNITZ>(P.T.Zd, P.T.sc, r, p, h, hh, Fpx, Fpy, Fm.x, Fm.y, F1.x,
                                                         F1.y, Sr, Sp, SA, SQ, r2);
            if (PHASE != ETRAJ_PHASE_CTMC) phase = fma(pc.cA, SA, fma(pc.cQ, SQ, phase));
            if (PHASE == ETRAJ_PHASE_QTMC) phase += pc.c0;
        } else {
            double SPh;
            gen_rk4_sums<double, DIM, PHASE, TGT>(P.T, saeq, r, p, h, Fpx, Fpy, Fm.x, Fm.y, F1.x, F1.y, Sr, Sp, SPh,
                                                  r2);
            if (PHASE != ETRAJ_PHASE_CTMC) phase = fma(h6, SPh, phase);
        }
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            r[i] = fma(h6, Sr[i], r[i]);
            p[i] = fma(h6, Sp[i], p[i]);
        }
        Fpx = F1.x;
        Fpy = F1.y;
    };
    // steps in pairs, the next step's entries loaded while this one computes
    double2 ma = row_load_d2(T64, 0), ea = row_load_d2(T64, 16);
    for (int64_t s = 0; s < n; s += 2) {
        const double2 mb = row_load_d2(T64, (int)((s + 1) * 32)), eb = row_load_d2(T64, (int)((s + 1) * 32 + 16));
        step(ma, ea);
        if (s + 1 < n) {
            ma = row_load_d2(T64, (int)((s + 2) * 32));
            ea = row_load_d2(T64, (int)((s + 2) * 32 + 16));
            step(mb, eb);
        }
    }
}

// 64 electrons of one segment (equal t0) per wavefront. Segment s covers
// columns [seg_start[s], seg_start[s] + seg_count[s]); its field table row is
// FT.f32/f64 + s * FT.stride, read with wave-uniform (scalar) loads.
//
// MIXED (rk4_fp32_pass): FP32 state with a Kahan compensation term per
// component; phase increments are Kahan-summed in FP32 over 32 steps and the
// partial sums Kahan-summed in FP64 (the SCTS phase reaches 1e3 rad). An electron that starts
// any step within restep_radius of the core is not binned: flags[col] = 1
// marks it for the FP64 pass (k_compact_flagged + this kernel with
// MIXED = false), which integrates it again from t0. Errors a close encounter
// amplifies are accumulated before it, so the whole trajectory is redone.
// out_index (may be NULL) maps a column to the index of its record/contrib.
template <int DIM, int PHASE, bool MIXED, int OUT, int TGT>
__device__ inline void rk4_chunk(const IntegP& P, const OutP& O, const double* __restrict__ init,
                                 const int64_t* __restrict__ seg_start, const int32_t* __restrict__ seg_count,
                                 int64_t seg, int chunk, const FieldTab& FT, uint8_t* __restrict__ flags,
                                 const int64_t* __restrict__ out_index) {
    const int lane = lane_id();
    const int32_t cnt = seg_count[seg];
    const int first = chunk * 64;
    if (first >= cnt) return;
    const int64_t cbase = seg_start[seg];
    const bool active = first + lane < cnt;
    const int64_t col_idx = cbase + first + (active ? lane : 0);
    OutP Oj = O;
    if (O.seg_job) {
        const int job = O.seg_job[seg];
        Oj.pmd = O.pmd + job * O.job_stride;
        Oj.scal = O.scal + 2 * kAccSlots * job;
        Oj.cnt = O.cnt + (int64_t)job * CNT_N * kAccSlots;
        if (O.pmdq) {
            Oj.pmdq = O.pmdq + 2 * job * O.job_stride;
            Oj.scalq = O.scalq + 4 * kAccSlots * job;
        }
        if (O.hits) Oj.hits = O.hits + job * O.npix;
        Oj.cell0 = job * O.npix;
    }

    Lane<DIM, PHASE> e;
    load_column<DIM, PHASE>(init + col_idx * P.rows, e);
    const double t0 = uniform_f64(init[(cbase + first) * P.rows + 2 * DIM]);
    const double span = P.tf - t0;
    int64_t n = (int64_t)ceil(span / P.dt);
    if (n < 1) n = 1;
    n = uniform_i64(n);   // wave-uniform: loop counters stay scalar
    const double h = span / (double)n;
    const bool saeq = (DIM == 2 && PHASE == ETRAJ_PHASE_SCTS && (P.quirks & ETRAJ_QUIRK_SAE2D_SCTS) != 0);
    const __amdgpu_buffer_rsrc_t T32 = row_rsrc(FT.f32 + seg * FT.stride, FT.stride * (int64_t)sizeof(float4));
    const __amdgpu_buffer_rsrc_t T64 = row_rsrc(FT.f64 + 2 * seg * FT.stride, FT.stride * (int64_t)(2 * sizeof(double2)));
    const double2 F0 = FT.f0[seg];
    const bool unitz = P.T.Zd == 1.0;
    bool bad = false, close = false;

    double rf[DIM], pf[DIM], phase;
    if constexpr (!MIXED) {
        if constexpr (is_hydrogen_like(TGT)) {
            if (unitz) rk4_fp64_pass<DIM, PHASE, TGT, true>(P, saeq, e, n, h, T64, F0, rf, pf, phase);
            else rk4_fp64_pass<DIM, PHASE, TGT, false>(P, saeq, e, n, h, T64, F0, rf, pf, phase);
        } else {
            rk4_fp64_pass<DIM, PHASE, TGT, false>(P, saeq, e, n, h, T64, F0, rf, pf, phase);
        }
    } else {
        if constexpr (is_hydrogen_like(TGT)) {
            if (P.T.sc > 1e-6) {
                if (unitz) rk4_fp32_pass<DIM, PHASE, kTargetHSoft, true>(P, saeq, e, n, h, T32, F0, close, rf, pf, phase);
                else rk4_fp32_pass<DIM, PHASE, kTargetHSoft, false>(P, saeq, e, n, h, T32, F0, close, rf, pf, phase);
            } else {
                if (unitz) rk4_fp32_pass<DIM, PHASE, TGT, true>(P, saeq, e, n, h, T32, F0, close, rf, pf, phase);
                else rk4_fp32_pass<DIM, PHASE, TGT, false>(P, saeq, e, n, h, T32, F0, close, rf, pf, phase);
            }
        } else {
            rk4_fp32_pass<DIM, PHASE, TGT, false>(P, saeq, e, n, h, T32, F0, close, rf, pf, phase);
        }
        if (active && flags) flags[col_idx] = close ? 1 : 0;
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i) bad = bad || isnan(rf[i]) || isnan(pf[i]);
    if (PHASE != ETRAJ_PHASE_CTMC) bad = bad || isnan(phase);
    const bool done = active && !close;
    const int64_t out_idx = (out_index && done) ? out_index[col_idx] : col_idx;
    emit<DIM, PHASE, OUT>(P, Oj, out_idx, done, bad, rf, pf, phase, t0, e.prob);
    if (OUT == OUT_PMD) {
        // steps: every active lane of this pass; FP64 steps also counted apart
        const int64_t nact = wave_sum_i64(active ? 1 : 0);
        const int64_t nclose = wave_sum_i64((active && close) ? 1 : 0);
        if (lane == 0) {
            int64_t* cs = Oj.cnt + CNT_N * acc_slot();
            atomicAdd((unsigned long long*)(cs + CNT_STEPS), (unsigned long long)(n * nact));
            atomicAdd((unsigned long long*)(cs + CNT_ATTEMPTS), (unsigned long long)(n * nact));
            if (!MIXED) atomicAdd((unsigned long long*)(cs + CNT_STEPS64), (unsigned long long)(n * nact));
            if (nclose) atomicAdd((unsigned long long*)(cs + CNT_RESTEP), (unsigned long long)nclose);
        }
    }
}

// Static launch: wave w takes chunk w % chunks_per_seg of segment
// w / chunks_per_seg (waves of empty chunks leave at once).
template <int DIM, int PHASE, bool MIXED, int OUT, int TGT>
__global__ __launch_bounds__(256) void k_rk4(IntegP P, OutP O, const double* __restrict__ init,
                                             const int64_t* __restrict__ seg_start,
                                             const int32_t* __restrict__ seg_count, int64_t n_seg,
                                             int chunks_per_seg, FieldTab FT, uint8_t* __restrict__ flags,
                                             const int64_t* __restrict__ out_index) {
    const int64_t gw = uniform_i64((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int64_t seg = gw / chunks_per_seg;
    if (seg >= n_seg) return;
    rk4_chunk<DIM, PHASE, MIXED, OUT, TGT>(P, O, init, seg_start, seg_count, seg, (int)(gw - seg * chunks_per_seg), FT,
                                            flags, out_index);
}

// The Kepler asymptote and binning of the final states that k_tsit5 (OUT_STATE)
// left in the init columns (x = NaN: the integration failed). Wavefront w takes
// chunk w % cps (64 columns) of segment w / cps, as k_rk4; every lane of a
// wavefront takes part in emit's wave sums. flags (may be null): columns with
// flags[col] != 0 were stopped for the FP64 re-integration and are not binned.
template <int DIM, int PHASE>
__global__ __launch_bounds__(256) void k_ts5_finish(IntegP P, OutP O, const double* __restrict__ init,
                                                    const int64_t* __restrict__ seg_start,
                                                    const int32_t* __restrict__ seg_count, int64_t n_seg, int cps,
                                                    const uint8_t* __restrict__ flags) {
    const int64_t gw = uniform_i64((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int64_t seg = gw / cps;
    if (seg >= n_seg) return;
    const int first = (int)(gw - seg * cps) * 64;
    const int32_t cnt = seg_count[seg];
    if (first >= cnt) return;   // an empty chunk (wave-uniform)
    const int lane = lane_id();
    bool active = first + lane < cnt;
    const int64_t col = seg_start[seg] + first + (active ? lane : 0);
    if (active && flags && flags[col]) active = false;
    Lane<DIM, PHASE> e;
    load_column<DIM, PHASE>(init + col * P.rows, e);
    bool bad = false;
#pragma unroll
    for (int i = 0; i < DIM; ++i) bad = bad || isnan(e.r[i]) || isnan(e.p[i]);
    if (PHASE != ETRAJ_PHASE_CTMC) bad = bad || isnan(e.phase0);
    emit<DIM, PHASE, OUT_PMD>(P, O, col, active, bad, e.r, e.p, e.phase0, e.t0, e.prob);
}

// In-place, order-preserving compaction of the flagged columns of each
// segment (one workgroup of kCompactWaves wavefronts per segment): count_out[s]
// receives their number and orig[seg_start[s] + k] the original column (or its
// out_index entry) of the k-th. Every column of a chunk is read before any is
// written, and writes never pass the read position, so the compaction can
// overwrite in place. The flags of the next chunk are loaded while this one is
// compacted (they are never written here).
constexpr int kCompactWaves = 16;
__global__ __launch_bounds__(64 * kCompactWaves) void k_compact_flagged(double* __restrict__ init, int rows,
                                                                       const int64_t* __restrict__ seg_start,
                                                                       const int32_t* __restrict__ seg_count,
                                                                       const uint8_t* __restrict__ flags,
                                                                       const int64_t* __restrict__ out_index,
                                                                       int32_t* __restrict__ count_out,
                                                                       int64_t* __restrict__ orig) {
    constexpr int kChunk = 64 * kCompactWaves;
    __shared__ int wave_tot[kCompactWaves];
    __shared__ int base_sh;
    const int64_t seg = blockIdx.x;
    const int64_t base = seg_start[seg];
    const int32_t cnt = seg_count[seg];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) base_sh = 0;
    __syncthreads();
    double col[9];
    bool fnext = tid < cnt && flags[base + tid] != 0;
    for (int j0 = 0; j0 < cnt; j0 += kChunk) {
        const int j = j0 + tid;
        const bool keep = fnext;
        const int jn = j + kChunk;
        fnext = jn < cnt && flags[base + jn] != 0;
        if (keep)
            for (int r = 0; r < rows; ++r) col[r] = init[(base + j) * rows + r];
        const uint64_t m = __ballot(keep);
        const int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (lane == 0) wave_tot[wv] = __popcll(m);
        __syncthreads();
        int off = base_sh, tot = 0;
        for (int w = 0; w < kCompactWaves; ++w) {
            const int c = wave_tot[w];
            off += w < wv ? c : 0;
            tot += c;
        }
        if (keep) {
            const int64_t dst = base + off + pre;
            for (int r = 0; r < rows; ++r) init[dst * rows + r] = col[r];
            if (orig) orig[dst] = out_index ? out_index[base + j] : base + j;
        }
        __syncthreads();
        if (tid == 0) base_sh += tot;
        __syncthreads();
    }
    if (tid == 0) count_out[seg] = base_sh;
}

// Exclusive prefix sum of the segment counts (one workgroup):
// pre[k] = sum_{j<k} count[j], pre[n] = total.
__global__ __launch_bounds__(1024) void k_scan_counts(const int32_t* __restrict__ count, int64_t n,
                                                       int64_t* __restrict__ pre) {
    __shared__ int64_t part[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b = 0; b < n + 1; b += 1024) {
        const int64_t i = b + threadIdx.x;
        part[threadIdx.x] = (i < n) ? (int64_t)count[i] : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
            int64_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += v;
            __syncthreads();
        }
        const int64_t excl = carry + (threadIdx.x ? part[threadIdx.x - 1] : 0);
        if (i <= n) pre[i] = excl;
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
}

__global__ void k_fix_to_f64(const uint64_t* __restrict__ q, int64_t n, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = fix_to_double(q[2 * i], q[2 * i + 1]);
}

// Zeroes n 8-byte words (grid-stride): the clearing of the job's sums and PMD, a
// kernel rather than hipMemsetAsync so that the job's launches can be captured
// into a HIP graph.
__global__ void k_zero_words(uint64_t* __restrict__ p, int64_t n) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = 0ull;
}

__global__ void k_finalize_abs2(const double* __restrict__ amp, int64_t n, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double re = amp[2 * i], im = amp[2 * i + 1];
    out[i] = re * re + im * im;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int DIM, int PHASE>
static hipError_t launch_ts5_finish_p(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                                      const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                                      const uint8_t* flags) {
    hipLaunchKernelGGL((k_ts5_finish<DIM, PHASE>), dim3(grid_for(n_seg * (int64_t)cps, 4)), dim3(256), 0, st, P, O,
                       init, seg_start, seg_count, n_seg, cps, flags);
    return hipGetLastError();
}

template <int DIM>
static hipError_t launch_ts5_finish_d(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                                      const int64_t* ss, const int32_t* sc, int64_t n_seg, int cps,
                                      const uint8_t* flags) {
    switch (P.phase) {
        case ETRAJ_PHASE_CTMC:
            return launch_ts5_finish_p<DIM, ETRAJ_PHASE_CTMC>(st, P, O, init, ss, sc, n_seg, cps, flags);
        case ETRAJ_PHASE_QTMC:
            return launch_ts5_finish_p<DIM, ETRAJ_PHASE_QTMC>(st, P, O, init, ss, sc, n_seg, cps, flags);
        default: return launch_ts5_finish_p<DIM, ETRAJ_PHASE_SCTS>(st, P, O, init, ss, sc, n_seg, cps, flags);
    }
}

hipError_t launch_ts5_finish(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                             const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                             const uint8_t* flags) {
    if (n_seg <= 0 || cps <= 0) return hipSuccess;
    return P.dim == 3 ? launch_ts5_finish_d<3>(st, P, O, init, seg_start, seg_count, n_seg, cps, flags)
                      : launch_ts5_finish_d<2>(st, P, O, init, seg_start, seg_count, n_seg, cps, flags);
}

hipError_t launch_laser_field(hipStream_t st, const LaserP& L, const double* t, int64_t n, double* Fx, double* Fy,
                              double* Ax, double* Ay) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_laser_field, dim3(grid_for(n, 256)), dim3(256), 0, st, L, t, n, Fx, Fy, Ax, Ay);
    return hipGetLastError();
}

hipError_t launch_target_eval(hipStream_t st, const TargetP& T, const double* xyz, int64_t n, double* V, double* F) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_target_eval, dim3(grid_for(n, 256)), dim3(256), 0, st, T, xyz, n, V, F);
    return hipGetLastError();
}

template <int DIM, int PHASE, bool MIXED, int OUT>
static hipError_t launch_rk4_t(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                               const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                               const FieldTab& FT, uint8_t* flags, const int64_t* out_index) {
    const int block = (P.rk4_block == 64 || P.rk4_block == 128) ? P.rk4_block : 256;
    const int waves = block / 64;
    const unsigned blocks = (unsigned)((n_seg * (int64_t)cps + waves - 1) / waves);
    if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        hipLaunchKernelGGL((k_rk4<DIM, PHASE, MIXED, OUT, ETRAJ_TARGET_HYDROGEN_LIKE>), dim3(blocks), dim3(block), 0, st,
                           P, O, init, seg_start, seg_count, n_seg, cps, FT, flags, out_index);
    else
        hipLaunchKernelGGL((k_rk4<DIM, PHASE, MIXED, OUT, ETRAJ_TARGET_SAE>), dim3(blocks), dim3(block), 0, st, P, O,
                           init, seg_start, seg_count, n_seg, cps, FT, flags, out_index);
    return hipGetLastError();
}

template <int DIM, int PHASE, bool MIXED>
static hipError_t launch_rk4_o(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                               const int64_t* ss, const int32_t* sc, int64_t ns, int cps, const FieldTab& FT,
                               uint8_t* fl, const int64_t* oi) {
    switch (out) {
        case OUT_PMD: return launch_rk4_t<DIM, PHASE, MIXED, OUT_PMD>(st, P, O, init, ss, sc, ns, cps, FT, fl, oi);
        case OUT_RECORDS:
            return launch_rk4_t<DIM, PHASE, MIXED, OUT_RECORDS>(st, P, O, init, ss, sc, ns, cps, FT, fl, oi);
        default: return launch_rk4_t<DIM, PHASE, MIXED, OUT_CONTRIB>(st, P, O, init, ss, sc, ns, cps, FT, fl, oi);
    }
}

template <int DIM, int PHASE>
static hipError_t launch_rk4_m(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                               const int64_t* ss, const int32_t* sc, int64_t ns, int cps, const FieldTab& FT,
                               uint8_t* fl, const int64_t* oi) {
    return P.mixed ? launch_rk4_o<DIM, PHASE, true>(st, P, O, out, init, ss, sc, ns, cps, FT, fl, oi)
                   : launch_rk4_o<DIM, PHASE, false>(st, P, O, out, init, ss, sc, ns, cps, FT, fl, oi);
}

template <int DIM>
static hipError_t launch_rk4_p(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                               const int64_t* ss, const int32_t* sc, int64_t ns, int cps, const FieldTab& FT,
                               uint8_t* fl, const int64_t* oi) {
    switch (P.phase) {
        case ETRAJ_PHASE_CTMC:
            return launch_rk4_m<DIM, ETRAJ_PHASE_CTMC>(st, P, O, out, init, ss, sc, ns, cps, FT, fl, oi);
        case ETRAJ_PHASE_QTMC:
            return launch_rk4_m<DIM, ETRAJ_PHASE_QTMC>(st, P, O, out, init, ss, sc, ns, cps, FT, fl, oi);
        default: return launch_rk4_m<DIM, ETRAJ_PHASE_SCTS>(st, P, O, out, init, ss, sc, ns, cps, FT, fl, oi);
    }
}

hipError_t launch_rk4(hipStream_t st, const IntegP& P, const OutP& O, int out, const double* init,
                      const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                      const FieldTab& FT, uint8_t* flags, const int64_t* out_index) {
    if (n_seg <= 0 || cps <= 0) return hipSuccess;
    return P.dim == 3 ? launch_rk4_p<3>(st, P, O, out, init, seg_start, seg_count, n_seg, cps, FT, flags, out_index)
                      : launch_rk4_p<2>(st, P, O, out, init, seg_start, seg_count, n_seg, cps, FT, flags, out_index);
}

hipError_t launch_compact_flagged(hipStream_t st, double* init, int rows, const int64_t* seg_start,
                                  const int32_t* seg_count, int64_t n_seg, const uint8_t* flags,
                                  const int64_t* out_index, int32_t* count_out, int64_t* orig) {
    if (n_seg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_flagged, dim3((unsigned)n_seg), dim3(64 * kCompactWaves), 0, st, init, rows, seg_start,
                       seg_count, flags, out_index, count_out, orig);
    return hipGetLastError();
}

hipError_t launch_scan_counts(hipStream_t st, const int32_t* count, int64_t n, int64_t* pre) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, count, n, pre);
    return hipGetLastError();
}

hipError_t launch_field_table(hipStream_t st, const LaserP* lps, const int32_t* seg_job, const double* t0s,
                              int64_t nseg, double tf, double dt, const FieldTab& FT) {
    if (nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_field_table, dim3(grid_for(nseg * FT.stride, 256)), dim3(256), 0, st, lps, seg_job, t0s, nseg,
                       tf, dt, FT.stride, const_cast<float4*>(FT.f32), const_cast<double2*>(FT.f64),
                       const_cast<double2*>(FT.f0));
    return hipGetLastError();
}

hipError_t launch_fix_to_f64(hipStream_t st, const uint64_t* q, int64_t n, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_to_f64, dim3(grid_for(n, 256)), dim3(256), 0, st, q, n, out);
    return hipGetLastError();
}

hipError_t launch_zero(hipStream_t st, void* p, int64_t bytes) {
    if (bytes <= 0) return hipSuccess;
    if (bytes % 8 != 0 || reinterpret_cast<uintptr_t>(p) % 8 != 0) return hipErrorInvalidValue;
    const int64_t n = bytes / 8;
    hipLaunchKernelGGL(k_zero_words, dim3(std::min<unsigned>(grid_for(n, 256), 8192u)), dim3(256), 0, st,
                       static_cast<uint64_t*>(p), n);
    return hipGetLastError();
}

hipError_t launch_finalize_abs2(hipStream_t st, const double* amp, int64_t n, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize_abs2, dim3(grid_for(n, 256)), dim3(256), 0, st, amp, n, out);
    return hipGetLastError();
}

}  // namespace etraj
