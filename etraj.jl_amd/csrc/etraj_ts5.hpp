// etraj_ts5.hpp -- the adaptive Tsit5 step of one electron (device code,
// included by etraj_tsit5.hip only): start values, stage inputs, one
// attempted step with its error norm and step-size controller.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "etraj_physics.hpp"

namespace etraj {

// ---------------------------------------------------------------------------
// Adaptive Tsit5 (Tsitouras 2011) with the defaults OrdinaryDiffEq applies to
// solve(prob, Tsit5(), adaptive=true, reltol=traj_rtol) (TrajectorySimulation.jl:342):
// Hairer's initial step, PI controller beta1 = 7/50, beta2 = 2/25,
// qmin = 1/5, qmax = 10, gamma = 9/10, qoldinit = 1e-4, RMS error norm of
// utilde / (atol + max(|u|, |u_new|) rtol), NaN and maxiters checks.
// The state vector is (r, p[, phase]); the phase takes part in the error norm
// as in the reference's ODE.
// ---------------------------------------------------------------------------
namespace ts5 {
constexpr double c2 = 0.161, c3 = 0.327, c4 = 0.9, c5 = 0.9800255409045097;
constexpr double a21 = 0.161;
constexpr double a31 = -0.008480655492356989, a32 = 0.335480655492357;
constexpr double a41 = 2.897153057105493, a42 = -6.359448489975075, a43 = 4.3622954328695815;
constexpr double a51 = 5.325864828439257, a52 = -11.748883564062828, a53 = 7.4955393428898365,
                 a54 = -0.09249506636175525;
constexpr double a61 = 5.86145544294642, a62 = -12.92096931784711, a63 = 8.159367898576159,
                 a64 = -0.071584973281401, a65 = -0.028269050394068383;
constexpr double a71 = 0.09646076681806523, a72 = 0.01, a73 = 0.4798896504144996, a74 = 1.379008574103742,
                 a75 = -3.290069515436081, a76 = 2.324710524099774;
constexpr double b1 = -0.00178001105222577714, b2 = -0.0008164344596567469, b3 = 0.007880878010261995,
                 b4 = -0.1447110071732629, b5 = 0.5823571654525552, b6 = -0.45808210592918697,
                 b7 = 0.015151515151515152;
}  // namespace ts5

// Per-electron Tsit5 state, advanced one attempted step at a time so that a
// wavefront can refill lanes whose electron has finished. The EOM is second
// order (dr/dt = p, and no derivative depends on the phase), so the stages are
// kept in the Runge-Kutta-Nystrom arrangement of the same tableau: only the p
// and phase parts of k_j are stored; the r part of k_j is the stage's own p,
// whose weighted sums fold into C_s p + dt sum_l (A^2)_sl kp_l. The result is
// the Tsit5 step up to rounding, with about half the live stage registers.
constexpr double kTs5LogQmin = -9.210340371976184;   // log(qoldinit) = log(1e-4)

// 1/x from the hardware reciprocal and two Newton steps (within an ulp of 1/x for the
// positive, normal x of the error norm's denominators)
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

template <int DIM, int PHASE>
struct Ts5Lane {
    double r[DIM], p[DIM], ph;
    double k1p[DIM], k1ph;
    double t, dt, lqold;   // lqold = log(qold) of the PI controller
    // the simple laser's carrier and envelope sin/cos at t, carried from the last accepted step
    // (valid while base_ok != 0)
    double bsw, bcw, bsh, bch;
    int32_t accepted, attempted, base_ok;
};

template <int DIM, int PHASE, int TGT>
__device__ __forceinline__ void ts5_deriv(const TargetP& T, bool saeq, double Fx, double Fy,
                                                   const double* yr, const double* yp, double* kp, double& kph) {
    State<DIM, double> s, ds;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        s.r[i] = yr[i];
        s.p[i] = yp[i];
    }
    double dphi, r2;
    eom<DIM, PHASE, TGT, double>(T, saeq, s, Fx, Fy, ds, dphi, r2);
#pragma unroll
    for (int i = 0; i < DIM; ++i) kp[i] = ds.p[i];
    kph = dphi;
}

// The mixed-precision form of ts5_deriv: the derivative is evaluated in FP32
// (the FP32 EOM of the fixed-step mixed pass) from the FP64 stage values, and
// r2 (|r|^2, plus the soft core for hydrogen-like targets) is returned for the
// close-encounter test; stage values, sums and the error control stay FP64.
template <int DIM, int PHASE, int TGT>
__device__ __forceinline__ void ts5_deriv32(const TargetP& T, bool saeq, double Fx, double Fy,
                                                     const double* yr, const double* yp, double* kp, double& kph,
                                                     float& r2) {
    State<DIM, float> s, ds;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        s.r[i] = (float)yr[i];
        s.p[i] = (float)yp[i];
    }
    float dphi;
    eom<DIM, PHASE, TGT, float>(T, saeq, s, (float)Fx, (float)Fy, ds, dphi, r2);
#pragma unroll
    for (int i = 0; i < DIM; ++i) kp[i] = (double)ds.p[i];
    kph = (double)dphi;
}

// stage derivative in FP64, or in FP32 with the close-encounter test (MIXED)
template <int DIM, int PHASE, int TGT, bool MIXED>
__device__ __forceinline__ void ts5_stage_deriv(const TargetP& T, bool saeq, double Fx, double Fy,
                                                         const double* yr, const double* yp, double* kp, double& kph,
                                                         float rc2, bool& close) {
    if (MIXED) {
        float r2;
        ts5_deriv32<DIM, PHASE, TGT>(T, saeq, Fx, Fy, yr, yp, kp, kph, r2);
        close = close || (r2 < rc2);
    } else {
        ts5_deriv<DIM, PHASE, TGT>(T, saeq, Fx, Fy, yr, yp, kp, kph);
    }
}

// Hairer's initial step (OrdinaryDiffEq's ode_determine_initdt) and the first
// stage derivative. SL: the laser is simple_laser and b0 its phases at t0
// (mono_base), from which the fields at t0 and t0 + dt0 follow by angle
// addition. The norms multiply by reciprocals of their scales, and
// 10^(-(2 + log10 mx)/5) is taken as exp(-(2 ln 10 + ln mx)/5): all within ulps
// of the quotients and the power.
template <int DIM, int PHASE, int TGT, bool SL = false>
__device__ __forceinline__ void ts5_init(const LaserP& L, const TargetP& T, bool saeq, double t0, double tf,
                                         double rtol, double atol, Ts5Lane<DIM, PHASE>& s,
                                         const MonoBase* b0 = nullptr) {
    constexpr bool PH = PHASE != ETRAJ_PHASE_CTMC;
    constexpr int N = 2 * DIM + (PH ? 1 : 0);
    s.t = t0;
    s.lqold = kTs5LogQmin;
    s.base_ok = 0;
    s.accepted = 0;
    s.attempted = 0;
    const double dtmax = tf - t0;
    const double dtmin0 = fmax(0.0, fabs(t0) * 0x1.0p-52);
    const double smalldt = fmax(dtmin0, 1e-6);
    double u[N], isk[N], d0s = 0, d1s = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        u[i] = s.r[i];
        u[DIM + i] = s.p[i];
    }
    if (PH) u[N - 1] = s.ph;
    double Fx, Fy;
    if (SL) mono_field_at<true>(L.c[0], *b0, 0.0, Fx, Fy);
    else laser_field(L, t0, Fx, Fy);
    ts5_deriv<DIM, PHASE, TGT>(T, saeq, Fx, Fy, s.r, s.p, s.k1p, s.k1ph);
    double f0[N];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        f0[i] = s.p[i];
        f0[DIM + i] = s.k1p[i];
    }
    if (PH) f0[N - 1] = s.k1ph;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        isk[i] = recip(atol + fabs(u[i]) * rtol);
        const double a = u[i] * isk[i], b = f0[i] * isk[i];
        d0s += a * a;
        d1s += b * b;
    }
    const double d0 = sqrt(d0s * (1.0 / N)), d1 = sqrt(d1s * (1.0 / N));
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
    dt0 = fmin(dt0, dtmax);
    double yr[DIM], yp[DIM], k2p[DIM], k2ph;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        yr[i] = u[i] + dt0 * f0[i];
        yp[i] = u[DIM + i] + dt0 * f0[DIM + i];
    }
    if (SL) mono_field_at<true>(L.c[0], *b0, dt0, Fx, Fy);
    else laser_field(L, t0 + dt0, Fx, Fy);
    ts5_deriv<DIM, PHASE, TGT>(T, saeq, Fx, Fy, yr, yp, k2p, k2ph);
    double d2s = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double f1 = i < DIM ? yp[i] : (i < 2 * DIM ? k2p[i - DIM] : k2ph);
        const double a = (f1 - f0[i]) * isk[i];
        d2s += a * a;
    }
    const double d2 = sqrt(d2s * (1.0 / N)) / dt0;
    const double mx = fmax(d1, d2);
    constexpr double kLn10 = 2.302585092994045684;
    const double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : exp(-0.2 * (2 * kLn10 + log(mx)));
    s.dt = fmax(dtmin0, fmin(fmin(100 * dt0, dt1), dtmax));
}

namespace ts5n {
// Tsit5 tableau as arrays (row s = stage s+1, 0-based), and the products the
// second-order (r' = p) form needs: r-stage sums use sum_j a_sj yp_j with
// yp_j = p + dt sum_l a_jl kp_l, i.e. C_s p + dt sum_l A2_sl kp_l.
struct Tab {
    double a[7][7], c[7], A2[7][7], b[7], BA[7], B6;
};
constexpr Tab make() {
    Tab t{};
    using namespace ts5;
    const double A[7][7] = {{0, 0, 0, 0, 0, 0, 0},
                            {a21, 0, 0, 0, 0, 0, 0},
                            {a31, a32, 0, 0, 0, 0, 0},
                            {a41, a42, a43, 0, 0, 0, 0},
                            {a51, a52, a53, a54, 0, 0, 0},
                            {a61, a62, a63, a64, a65, 0, 0},
                            {a71, a72, a73, a74, a75, a76, 0}};
    const double B[7] = {b1, b2, b3, b4, b5, b6, b7};
    for (int s = 0; s < 7; ++s) {
        double cs = 0;
        for (int j = 0; j < 7; ++j) {
            t.a[s][j] = A[s][j];
            cs += A[s][j];
        }
        t.c[s] = cs;
        for (int l = 0; l < 7; ++l) {
            double v = 0;
            for (int j = l + 1; j < s; ++j) v += A[s][j] * A[j][l];
            t.A2[s][l] = v;
        }
        t.b[s] = B[s];
    }
    t.B6 = 0;
    for (int j = 0; j < 6; ++j) t.B6 += B[j];
    for (int l = 0; l < 7; ++l) {
        double v = 0;
        for (int j = l + 1; j < 6; ++j) v += B[j] * A[j][l];
        t.BA[l] = v;
    }
    return t;
}
constexpr Tab tab = make();
}  // namespace ts5n

// Inputs of stage S (1-based, 2..7): yp = p + dt sum_{l<S} a_Sl kp_l,
// yr = r + dt (C_S p + dt sum_{l<S-1} A2_Sl kp_l).
template <int S, int DIM>
__device__ __forceinline__ void ts5_inputs(const double* r, const double* p, const double (*kp)[DIM],
                                                    double dt, double* yr, double* yp) {
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double sp = ts5n::tab.a[S - 1][0] * kp[0][i];
#pragma unroll
        for (int l = 1; l < S - 1; ++l) sp += ts5n::tab.a[S - 1][l] * kp[l][i];
        yp[i] = p[i] + dt * sp;
        double sr = 0;
        if (S > 2) {
            sr = ts5n::tab.A2[S - 1][0] * kp[0][i];
#pragma unroll
            for (int l = 1; l < S - 2; ++l) sr += ts5n::tab.A2[S - 1][l] * kp[l][i];
        }
        yr[i] = r[i] + dt * (ts5n::tab.c[S - 1] * p[i] + dt * sr);
    }
}

// One attempted step. MIXED: stage derivatives in FP32 (ts5_deriv32); `close`
// is set when a stage position comes within the re-step radius (r2 < rc2).
// SL: simple_laser(L) holds (one cos^n colour of N >= 2 cycles, azimuth 0), which keeps the
// laser parameters the stage fields read few enough for the scalar registers.
template <int DIM, int PHASE, int TGT, bool MIXED = false, bool SL = false>
__device__ __forceinline__ int ts5_attempt(const LaserP& L, const TargetP& T, bool saeq, double t0,
                                                    double tf, double rtol, double atol, int64_t maxiters,
                                                    Ts5Lane<DIM, PHASE>& s, float rc2 = 0.0f,
                                                    bool* close_out = nullptr) {
    bool close = false;
    constexpr bool PH = PHASE != ETRAJ_PHASE_CTMC;
    constexpr int N = 2 * DIM + (PH ? 1 : 0);
    double dt = s.dt;
    const double t = s.t;
    if (tf - t < dt) dt = tf - t;
    if (++s.attempted > maxiters) return -1;
    double kp[6][DIM], kh[6];
#pragma unroll
    for (int i = 0; i < DIM; ++i) kp[0][i] = s.k1p[i];
    kh[0] = s.k1ph;
    double yr[DIM], yp[DIM];
    ts5_inputs<2, DIM>(s.r, s.p, kp, dt, yr, yp);
    // SL: the laser's phases at t carried from the start values (k_ts5_prep) or the last accepted
    // step's stage 7, by angle addition; their rounding errors add up by about an ulp per
    // accepted step (1e-13 after a thousand steps), the same whatever the launch's scheduling.
    // A step after one whose stage 7 lay before the pulse window takes them afresh.
    LaserBase lb;
    if (SL && s.base_ok) {
        lb.c[0].tp = t - L.c[0].tshift;
        lb.c[0].sw = s.bsw;
        lb.c[0].cw = s.bcw;
        lb.c[0].sh = s.bsh;
        lb.c[0].ch = s.bch;
        lb.c[1] = lb.c[0];
    } else {
        lb = laser_base<SL>(L, t);
    }
    // SL: the field of stages 6 and 7 (t + dt) first; stages 4 and 5 turn its phases back by
    // 0.1 and 0.02 of w dt (order-3 and order-2 Taylor series), and the carrier increments of
    // stages 2 and 3, 0.161 and 0.327 of w dt from t, take the order-4 and order-5 series,
    // wherever the last stage's |w dt| <= 1/2 takes the order-7 one
    double F1x, F1y;                     // stages 6 and 7 share t + dt (c6 = c7 = 1)
    MonoBase at;                         // SL: the phases at t + dt
    if (SL) mono_field_at<true>(L.c[0], lb.c[0], dt, F1x, F1y, &at);
    double Fx, Fy;
    if (SL) mono_field_at<true, 4>(L.c[0], lb.c[0], ts5::c2 * dt, Fx, Fy);
    else laser_field_at<SL>(L, lb, ts5::c2 * dt, Fx, Fy);
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, Fx, Fy, yr, yp, kp[1], kh[1], rc2, close);
    ts5_inputs<3, DIM>(s.r, s.p, kp, dt, yr, yp);
    if (SL) mono_field_at<true, 5>(L.c[0], lb.c[0], ts5::c3 * dt, Fx, Fy);
    else laser_field_at<SL>(L, lb, ts5::c3 * dt, Fx, Fy);
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, Fx, Fy, yr, yp, kp[2], kh[2], rc2, close);
    ts5_inputs<4, DIM>(s.r, s.p, kp, dt, yr, yp);
    if (SL) simple_field_back<3, 3>(L.c[0], lb.c[0], ts5::c4 * dt, at, (1.0 - ts5::c4) * dt, Fx, Fy);
    else laser_field_at<SL>(L, lb, ts5::c4 * dt, Fx, Fy);
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, Fx, Fy, yr, yp, kp[3], kh[3], rc2, close);
    ts5_inputs<5, DIM>(s.r, s.p, kp, dt, yr, yp);
    if (SL) simple_field_back<2, 2>(L.c[0], lb.c[0], ts5::c5 * dt, at, (1.0 - ts5::c5) * dt, Fx, Fy);
    else laser_field_at<SL>(L, lb, ts5::c5 * dt, Fx, Fy);
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, Fx, Fy, yr, yp, kp[4], kh[4], rc2, close);
    ts5_inputs<6, DIM>(s.r, s.p, kp, dt, yr, yp);
    if (!SL) laser_field_at<SL>(L, lb, dt, F1x, F1y);
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, F1x, F1y, yr, yp, kp[5], kh[5], rc2, close);
    double unr[DIM], unp[DIM], er[DIM], ep[DIM], unh = 0, eh = 0;
    ts5_inputs<7, DIM>(s.r, s.p, kp, dt, unr, unp);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double sr = ts5n::tab.BA[0] * kp[0][i], sp = ts5n::tab.b[0] * kp[0][i];
#pragma unroll
        for (int l = 1; l < 6; ++l) {
            if (l < 5) sr += ts5n::tab.BA[l] * kp[l][i];
            sp += ts5n::tab.b[l] * kp[l][i];
        }
        er[i] = ts5n::tab.B6 * s.p[i] + dt * sr;
        ep[i] = sp;
    }
    if (PH) {
        double sa = ts5n::tab.a[6][0] * kh[0], sb = ts5n::tab.b[0] * kh[0];
#pragma unroll
        for (int l = 1; l < 6; ++l) {
            sa += ts5n::tab.a[6][l] * kh[l];
            sb += ts5n::tab.b[l] * kh[l];
        }
        unh = s.ph + dt * sa;
        eh = sb;
    }
    double k7p[DIM], k7h;
    ts5_stage_deriv<DIM, PHASE, TGT, MIXED>(T, saeq, F1x, F1y, unr, unp, k7p, k7h, rc2, close);
    // RMS error norm and PI controller in FP64, arranged so that they need one log and one exp
    // and no division or square root: the norm's ratios use a Newton-refined hardware reciprocal
    // of the denominator (within an ulp of the quotient), EEst <= 1 is tested as EEst^2 <= 1 and
    // log(EEst) taken as log(EEst^2)/2, and the new step is dt times 1/q, with q = EEst^(7/50) /
    // qold^(2/25) clamped to [1/qmax, 1/qmin] after the division by gamma -- on a rejection
    // q11 = EEst^(7/50) alone -- evaluated as gamma exp(-log q) clamped to [qmin, qmax].
    const double b7 = ts5::b7;
    double ss = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        const double a = dt * (er[i] + b7 * unp[i]) * recip(atol + fmax(fabs(s.r[i]), fabs(unr[i])) * rtol);
        const double b = dt * (ep[i] + b7 * k7p[i]) * recip(atol + fmax(fabs(s.p[i]), fabs(unp[i])) * rtol);
        ss = fma(a, a, ss);
        ss = fma(b, b, ss);
    }
    if (PH) {
        const double a = dt * (eh + b7 * k7h) * recip(atol + fmax(fabs(s.ph), fabs(unh)) * rtol);
        ss = fma(a, a, ss);
    }
    if (MIXED && close_out) *close_out = *close_out || close;
    const double E2 = ss * (1.0 / N);   // EEst^2
    if (isnan(E2)) return -1;
    const bool accept = E2 <= 1.0;
    const double lE = E2 == 0.0 ? kTs5LogQmin : 0.5 * log(E2);   // log(EEst)
    // 1/q (accept) or 1/min(1/qmin, q11/gamma) (reject)
    const double iq = E2 == 0.0 ? 10.0
                    : accept ? fmin(10.0, fmax(0.2, 0.9 * exp((2.0 / 25.0) * s.lqold - (7.0 / 50.0) * lE)))
                             : fmax(0.2, 0.9 * exp(-(7.0 / 50.0) * lE));
    if (accept) {
        s.t = (tf - t == dt) ? tf : t + dt;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            s.r[i] = unr[i];
            s.p[i] = unp[i];
            s.k1p[i] = k7p[i];
        }
        s.ph = unh;
        s.k1ph = k7h;
        ++s.accepted;
        if (SL) {
            // at.tp: 1 phases at t + dt; 2 past the pulse window, where no later stage needs them
            // (the window is exactly zero there and stays so); 0 before the window
            s.base_ok = at.tp != 0.0;
            if (at.tp == 1.0) {
                s.bsw = at.sw;
                s.bcw = at.cw;
                s.bsh = at.sh;
                s.bch = at.ch;
            }
        }
        s.lqold = fmax(lE, kTs5LogQmin);   // log(max(EEst, qoldinit))
        s.dt = fmin(dt * iq, tf - t0);
    } else {
        s.dt = dt * iq;
    }
    if (!(s.t < tf)) return 1;
    if (fabs(s.dt) < fmax(fabs(t0), fabs(tf)) * 0x1.0p-52) return -1;
    return 0;
}

}  // namespace etraj
