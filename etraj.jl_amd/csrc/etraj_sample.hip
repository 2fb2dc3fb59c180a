// etraj_sample.hip -- the electron samplers (k_adk_sample: ADK, SFA-SPA,
// SFA-SPANE, WFAT) and their launcher.
//
// A file of its own because it is built with machine-level loop-invariant code
// motion off (-mllvm -disable-machine-licm), as etraj_tsit5.hip is: hoisted out
// of the saddle-point and quadrature loops, the FP64 constants of the laser's
// complex-time evaluation pin so many registers that the SFA-SPA form spills
// about 100 values to scratch; rematerialised inside the loops they cost scalar
// moves only. The RK4 kernels (etraj_kernels.hip) run faster with the motion on.
#include "etraj_engine.hpp"
#include "etraj_device.hpp"

#include <hip/hip_runtime.h>

namespace etraj {

// ---------------------------------------------------------------------------
// ADK sampler (ADKSampler.jl:167-396), atoms with l = 0: the prefactor's
// Wigner-D and Y_00 factors are constants, so amp is real and positive and
// the initial phase angle(amp) is 0.
// ---------------------------------------------------------------------------
struct AdkBatch {
    double tr, F, cphi, sphi, c_cc, dkdt;
    double g2;          // WFAT table form: |G00(beta, gamma)|^2 of this batch's field direction
    double beta, gamma; // WFAT: Euler angles FF -> MF of this batch (WFATSampler.jl:143-145)
    const double* wB;   // WFAT: B_|m| (LDS), rate = dkdt E(k) sum_|m| B_|m| k^(2|m|)
    int wM;             // WFAT: m_max of wB
    double alpha;       // prefactor beyond l = 0: Euler angles FF -> frame of C_lm (with beta, gamma)
    double ex, ey;      // F/|F| at tr
    double Fxr, Fyr;    // F at tr
    const double* pT;   // prefactor beyond l = 0: the batch's monomial coefficients (LDS, (re, im))
    const LaserTr* lt;  // SPA: the laser at tr, tr + 0.01, tr - 0.01 (LDS; complex-time evaluations)
    double dFx, dFy;    // SPANE: dF/dt at tr
    double cross;       // SPANE: Fx dFy - dFx Fy (SPANESampler.jl:238)
    double iF, i3F, sqF;   // 1/F, 1/(3F) and sqrt(F) at tr
    bool skip;          // SPA: F(tr) <= 1e-3 empties the batch (SPASampler.jl:167-169)
};

// params(RotZYZ(MF^-1 FF)) (shared_methods.jl:31-38, Rotations.jl ZYZ
// decomposition with beta in [0, pi]); returns (beta, gamma).
__device__ inline void ff_mf_euler(const double* mr, double Fx, double Fy, double& alpha, double& beta, double& gamma) {
    const double F = hypot(Fx, Fy);
    const double FF[3][3] = {{0, Fy / F, Fx / F}, {0, -Fx / F, Fy / F}, {1, 0, 0}};
    double sa, ca, sb, cb, sc, cc;
    sincos(mr[0], &sa, &ca);
    sincos(mr[1], &sb, &cb);
    sincos(mr[2], &sc, &cc);
    // MF = Rz(a) Ry(b) Rz(c)
    const double MF[3][3] = {{ca * cb * cc - sa * sc, -ca * cb * sc - sa * cc, ca * sb},
                             {sa * cb * cc + ca * sc, -sa * cb * sc + ca * cc, sa * sb},
                             {-sb * cc, sb * sc, cb}};
    double R[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) R[i][j] = MF[0][i] * FF[0][j] + MF[1][i] * FF[1][j] + MF[2][i] * FF[2][j];
    const double t1 = atan2(R[1][2], R[0][2]);
    double st, ct;
    sincos(t1, &st, &ct);
    alpha = t1;
    beta = atan2(ct * R[0][2] + st * R[1][2], R[2][2]);
    gamma = atan2(-st * R[0][0] + ct * R[1][0], -st * R[0][1] + ct * R[1][1]);
}

// Wigner small-d d^l_{m m'}(beta) = <l m| exp(-i beta J_y) |l m'> by Wigner's
// formula: sum_s (-1)^(m-m'+s) sqrt((l+m)!(l-m)!(l+m')!(l-m')!) /
// ((l+m'-s)! s! (m-m'+s)! (l-m-s)!) cos(beta/2)^(2l+m'-m-2s) sin(beta/2)^(m-m'+2s).
// l <= 30 (factorials to 60! stay finite in FP64).
__device__ inline double wigner_d(int l, int m, int mp, double beta) {
    double c, sn;
    sincos(0.5 * beta, &sn, &c);
    auto fact = [](int n) {
        double f = 1.0;
        for (int i = 2; i <= n; ++i) f *= i;
        return f;
    };
    const double pre = sqrt(fact(l + m) * fact(l - m) * fact(l + mp) * fact(l - mp));
    const int s0 = max(0, mp - m), s1 = min(l + mp, l - m);
    double sum = 0.0;
    for (int k = s0; k <= s1; ++k) {
        const double den = fact(l + mp - k) * fact(k) * fact(m - mp + k) * fact(l - m - k);
        const double t = pre / den * pow(c, 2 * l + mp - m - 2 * k) * pow(sn, m - mp + 2 * k);
        sum += ((m - mp + k) & 1) ? -t : t;
    }
    return sum;
}

// |G_{nxi m}(beta, gamma)|^2 exp(-2 kappa mu_z) of one channel from the WFAT
// integrals (MolWFATStructureFactor_G, GenericMolecule.jl:236-260). The
// integrals are Julia's Array{ComplexF64,4} I[nxi, m + m_max, l, m' + l]
// (0-based), column-major, (re, im).
__device__ inline double wfat_channel_g2(const SamplerP& S, double kappa, double beta, double gamma, int nxi, int m) {
    const int D0 = S.wfat_dims[0], D1 = S.wfat_dims[1], D2 = S.wfat_dims[2];
    const int mM = (D1 - 1) / 2, L = D2 - 1;
    double gre = 0.0, gim = 0.0;
    for (int l = abs(m); l <= L; ++l)
        for (int mp = -l; mp <= l; ++mp) {
            const double* I = S.wfat_coeff + 2 * (nxi + (int64_t)D0 * ((m + mM) + (int64_t)D1 * (l + (int64_t)D2 * (mp + l))));
            const double d = wigner_d(l, m, mp, beta);
            double sg, cg;
            sincos(mp * gamma, &sg, &cg);   // e^{-i m' gamma} = cg - i sg
            gre += d * (I[0] * cg + I[1] * sg);
            gim += d * (I[1] * cg - I[0] * sg);
        }
    const double muz = -sin(beta) * S.wfat_mu[0] + cos(beta) * S.wfat_mu[2];   // (RotZYZ(gamma, beta, 0) mu)_z
    return (gre * gre + gim * gim) * exp(-2.0 * kappa * muz);
}

// Field-factor prefactor of channel (nxi, |m|) (WFATSampler.jl:157 without
// its k-dependent factors): kappa^(|m|+2) / (2 F^(|m|+1) |m|!) (4 kappa^2/F)^(2n - 2nxi - |m| - 1).
__device__ inline double wfat_field_prefactor(double kappa, double F, double nstar, int nxi, int am) {
    double f = 1.0;
    for (int i = 2; i <= am; ++i) f *= i;
    return pow(kappa, am + 2) / (2.0 * pow(F, am + 1) * f) * pow(4 * kappa * kappa / F, 2 * nstar - 2 * nxi - am - 1);
}

// |G00|^2 with G00 = sum_{a,b} c_ab P_a(cos beta) e^{-i b gamma}
// (Legendre recurrence in a, complex power recurrence in b)
__device__ inline double wfat_g2(const SamplerP& S, double beta, double gamma) {
    const double x = cos(beta);
    double sg, cg;
    sincos(gamma, &sg, &cg);
    double gre = 0.0, gim = 0.0, pm1 = 0.0, p = 1.0;
    for (int a = 0; a < S.wfat_na; ++a) {
        if (a == 1) { pm1 = 1.0; p = x; }
        else if (a > 1) { const double pn = ((2 * a - 1) * x * p - (a - 1) * pm1) / a; pm1 = p; p = pn; }
        double er = 1.0, ei = 0.0;   // e^{-i b gamma}
        for (int b = 0; b < S.wfat_nb; ++b) {
            const double* c = S.wfat_coeff + 2 * (a * S.wfat_nb + b);
            gre += p * (c[0] * er - c[1] * ei);
            gim += p * (c[0] * ei + c[1] * er);
            const double nr = er * cg + ei * sg, ni = ei * cg - er * sg;
            er = nr;
            ei = ni;
        }
    }
    return gre * gre + gim * gim;
}

template <int M, bool PRE>
__device__ inline AdkBatch adk_batch_setup(const SamplerP& S, const LaserP& L, const TargetP& T, double tr) {
    AdkBatch B;
    B.tr = tr;
    double Fx, Fy;
    B.dFx = B.dFy = B.cross = 0.0;
    if (M == ETRAJ_INIT_SPANE) {
        laser_field_deriv(L, tr, Fx, Fy, B.dFx, B.dFy);              // SPANESampler.jl:164-167
        B.cross = Fx * B.dFy - B.dFx * Fy;
    } else {
        laser_field(L, tr, Fx, Fy);
    }
    B.F = hypot(Fx, Fy);                                            // :172
    B.iF = 1.0 / B.F;
    B.i3F = 1.0 / (3 * B.F);
    B.sqF = sqrt(B.F);
    double phi = atan2(-Fy, -Fx);                                   // :173
    sincos(phi, &B.sphi, &B.cphi);
    double Ip = S.Ip;
    double gam = 0.0;                                               // :183-187
    if (L.n == 1) gam = L.c[0].w * sqrt(2 * Ip) / (L.c[0].F0 * mono_envelope(L.c[0], tr));
    double kappa = sqrt(2 * Ip), nn = T.Zd / kappa;
    const double e0 = 2.71828182845904523;
    B.c_cc = pow(2.0, 3 * nn / 2 + 1) * pow(kappa, 5 * nn + 0.5) * pow(B.F, -nn) * pow(1 + 2 * gam / e0, -nn);  // :235
    B.dkdt = S.dkdt_base;
    B.g2 = 0.0;
    B.alpha = B.beta = B.gamma = 0.0;
    B.wB = nullptr;
    B.wM = 0;
    B.pT = nullptr;
    B.lt = nullptr;
    B.ex = Fx / B.F;
    B.ey = Fy / B.F;
    B.Fxr = Fx;
    B.Fyr = Fy;
    B.skip = (M == ETRAJ_INIT_SPA) && !(B.F > 1e-3);
    if (M == ETRAJ_INIT_WFAT) {
        ff_mf_euler(S.mol_rot, Fx, Fy, B.alpha, B.beta, B.gamma);   // WFATSampler.jl:143-145
        if (S.wfat_form == ETRAJ_WFAT_TABLE) B.g2 = wfat_g2(S, B.beta, B.gamma);   // :151-155, channel (0, 0)
    } else if (PRE) {
        ff_mf_euler(S.pre_rot, Fx, Fy, B.alpha, B.beta, B.gamma);  // ADKSampler.jl:198-204
    }
    return B;
}

// ---- tunnelling prefactor beyond l = 0 (ADKSampler.jl:188-265, SPASampler.jl:
// 204-278, SPANESampler.jl:208-266): N(k) = sum_{l,m,m'} C_lm D^l_{m' m}(alpha,
// beta, gamma) Y_{l m'}(k'), Y of shared_methods.jl:15-24,
//   Y_{lm'} = (-1)^m' sqrt((2l+1)/4pi (l-m')!/(l+m')!) (i/kappa)^l (l+m')!
//             sum_k (-1)^k / (2^(2k+m') (k+m')! k! (l-m'-2k)!) u^(k+m') v^k z^(l-m'-2k),
// u = x + iy, v = x - iy. Each (l, m', k) is one monomial u^a v^b z^c with
// a = k+m', b = k, c = l-m'-2k (and back: l = a+b+c, m' = a-b, k = b), so per
// batch the sum is regrouped into one coefficient T_abc per monomial, a+b+c <= L,
// stored in the order of pre_eval's loops.
__device__ inline double factd(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

__device__ inline void pre_monomial(const SamplerP& S, const AdkBatch& B, int idx, double out[2]) {
    const int L = S.pre_L;
    int a = 0, b = 0, c = 0, i = idx;
    for (a = 0; a <= L; ++a) {                       // idx -> (a, b, c) in pre_eval's order
        const int na = (L - a + 1) * (L - a + 2) / 2;
        if (i < na) break;
        i -= na;
    }
    for (b = 0; b <= L - a; ++b) {
        const int nb = L - a - b + 1;
        if (i < nb) break;
        i -= nb;
    }
    c = i;
    const int l = a + b + c, m = a - b, k = b;
    // A_{l m} = sum_{m0} C_{l m0} e^{-i m alpha} d^l_{m m0}(beta) e^{-i m0 gamma}
    double are = 0.0, aim = 0.0;
    for (int m0 = -l; m0 <= l; ++m0) {
        const double* C = S.pre_C + 2 * (l + (int64_t)(L + 1) * (m0 + l));
        if (C[0] == 0.0 && C[1] == 0.0) continue;                     // :258-259
        const double d = wigner_d(l, m, m0, B.beta);
        double sp, cp;
        sincos(-(m * B.alpha + m0 * B.gamma), &sp, &cp);
        are += d * (C[0] * cp - C[1] * sp);
        aim += d * (C[0] * sp + C[1] * cp);
    }
    const double kappa = sqrt(2 * S.Ip);
    double coef = sqrt((2 * l + 1) / (4 * kPi) * factd(l - m) / factd(l + m)) * factd(l + m) /
                  (pow(2.0, 2 * k + m) * factd(k + m) * factd(k) * factd(l - m - 2 * k)) / pow(kappa, l);
    if ((m + k) & 1) coef = -coef;                                   // (-1)^m' (-1)^k
    double pre = are * coef, pim = aim * coef;                       // times i^l
    for (int q = 0; q < (l & 3); ++q) {
        const double t = pre;
        pre = -pim;
        pim = t;
    }
    out[0] = pre;
    out[1] = pim;
}

// N(kts): the FF components of conj(kts) (obtain_xyz_FF_LF; LinearAlgebra.dot
// conjugates its first argument), then sum_abc T_abc u^a v^b z^c.
__device__ inline Cx pre_eval(const SamplerP& S, const AdkBatch& B, Cx kx, Cx ky, Cx kz) {
    const Cx cx{kx.re, -kx.im}, cy{ky.re, -ky.im}, cz{kz.re, -kz.im};
    const Cx nx = cz;                                                   // x' = (0, 0, 1)
    const Cx ny = B.ey * cx - B.ex * cy;                                // y' = (Fy, -Fx, 0)/F
    const Cx nz = B.ex * cx + B.ey * cy;                                // z' = (Fx, Fy, 0)/F
    const Cx u{nx.re - ny.im, nx.im + ny.re}, v{nx.re + ny.im, nx.im - ny.re};
    const int L = S.pre_L;
    Cx sum{0.0, 0.0}, ua{1.0, 0.0};
    int idx = 0;
    for (int a = 0; a <= L; ++a) {
        Cx uvb = ua;
        for (int b = 0; b <= L - a; ++b) {
            Cx t = uvb;
            for (int c = 0; c <= L - a - b; ++c) {
                const double* T = B.pT + 2 * idx++;
                sum = sum + Cx{T[0], T[1]} * t;
                t = t * nz;
            }
            uvb = uvb * v;
        }
        ua = ua * u;
    }
    return sum;
}

// SFA-SPA (SPASampler.jl:230-292). With px, py of :231-232 the saddle-point
// bracket (px + Ax)^2 + (py + Ay)^2 is real and equals kd^2 - |Im A(ts)|^2,
// so the equation of :233-238 is f(ti) = (kd^2 + kz^2 - |Im A|^2)/2 + Ip = 0
// with f'(ti) = Im Ax Re Fx + Im Ay Re Fy (dA/dti = -i F). Newton from the
// reference's initial guess (:240), or from ti0 > 0 (a start near the root: the
// Jacobian's solves start at the sample's own root moved to first order in the
// change of tr or kd, and reach their roots in fewer iterations, to the same
// convergence test); accepted if
// |f| <= 1e-8 (NLsolve's ftol) and ti > 0 (:241-246), else 0. Fx, Fy, Ax, Ay
// receive the laser at tr + i t for the returned t (the acceptance test's
// evaluation, which the callers would otherwise repeat); for a failed solve
// only with AT_FAIL.
template <bool AT_FAIL>
__device__ inline double spa_solve(const LaserP& L, const LaserTr& LT, double tr, double Ftr, double omega, double Ip,
                                   double kd, double kz, Cx& Fx, Cx& Fy, Cx& Ax, Cx& Ay, double ti0 = 0.0) {
    const double k2 = kd * kd + kz * kz;
    double ti = ti0 > 0.0 ? ti0 : asinh(omega / Ftr * sqrt(k2 + 2 * Ip)) / omega;
    bool ok = true;
#pragma unroll 1
    for (int it = 0; it < 100; ++it) {
        laser_field_pot_tau(L, LT, tr, ti, Fx, Fy, Ax, Ay);
        const double f = (k2 - (Ax.im * Ax.im + Ay.im * Ay.im)) / 2 + Ip;
        const double step = f / (Ax.im * Fx.re + Ay.im * Fy.re);
        if (!isfinite(step)) {
            ok = false;
            break;
        }
        ti -= step;
        if (fabs(step) <= 1e-15 * fabs(ti)) break;
    }
    if (ok) {
        laser_field_pot_tau(L, LT, tr, ti, Fx, Fy, Ax, Ay);
        const double f = (k2 - (Ax.im * Ax.im + Ay.im * Ay.im)) / 2 + Ip;
        if (fabs(f) <= 1e-8 && ti > 0) return ti;
    }
    if (AT_FAIL) laser_field_pot_tau(L, LT, tr, 0.0, Fx, Fy, Ax, Ay);   // a failed solve: the laser at t = 0
    return 0.0;
}

// px_/py_ (SPASampler.jl:231-232)
__device__ inline void spa_p(double kd, Cx Ax, Cx Ay, double& px, double& py) {
    const double rho = sqrt(Ax.im * Ax.im + Ay.im * Ay.im);
    px = kd * Ay.im / rho - Ax.re;
    py = -kd * Ax.im / rho - Ay.re;
}

// SFA-SPA in two phases per sample, each on the sample's own lane, with the
// saddle-point results passed through LDS (SpaSlots, one column per thread):
// spa_saddles solves the five saddle-point equations of a sample (its own,
// and the four at tr +- dtr, kd +- dkd of the Jacobian) and spa_amp does the
// rest. Every operation and its order are those of one pass over the sample;
// the split keeps the Newton solver and the quadrature from being live at once
// (one pass takes 256 VGPRs and spills about 100 values to scratch).
constexpr int kSpaSlots = 19;   // ti, px, py, (Fx, Fy, Ax, Ay)(ts), pq[4][2]
struct SpaSlots {
    double* v;                  // LDS: slot k of this thread at v[k * 256]
    __device__ double& operator[](int k) const { return v[k * 256]; }
};

// SFA-SPA phase 1 (SPASampler.jl:230-246, :280-292): the saddle point ts = tr +
// i ti of (kd, kz), F and A at ts, the momentum (px, py) of :231-232, and the
// momenta at the saddle points of (tr +- dtr, kd) and (tr, kd +- dkd). A failed
// solve gives ti = 0 (the sample is discarded) and, in the Jacobian, a NaN
// momentum, as in the reference.
template <bool JAC>
__device__ inline void spa_saddles(const SamplerP& S, const LaserP& L, const AdkBatch& B, double kd, double kz,
                                   SpaSlots o) {
    const double Ip = S.Ip, tr = B.tr;
    Cx Fx, Fy, Ax, Ay;
    const double ti = spa_solve<false>(L, B.lt[0], tr, B.F, S.omega, Ip, kd, kz, Fx, Fy, Ax, Ay);
    o[0] = ti;
    if (ti == 0.0) return;
    double px, py;
    spa_p(kd, Ax, Ay, px, py);
    o[1] = px;
    o[2] = py;
    o[3] = Fx.re; o[4] = Fx.im; o[5] = Fy.re; o[6] = Fy.im;
    o[7] = Ax.re; o[8] = Ax.im; o[9] = Ay.re; o[10] = Ay.im;
    if (!JAC) return;
    const double dtr = 0.01, dkd = 0.01;
    // start values of the four solves: the root moved to first order in dtr (df/dtr =
    // Im Ax Im Fx + Im Ay Im Fy, as dA/dtr = -F) and exactly in the kd-dependence of f
    // (kd dkd + dkd^2 / 2); the solves converge from there to the same test
    const double dfdti = Ax.im * Fx.re + Ay.im * Fy.re;
    const double dfdtr = Ax.im * Fx.im + Ay.im * Fy.im;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const double trq = tr + (q == 0 ? dtr : q == 1 ? -dtr : 0.0);
        const double kdq = kd + (q == 2 ? dkd : q == 3 ? -dkd : 0.0);
        const LaserTr& LTq = B.lt[q == 0 ? 1 : q == 1 ? 2 : 0];   // laser at trq
        const double df = q < 2 ? dfdtr * (trq - tr) : (kdq * kdq - kd * kd) / 2;
        double t0 = ti - df / dfdti;
        if (!(t0 > 0.0) || !isfinite(t0)) t0 = ti;
        Cx f1, f2, a1, b1;
        (void)spa_solve<true>(L, LTq, trq, B.F, S.omega, Ip, kdq, kz, f1, f2, a1, b1, t0);
        double qx, qy;
        spa_p(kdq, a1, b1, qx, qy);
        o[11 + 2 * q] = qx;
        o[12 + 2 * q] = qy;
    }
}

// SFA-SPA phase 2, for a sample whose saddle point was found: the exit
// position x0 = int_0^ti Im Ax(tr + i tau) (:350-351), the tunnelling integral
// G = int_0^ti g(tr + i tau) dtau with g = ((px+Ax)^2 + (py+Ay)^2 + kz^2)/2 + Ip,
// so exp(i S_tun) = exp(-G) (:279), by kSpaGlPoints-point Gauss-Legendre on
// [0, ti] (S.glq: nodes, weights); the prefactor k(ts).F(ts) (:271-278) and the
// finite-difference Jacobian (:280-292).
template <bool PRE>
__device__ inline void spa_amp(const SamplerP& S, const LaserP& L, const TargetP& T, const AdkBatch& B, double kd,
                               double kz, double w, SpaSlots o, double& rate, double& phase, double col[6]) {
    const double Ip = S.Ip, tr = B.tr;
    const double ti = o[0], px = o[1], py = o[2];
    const Cx Fx{o[3], o[4]}, Fy{o[5], o[6]}, Ax{o[7], o[8]}, Ay{o[9], o[10]};
    const Cx kdotF = Cx{px + Ax.re, Ax.im} * Fx + Cx{py + Ay.re, Ay.im} * Fy;
    double x0 = 0.0, y0 = 0.0;
    Cx G{0.0, 0.0};
    const double h = ti / 2;
#pragma unroll 1
    for (int i = 0; i < kSpaGlPoints; ++i) {
        Cx fx, fy, ax, ay;
        laser_field_pot_tau(L, B.lt[0], tr, h + h * S.glq[2 * i], fx, fy, ax, ay);
        const double wi = S.glq[2 * i + 1];
        x0 += wi * ax.im;
        y0 += wi * ay.im;
        const Cx kx{px + ax.re, ax.im}, ky{py + ay.re, ay.im};
        const Cx g = kx * kx + ky * ky;
        G = G + wi * Cx{g.re / 2 + kz * kz / 2 + Ip, g.im / 2};
    }
    x0 *= h;
    y0 *= h;
    G = h * G;
    double Axr, Ayr;
    laser_potential(L, tr, Axr, Ayr);
    // amp = sqrt(w) exp(-G) [c numer / (k.F)^((n+1)/2)] [sqrt(jac)]  (:296-309)
    double lmag = 0.5 * log(w) - G.re, arg = -G.im;
    if (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
        const double cpre = (S.prefix & ETRAJ_PREFIX_PRE) ? S.c_pre : B.c_cc;
        const double p = (S.nstar + 1) / 2;
        if (!PRE) {
            lmag += log(cpre * S.numer);
        } else {   // k_ts = (px + Ax(ts), py + Ay(ts), kz) (:250)
            const Cx N = pre_eval(S, B, Cx{px + Ax.re, Ax.im}, Cx{py + Ay.re, Ay.im}, Cx{kz, 0.0});
            lmag += log(cpre) + log(hypot(N.re, N.im));
            arg += atan2(N.im, N.re);
        }
        lmag -= p * log(hypot(kdotF.re, kdotF.im));
        arg -= p * atan2(kdotF.im, kdotF.re);
    }
    if (S.prefix & ETRAJ_PREFIX_JAC) {
        const double dtr = 0.01, dkd = 0.01;
        const double dpxdtr = o[11] - o[13], dpydtr = o[12] - o[14];
        const double dpxdkd = o[15] - o[17], dpydkd = o[16] - o[18];
        lmag += 0.5 * log(fabs(dpxdtr * dpydkd - dpxdkd * dpydtr) / (4 * dtr * dkd));
    }
    rate = exp(2 * lmag);
    phase = atan2(sin(arg), cos(arg));     // angle(amp) in (-pi, pi]
    col[0] = x0;
    col[1] = y0;
    col[2] = px + Axr;                     // :352-353
    col[3] = py + Ayr;
}

// The sample point (kd0, kz0) of sample j of batch b and its weight w (the
// step grid of :316-320, shared_methods.jl:4-12 Monte-Carlo, or the stratified
// Gaussian extension).
// x^1.5 as x sqrt(x), within two ulp of pow(x, 1.5)
__device__ inline double pow15(double x) { return x * sqrt(x); }

// x^e for the prefactor's exponent e = (n* + 1)/4 of a job: a square root for e = 1/2 (n* = 1,
// the hydrogen-like ground states), x for e = 1, else exp(e log x); within a few ulp of pow(x, e)
__device__ inline double pow_nstar(double x, double e) {
    if (e == 0.5) return sqrt(x);
    if (e == 1.0) return x;
    return exp(e * log(x));
}

__device__ inline void sample_point(const SamplerP& S, const AdkBatch& B, int b, int64_t j, double& kd0, double& kz0,
                                    double& w) {
    kz0 = 0.0;
    w = B.dkdt;
    if (S.sampling == ETRAJ_SAMPLING_STEP) {
        int64_t ikd = j / S.kz_num, ikz = j - ikd * S.kz_num;      // ikd outer, ikz inner (:318-319)
        kd0 = grid_point(-S.kd_max, S.kd_max, S.kd_num, (int)ikd);
        if (S.dim == 3) kz0 = grid_point(-S.kz_max, S.kz_max, S.kz_num, (int)ikz);
    } else {
        U64x4 o = philox((uint64_t)j, (uint64_t)b, 0, kTagK, S.seed, 0);
        if (S.sampling == ETRAJ_SAMPLING_MONTECARLO) {           // shared_methods.jl:4-12
            kd0 = (unit_co(o.v[0]) - 0.5) * 2 * S.mc_kd_max;
            if (S.dim == 3) {
                kz0 = (unit_co(o.v[1]) - 0.5) * 2 * S.mc_kz_max;
                if (S.mc_kd_max == 0) kd0 = 0.0;
                if (S.mc_kz_max == 0) kz0 = 0.0;
            }
        } else {                                                  // stratified Gaussian (extension)
            // Box-Muller; w is divided by the normal densities of kd0 (and kz0), i.e. multiplied
            // by sigma sqrt(2 pi) exp(z^2 / 2) with z = kd0 / sigma = r cos a (r sin a)
            const double r = sqrt(-2.0 * log(unit_oc(o.v[0])));
            double sa, ca;
            sincos(2 * kPi * unit_co(o.v[1]), &sa, &ca);         // a = 2 pi u
            const double zd = r * ca;
            kd0 = S.kd_sigma * zd;
            constexpr double kSqrt2Pi = 2.5066282746310002;
            if (S.dim == 3) {
                const double zz = r * sa;
                kz0 = S.kz_sigma * zz;
                w = w * ((S.kd_sigma * kSqrt2Pi) * (S.kz_sigma * kSqrt2Pi)) * exp(0.5 * (zd * zd + zz * zz));
            } else {
                w = w * (S.kd_sigma * kSqrt2Pi) * exp(0.5 * (zd * zd));
            }
        }
    }
}

// The columns of an SFA-SPA sample from its phase-2 results (false: discarded).
__device__ inline bool spa_columns(const SamplerP& S, const AdkBatch& B, double kz0, double rate, double phase,
                                   const double v[6], double* col) {
    if (isnan(rate) || rate < S.cutoff) return false;                // :332
    if (S.dim == 3) {
        col[0] = v[0]; col[1] = v[1]; col[2] = 0.0;
        col[3] = v[2]; col[4] = v[3]; col[5] = kz0; col[6] = B.tr; col[7] = rate;
        if (S.phase != ETRAJ_PHASE_CTMC) col[8] = phase;
    } else {
        col[0] = v[0]; col[1] = v[1];
        col[2] = v[2]; col[3] = v[3]; col[4] = B.tr; col[5] = rate;
        if (S.phase != ETRAJ_PHASE_CTMC) col[6] = phase;
    }
    return true;
}

// Evaluates sample j of batch b; returns false if it is discarded.
template <int M, bool PRE>
__device__ inline bool adk_sample(const SamplerP& S, const LaserP& L, const TargetP& T, const AdkBatch& B, int b,
                                  int64_t j, double* col) {
    double kd0, kz0, w;
    sample_point(S, B, b, j, kd0, kz0, w);
    const double Ip = S.Ip;
    double rate, phase = 0.0;
    double x0, y0, kx0, ky0;
    static_assert(M != ETRAJ_INIT_SPA, "SFA-SPA samples in two phases (spa_saddles, spa_amp)");
    if (M == ETRAJ_INIT_SPANE) {
        if (fabs(kd0) < 1e-4) return false;                           // SPANESampler.jl:323
        kx0 = kd0 * -B.sphi;
        ky0 = kd0 * B.cphi;
        const double k2 = kx0 * kx0 + ky0 * ky0 + kz0 * kz0;
        const double f2 = B.F * B.F - (kx0 * B.dFx + ky0 * B.dFy);   // F2eff :239
        const double r0 = B.F / 2 * (k2 + 2 * Ip) / f2;               // :240
        x0 = r0 * B.cphi;
        y0 = r0 * B.sphi;
        if (f2 <= B.F * B.F / (S.sampling == ETRAJ_SAMPLING_STEP ? 9.0 : 25.0)) return false;   // :332, :378
        double amp = sqrt(w) * exp(-pow15(kd0 * kd0 + kz0 * kz0 + 2 * Ip) / (3 * sqrt(f2)));
        double aim = 0.0;
        if (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
            const double cpre = (S.prefix & ETRAJ_PREFIX_PRE) ? S.c_pre : B.c_cc;
            const double den = pow_nstar((k2 + 2 * Ip) * f2, (S.nstar + 1) / 4);   // :262-269
            if (!PRE) {
                amp *= cpre * S.numer / den;
            } else {   // ti = sqrt((kappa^2+kd^2+kz^2)/F2eff), k_ts of :241-242
                const double ti = sqrt((2 * Ip + kd0 * kd0 + kz0 * kz0) / f2);
                const Cx N = pre_eval(S, B, Cx{kx0 + ti * ti / 2 * B.dFx, -ti * B.Fxr},
                                      Cx{ky0 + ti * ti / 2 * B.dFy, -ti * B.Fyr}, Cx{kz0, 0.0});
                const double f = amp * cpre / den;
                amp = f * N.re;
                aim = f * N.im;
            }
        }
        if (S.prefix & ETRAJ_PREFIX_JAC) {
            const double jf = sqrt(fabs(B.F + sqrt(kd0 * kd0 + kz0 * kz0) / (B.F * B.F) * B.cross));   // :270
            amp *= jf;
            aim *= jf;
        }
        rate = amp * amp + aim * aim;
        if (PRE) phase = atan2(aim, amp);
        if (isnan(rate) || rate / w < S.cutoff) return false;        // :337: rate/dkdt (Q7)
    } else if (M == ETRAJ_INIT_WFAT) {
        if (kd0 * kd0 + kz0 * kz0 < 1e-4 * 1e-4) return false;    // WFATSampler.jl:188
        // sum over channels of |G|^2 W(nxi, |m|, kd, kz) dkdt (WFATSampler.jl:157,:169),
        // W = prefactor (kd^2+kz^2)^|m| exp(-2 (kappa^2+kd^2+kz^2)^1.5 / 3F)
        const double k2 = 2 * Ip, q2 = kd0 * kd0 + kz0 * kz0;
        double poly = 0.0;
        for (int am = B.wM; am >= 0; --am) poly = poly * q2 + B.wB[am];
        rate = poly * exp(-2 * pow15(k2 + q2) * B.i3F) * w;
    } else {
        if (fabs(kd0) < 1e-4) return false;                           // ADKSampler.jl:312,:321
        const double kx = kd0 * -B.sphi, ky = kd0 * B.cphi;
        double amp = sqrt(w) * exp(-pow15(kd0 * kd0 + kz0 * kz0 + 2 * Ip) * B.i3F);  // :198
        double aim = 0.0;                                             // Im amp (prefactor beyond l = 0)
        if (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {    // :258-265
            double cpre = (S.prefix & ETRAJ_PREFIX_PRE) ? S.c_pre : B.c_cc;
            const double den = pow_nstar((kx * kx + ky * ky + kz0 * kz0 + 2 * Ip) * (B.F * B.F), (S.nstar + 1) / 4);
            if (!PRE) {
                amp *= cpre * S.numer / den;
            } else {   // k_ts at ts = tr + i ti, ti = sqrt(kappa^2+kd^2+kz^2)/F (:237-238)
                const double ti = sqrt(2 * Ip + kd0 * kd0 + kz0 * kz0) / B.F;
                const Cx N = pre_eval(S, B, Cx{kx, -ti * B.Fxr}, Cx{ky, -ti * B.Fyr}, Cx{kz0, 0.0});
                const double f = amp * cpre / den;
                amp = f * N.re;
                aim = f * N.im;
            }
        }
        if (S.prefix & ETRAJ_PREFIX_JAC) {                            // :266,:283
            amp *= B.sqF;
            aim *= B.sqF;
        }
        rate = amp * amp + aim * aim;
        if (PRE) phase = atan2(aim, amp);                    // angle(amp)
    }
    if (M == ETRAJ_INIT_ADK || M == ETRAJ_INIT_WFAT) {
        kx0 = kd0 * -B.sphi;                                          // :324-325
        ky0 = kd0 * B.cphi;
        const double r0 = (Ip + (kd0 * kd0 + kz0 * kz0) / 2) * B.iF;  // :326
        x0 = r0 * B.cphi;
        y0 = r0 * B.sphi;
    }
    // :332 (SPANE applied its own rate/dkdt cut above, Q7)
    if (M != ETRAJ_INIT_SPANE && (isnan(rate) || rate < S.cutoff)) return false;
    if (S.dim == 3) {
        col[0] = x0; col[1] = y0; col[2] = 0.0;
        col[3] = kx0; col[4] = ky0; col[5] = kz0; col[6] = B.tr; col[7] = rate;
        if (S.phase != ETRAJ_PHASE_CTMC) col[8] = phase;
    } else {
        col[0] = x0; col[1] = y0;
        col[2] = kx0; col[3] = ky0; col[4] = B.tr; col[5] = rate;
        if (S.phase != ETRAJ_PHASE_CTMC) col[6] = phase;
    }
    return true;
}

// One workgroup per batch; survivors are compacted in sample order with a
// wave ballot + LDS scan, so column order equals the reference's loop order.
// Workgroup g samples batch G.batch[g] of job G.job[g] (per-job parameters
// in sps / lps) into the columns from G.start[g]. At least three waves per SIMD:
// the SFA-SPA form then fits 168 VGPRs with two values spilled to scratch and
// samples the Ar 3D grid 6 % faster than with two waves (214 VGPRs, no spill).
template <int M, bool PRE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_adk_sample(
    const SamplerP* __restrict__ sps, const LaserP* __restrict__ lps, TargetP T, SegTab G, double* __restrict__ out,
    int32_t* __restrict__ counts, int64_t* __restrict__ cnt) {
    __shared__ int wave_tot[4];
    __shared__ int base_sh;
    const int g = blockIdx.x;
    const int job = G.job[g];
    const int b = G.batch[g];
    const SamplerP S = sps[job];
    const LaserP& L = lps[job];
    AdkBatch B = adk_batch_setup<M, PRE>(S, L, T, G.t0[g]);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double* bout = out + G.start[g] * S.rows;
    if (tid == 0) base_sh = 0;
    if (M == ETRAJ_INIT_SPA) {   // the laser at the three real times of this batch's complex-time evaluations
        __shared__ LaserTr spa_lt[3];
        if (tid < 3) spa_lt[tid] = laser_tr(L, tid == 0 ? B.tr : tid == 1 ? B.tr + 0.01 : B.tr - 0.01);
        B.lt = spa_lt;
    }
    if (M == ETRAJ_INIT_WFAT) {
        // B_|m| = sum_nxi sum_{m = +-|m|} |G_{nxi m}|^2 prefactor(nxi, |m|): one
        // channel per thread, then summed in a fixed order (deterministic)
        __shared__ double wfat_v[kWfatMaxChannels];
        __shared__ double wfat_B[kWfatMaxM + 1];
        const double kappa = sqrt(2 * S.Ip);
        if (S.wfat_form == ETRAJ_WFAT_INTDATA) {
            const int D0 = S.wfat_dims[0], D1 = S.wfat_dims[1], mM = (D1 - 1) / 2;
            for (int ch = tid; ch < D0 * D1; ch += blockDim.x) {
                const int nxi = ch / D1, m = ch % D1 - mM;
                wfat_v[ch] = wfat_channel_g2(S, kappa, B.beta, B.gamma, nxi, m) *
                             wfat_field_prefactor(kappa, B.F, S.nstar, nxi, abs(m));
            }
            __syncthreads();
            if (tid <= mM) {
                double acc = 0.0;
                for (int nxi = 0; nxi < D0; ++nxi) {
                    acc += wfat_v[nxi * D1 + mM + tid];
                    if (tid > 0) acc += wfat_v[nxi * D1 + mM - tid];
                }
                wfat_B[tid] = acc;
            }
            B.wM = mM;
        } else {
            if (tid == 0) wfat_B[0] = B.g2 * wfat_field_prefactor(kappa, B.F, S.nstar, 0, 0);
            B.wM = 0;
        }
        B.wB = wfat_B;
    } else if (PRE && (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC))) {
        __shared__ double pre_T[2 * kPreMaxMono];
        const int nmono = (S.pre_L + 1) * (S.pre_L + 2) * (S.pre_L + 3) / 6;
        for (int i = tid; i < nmono; i += blockDim.x) pre_monomial(S, B, i, pre_T + 2 * i);
        B.pT = pre_T;
    }
    __syncthreads();
    double col[9];
    for (int64_t j0 = 0; j0 < S.nk; j0 += 256) {
        int64_t j = j0 + tid;
        bool keep;
        if constexpr (M == ETRAJ_INIT_SPA) {
            __shared__ double spa_slots[kSpaSlots * 256];
            const SpaSlots o{spa_slots + tid};
            double kd0 = 0.0, kz0 = 0.0, w = 0.0;
            keep = (j < S.nk) && !B.skip;
            if (keep) {
                sample_point(S, B, b, j, kd0, kz0, w);
                keep = fabs(kd0) >= 1e-4;                            // SPASampler.jl:345
            }
            if (keep) {
                if (S.prefix & ETRAJ_PREFIX_JAC) spa_saddles<true>(S, L, B, kd0, kz0, o);
                else spa_saddles<false>(S, L, B, kd0, kz0, o);
            }
            __syncthreads();   // a fence between the phases (each thread reads back only its own slots)
            if (keep && o[0] != 0.0) {
                double rate, phase, v[6];
                spa_amp<PRE>(S, L, T, B, kd0, kz0, w, o, rate, phase, v);
                keep = spa_columns(S, B, kz0, rate, phase, v, col);
            } else {
                keep = false;
            }
        } else {
            keep = (j < S.nk) && !B.skip && adk_sample<M, PRE>(S, L, T, B, b, j, col);
        }
        uint64_t m = __ballot(keep);
        int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (lane == 0) wave_tot[wv] = __popcll(m);
        __syncthreads();
        int off = base_sh;
        for (int w = 0; w < wv; ++w) off += wave_tot[w];
        int tot = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        if (keep) {
            double* c = bout + (int64_t)(off + pre) * S.rows;
            for (int r = 0; r < S.rows; ++r) c[r] = col[r];
        }
        __syncthreads();
        if (tid == 0) base_sh += tot;
        __syncthreads();
    }
    if (tid == 0) {
        counts[g] = base_sh;
        if (cnt && base_sh)
            atomicAdd((unsigned long long*)(cnt + ((int64_t)job * kAccSlots + acc_slot()) * CNT_N + CNT_LAUNCHED),
                      (unsigned long long)base_sh);
    }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
template <int M>
static void launch_sample_m(hipStream_t st, bool pre, const SamplerP* sps, const LaserP* lps, const TargetP& T,
                            const SegTab& G, int nb, double* out, int32_t* counts, int64_t* cnt) {
    if (pre)
        hipLaunchKernelGGL((k_adk_sample<M, true>), dim3(nb), dim3(256), 0, st, sps, lps, T, G, out, counts, cnt);
    else
        hipLaunchKernelGGL((k_adk_sample<M, false>), dim3(nb), dim3(256), 0, st, sps, lps, T, G, out, counts, cnt);
}

hipError_t launch_adk_sample(hipStream_t st, int method, bool pre, const SamplerP* sps, const LaserP* lps,
                             const TargetP& T, const SegTab& G, int nb, double* out, int32_t* counts, int64_t* cnt) {
    if (nb <= 0) return hipSuccess;
    switch (method) {
        case ETRAJ_INIT_SPA: launch_sample_m<ETRAJ_INIT_SPA>(st, pre, sps, lps, T, G, nb, out, counts, cnt); break;
        case ETRAJ_INIT_SPANE: launch_sample_m<ETRAJ_INIT_SPANE>(st, pre, sps, lps, T, G, nb, out, counts, cnt); break;
        case ETRAJ_INIT_WFAT:
            hipLaunchKernelGGL((k_adk_sample<ETRAJ_INIT_WFAT, false>), dim3(nb), dim3(256), 0, st, sps, lps, T, G, out,
                               counts, cnt);
            break;
        default: launch_sample_m<ETRAJ_INIT_ADK>(st, pre, sps, lps, T, G, nb, out, counts, cnt);
    }
    return hipGetLastError();
}

}  // namespace etraj
