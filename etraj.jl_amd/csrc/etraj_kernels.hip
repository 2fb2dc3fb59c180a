// etraj_kernels.hip -- gfx950 kernels of the trajectory engine.
//
// Pipeline (one call of etraj_run_job):
//   k_adk_sample      one workgroup per t0 batch: ADK tunnel exit, weight,
//                     cutoff and order-preserving compaction into the init
//                     buffer (Julia column layout, one column per electron).
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
// reference's initial guess (:240); accepted if |f| <= 1e-8 (NLsolve's ftol)
// and ti > 0 (:241-246), else 0.
__device__ inline double spa_solve(const LaserP& L, const LaserTr& LT, double tr, double Ftr, double omega, double Ip,
                                   double kd, double kz) {
    const double k2 = kd * kd + kz * kz;
    double ti = asinh(omega / Ftr * sqrt(k2 + 2 * Ip)) / omega;
    Cx Fx, Fy, Ax, Ay;
#pragma unroll 1
    for (int it = 0; it < 100; ++it) {
        laser_field_pot_tau(L, LT, tr, ti, Fx, Fy, Ax, Ay);
        const double f = (k2 - (Ax.im * Ax.im + Ay.im * Ay.im)) / 2 + Ip;
        const double step = f / (Ax.im * Fx.re + Ay.im * Fy.re);
        if (!isfinite(step)) return 0.0;
        ti -= step;
        if (fabs(step) <= 1e-15 * fabs(ti)) break;
    }
    laser_field_pot_tau(L, LT, tr, ti, Fx, Fy, Ax, Ay);
    const double f = (k2 - (Ax.im * Ax.im + Ay.im * Ay.im)) / 2 + Ip;
    return (fabs(f) <= 1e-8 && ti > 0) ? ti : 0.0;
}

// px_/py_ (SPASampler.jl:231-232)
__device__ inline void spa_p(double kd, Cx Ax, Cx Ay, double& px, double& py) {
    const double rho = sqrt(Ax.im * Ax.im + Ay.im * Ay.im);
    px = kd * Ay.im / rho - Ax.re;
    py = -kd * Ax.im / rho - Ay.re;
}

// One SPA sample: saddle point, exit position x0 = int_0^ti Im Ax(tr + i tau)
// (:350-351), the tunnelling integral G = int_0^ti g(tr + i tau) dtau with
// g = ((px+Ax)^2 + (py+Ay)^2 + kz^2)/2 + Ip, so exp(i S_tun) = exp(-G) (:279),
// by kSpaGlPoints-point Gauss-Legendre on [0, ti] (S.glq: nodes, weights); the
// prefactor k(ts).F(ts) (:271-278) and the finite-difference Jacobian
// (:280-292). Returns false if the solve fails.
template <bool PRE>
__device__ inline bool spa_eval(const SamplerP& S, const LaserP& L, const TargetP& T, const AdkBatch& B, double kd,
                                double kz, double w, double& rate, double& phase, double col[6]) {
    const double Ip = S.Ip, tr = B.tr;
    const double ti = spa_solve(L, B.lt[0], tr, B.F, S.omega, Ip, kd, kz);
    if (ti == 0.0) return false;
    Cx Fx, Fy, Ax, Ay;
    laser_field_pot_tau(L, B.lt[0], tr, ti, Fx, Fy, Ax, Ay);
    double px, py;
    spa_p(kd, Ax, Ay, px, py);
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
        // saddle points at tr +- dtr and kd +- dkd; a failed solve gives
        // ti = 0 and a NaN momentum, as in the reference
        const double dtr = 0.01, dkd = 0.01;
        double pq[4][2];
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const double trq = tr + (q == 0 ? dtr : q == 1 ? -dtr : 0.0);
            const double kdq = kd + (q == 2 ? dkd : q == 3 ? -dkd : 0.0);
            const LaserTr& LTq = B.lt[q == 0 ? 1 : q == 1 ? 2 : 0];   // laser at trq
            const double tq = spa_solve(L, LTq, trq, B.F, S.omega, Ip, kdq, kz);
            Cx f1, f2, a1, b1;
            laser_field_pot_tau(L, LTq, trq, tq, f1, f2, a1, b1);
            spa_p(kdq, a1, b1, pq[q][0], pq[q][1]);
        }
        const double dpxdtr = pq[0][0] - pq[1][0], dpydtr = pq[0][1] - pq[1][1];
        const double dpxdkd = pq[2][0] - pq[3][0], dpydkd = pq[2][1] - pq[3][1];
        lmag += 0.5 * log(fabs(dpxdtr * dpydkd - dpxdkd * dpydtr) / (4 * dtr * dkd));
    }
    rate = exp(2 * lmag);
    phase = atan2(sin(arg), cos(arg));     // angle(amp) in (-pi, pi]
    col[0] = x0;
    col[1] = y0;
    col[2] = px + Axr;                     // :352-353
    col[3] = py + Ayr;
    return true;
}

// Evaluates sample j of batch b; returns false if it is discarded.
template <int M, bool PRE>
__device__ inline bool adk_sample(const SamplerP& S, const LaserP& L, const TargetP& T, const AdkBatch& B, int b,
                                  int64_t j, double* col) {
    double kd0, kz0 = 0.0, w = B.dkdt;
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
            double r = sqrt(-2.0 * log(unit_oc(o.v[0])));
            double a = 2 * kPi * unit_co(o.v[1]);
            double sa, ca;
            sincos(a, &sa, &ca);
            kd0 = S.kd_sigma * (r * ca);
            w = w / (exp(-0.5 * (kd0 / S.kd_sigma) * (kd0 / S.kd_sigma)) / (S.kd_sigma * sqrt(2 * kPi)));
            if (S.dim == 3) {
                kz0 = S.kz_sigma * (r * sa);
                w = w / (exp(-0.5 * (kz0 / S.kz_sigma) * (kz0 / S.kz_sigma)) / (S.kz_sigma * sqrt(2 * kPi)));
            }
        }
    }
    const double Ip = S.Ip;
    double rate, phase = 0.0;
    double x0, y0, kx0, ky0;
    if (M == ETRAJ_INIT_SPA) {
        if (fabs(kd0) < 1e-4) return false;                           // SPASampler.jl:345
        double v[6];
        if (!spa_eval<PRE>(S, L, T, B, kd0, kz0, w, rate, phase, v)) return false;
        x0 = v[0]; y0 = v[1]; kx0 = v[2]; ky0 = v[3];
    } else if (M == ETRAJ_INIT_SPANE) {
        if (fabs(kd0) < 1e-4) return false;                           // SPANESampler.jl:323
        kx0 = kd0 * -B.sphi;
        ky0 = kd0 * B.cphi;
        const double k2 = kx0 * kx0 + ky0 * ky0 + kz0 * kz0;
        const double f2 = B.F * B.F - (kx0 * B.dFx + ky0 * B.dFy);   // F2eff :239
        const double r0 = B.F / 2 * (k2 + 2 * Ip) / f2;               // :240
        x0 = r0 * B.cphi;
        y0 = r0 * B.sphi;
        if (f2 <= B.F * B.F / (S.sampling == ETRAJ_SAMPLING_STEP ? 9.0 : 25.0)) return false;   // :332, :378
        double amp = sqrt(w) * exp(-pow(kd0 * kd0 + kz0 * kz0 + 2 * Ip, 1.5) / (3 * sqrt(f2)));
        double aim = 0.0;
        if (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
            const double cpre = (S.prefix & ETRAJ_PREFIX_PRE) ? S.c_pre : B.c_cc;
            const double den = pow((k2 + 2 * Ip) * f2, (S.nstar + 1) / 4);   // :262-269
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
        rate = poly * exp(-2 * pow(k2 + q2, 1.5) / (3 * B.F)) * w;
    } else {
        if (fabs(kd0) < 1e-4) return false;                           // ADKSampler.jl:312,:321
        const double kx = kd0 * -B.sphi, ky = kd0 * B.cphi;
        double amp = sqrt(w) * exp(-pow(kd0 * kd0 + kz0 * kz0 + 2 * Ip, 1.5) / (3 * B.F));  // :198
        double aim = 0.0;                                             // Im amp (prefactor beyond l = 0)
        if (S.prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {    // :258-265
            double cpre = (S.prefix & ETRAJ_PREFIX_PRE) ? S.c_pre : B.c_cc;
            const double den = pow((kx * kx + ky * ky + kz0 * kz0 + 2 * Ip) * (B.F * B.F), (S.nstar + 1) / 4);
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
            amp *= sqrt(B.F);
            aim *= sqrt(B.F);
        }
        rate = amp * amp + aim * aim;
        if (PRE) phase = atan2(aim, amp);                    // angle(amp)
    }
    if (M == ETRAJ_INIT_ADK || M == ETRAJ_INIT_WFAT) {
        kx0 = kd0 * -B.sphi;                                          // :324-325
        ky0 = kd0 * B.cphi;
        const double r0 = (Ip + (kd0 * kd0 + kz0 * kz0) / 2) / B.F;  // :326
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
// in sps / lps) into the columns from G.start[g]. At least two waves per SIMD:
// the SFA-SPA form would otherwise take 256 VGPRs + 130 AGPRs (one wave); with
// ~140 values spilled to scratch it samples the Ar 3D grid 7 % faster.
template <int M, bool PRE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_adk_sample(
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
        bool keep = (j < S.nk) && !B.skip && adk_sample<M, PRE>(S, L, T, B, b, j, col);
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
        if (cnt && base_sh) atomicAdd((unsigned long long*)(cnt + (int64_t)job * CNT_N + CNT_LAUNCHED),
                                      (unsigned long long)base_sh);
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
        if (O.pmdq) {   // fixed point: exact integer sums (wave-uniform branch)
            bool ok1 = true, ok2 = true, ok3, ok4;
            if (inb) {
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
                fix_atomic_add(O.scalq + 0, ip);
                fix_atomic_add(O.scalq + 2, iu);
            }
        } else {
            if (inb) {
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
                if (ip != 0.0) unsafeAtomicAdd(O.scal + 0, ip);
                if (iu != 0.0) unsafeAtomicAdd(O.scal + 1, iu);
            }
        }
        if (O.hits && inb) atomicAdd((unsigned long long*)(O.hits + lin), 1ull);
        int64_t nf = wave_sum_i64((active && failed) ? 1 : 0);
        int64_t nb = wave_sum_i64(bound ? 1 : 0);
        if (lane_id() == 0) {
            if (nf) atomicAdd((unsigned long long*)(O.cnt + CNT_FAILED), (unsigned long long)nf);
            if (nb) atomicAdd((unsigned long long*)(O.cnt + CNT_BOUND), (unsigned long long)nb);
            if (nr) atomicAdd((unsigned long long*)(O.cnt + CNT_FIXRANGE), (unsigned long long)nr);
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
                hl_rk4_sums<float, DIM, PHASE, SOFT, UNITZ>(Zf, af, r, p, hf, Fpx, Fpy, f.x, f.y, f.z, f.w, Sr, Sp, SA,
                                                            SQ, r2);
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
    const double h6 = h / 6;
    const HlPhase<double> pc = hl_phase<double, PHASE, true>(h, P.T.Ip);
    double Fpx = F0.x, Fpy = F0.y;
    auto step = [&](const double2 Fm, const double2 F1) {
        double Sr[DIM], Sp[DIM], r2;
        if constexpr (HL) {
            double SA, SQ;
            hl_rk4_sums<double, DIM, PHASE, true, UNITZ>(P.T.Zd, P.T.sc, r, p, h, Fpx, Fpy, Fm.x, Fm.y, F1.x, F1.y, Sr,
                                                         Sp, SA, SQ, r2);
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
        Oj.scal = O.scal + 2 * job;
        Oj.cnt = O.cnt + (int64_t)job * CNT_N;
        if (O.pmdq) {
            Oj.pmdq = O.pmdq + 2 * job * O.job_stride;
            Oj.scalq = O.scalq + 4 * job;
        }
        if (O.hits) Oj.hits = O.hits + job * O.npix;
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
            atomicAdd((unsigned long long*)(Oj.cnt + CNT_STEPS), (unsigned long long)(n * nact));
            atomicAdd((unsigned long long*)(Oj.cnt + CNT_ATTEMPTS), (unsigned long long)(n * nact));
            if (!MIXED) atomicAdd((unsigned long long*)(Oj.cnt + CNT_STEPS64), (unsigned long long)(n * nact));
            if (nclose) atomicAdd((unsigned long long*)(Oj.cnt + CNT_RESTEP), (unsigned long long)nclose);
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

// In-place, order-preserving compaction of the flagged columns of each
// segment (one workgroup per segment): count_out[s] receives their number and
// orig[seg_start[s] + k] the original column (or its out_index entry) of the
// k-th. Every column of a chunk is read before any is written, and writes
// never pass the read position, so the compaction can overwrite in place.
__global__ __launch_bounds__(256) void k_compact_flagged(double* __restrict__ init, int rows,
                                                         const int64_t* __restrict__ seg_start,
                                                         const int32_t* __restrict__ seg_count,
                                                         const uint8_t* __restrict__ flags,
                                                         const int64_t* __restrict__ out_index,
                                                         int32_t* __restrict__ count_out,
                                                         int64_t* __restrict__ orig) {
    __shared__ int wave_tot[4];
    __shared__ int base_sh;
    const int64_t seg = blockIdx.x;
    const int64_t base = seg_start[seg];
    const int32_t cnt = seg_count[seg];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) base_sh = 0;
    __syncthreads();
    double col[9];
    for (int j0 = 0; j0 < cnt; j0 += 256) {
        const int j = j0 + tid;
        const bool keep = j < cnt && flags[base + j] != 0;
        if (keep)
            for (int r = 0; r < rows; ++r) col[r] = init[(base + j) * rows + r];
        const uint64_t m = __ballot(keep);
        const int pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (lane == 0) wave_tot[wv] = __popcll(m);
        __syncthreads();
        int off = base_sh;
        for (int w = 0; w < wv; ++w) off += wave_tot[w];
        const int tot = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
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

template <int DIM, int PHASE, bool MIXED, int OUT>
static hipError_t launch_rk4_t(hipStream_t st, const IntegP& P, const OutP& O, const double* init,
                               const int64_t* seg_start, const int32_t* seg_count, int64_t n_seg, int cps,
                               const FieldTab& FT, uint8_t* flags, const int64_t* out_index) {
    const unsigned blocks = (unsigned)((n_seg * (int64_t)cps + 3) / 4);
    if (P.T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        hipLaunchKernelGGL((k_rk4<DIM, PHASE, MIXED, OUT, ETRAJ_TARGET_HYDROGEN_LIKE>), dim3(blocks), dim3(256), 0, st,
                           P, O, init, seg_start, seg_count, n_seg, cps, FT, flags, out_index);
    else
        hipLaunchKernelGGL((k_rk4<DIM, PHASE, MIXED, OUT, ETRAJ_TARGET_SAE>), dim3(blocks), dim3(256), 0, st, P, O,
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
    hipLaunchKernelGGL(k_compact_flagged, dim3((unsigned)n_seg), dim3(256), 0, st, init, rows, seg_start, seg_count,
                       flags, out_index, count_out, orig);
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

hipError_t launch_finalize_abs2(hipStream_t st, const double* amp, int64_t n, double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize_abs2, dim3(grid_for(n, 256)), dim3(256), 0, st, amp, n, out);
    return hipGetLastError();
}

}  // namespace etraj
