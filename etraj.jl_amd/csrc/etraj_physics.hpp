// etraj_physics.hpp -- host+device physics of the trajectory engine.
//
// Everything here is `__host__ __device__` so the same arithmetic runs in the
// gfx950 kernels and, for development checks, on the host. Field and target
// parameters are resolved once on the host from the C-ABI descriptors into
// the flat `LaserP` / `TargetP` structs below (passed by value as kernel
// arguments, so they live in SGPRs / the kernarg segment).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/etraj.h"

namespace etraj {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon et al., SC'11). Counter = (index, batch, 0, tag),
// key = (seed, 0).
// ---------------------------------------------------------------------------
struct U64x4 { uint64_t v[4]; };

__host__ __device__ inline void mulhilo(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    hi = __umul64hi(a, b);
    lo = a * b;
#else
    unsigned __int128 p = (unsigned __int128)a * b;
    hi = (uint64_t)(p >> 64);
    lo = (uint64_t)p;
#endif
}

__host__ __device__ inline U64x4 philox(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t h0, l0, h1, l1;
        mulhilo(0xD2E7470EE14C6C93ULL, c0, h0, l0);
        mulhilo(0xCA5A826395121157ULL, c2, h1, l1);
        uint64_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0; c1 = l1; c2 = n2; c3 = l0;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    U64x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

constexpr uint64_t kTagT0 = 1, kTagK = 2;

__host__ __device__ inline double unit_co(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }          // [0,1)
__host__ __device__ inline double unit_oc(uint64_t x) { return ((double)(x >> 11) + 1.0) * 0x1.0p-53; }  // (0,1]

// range(lo, hi; length=n)[i] (0-based): (lo*(n-1-i) + hi*i)/(n-1)
__host__ __device__ inline double grid_point(double lo, double hi, int n, int i) {
    if (n <= 1) return lo;
    return (lo * (double)(n - 1 - i) + hi * (double)i) / (double)(n - 1);
}

// ---------------------------------------------------------------------------
// Lasers (src/Lasers/*Laser.jl). F = -dA/dt in closed form; the cos^n
// envelopes carry the tanh cut-off window of Cos2Laser.jl:93.
// ---------------------------------------------------------------------------
struct MonoP {
    int kind;        // ETRAJ_LASER_*
    int azi_nz;      // azimuth != 0 selects the rotated branch (Cos2Laser.jl:99,127)
    double F0, A0, w, N, cep, tshift, ellip, cazi, sazi;
    double k;        // 200/Period
    double halfw;    // N*pi/w
    double coefN;    // 1/N (Cos2) or 2/N (Cos4)
    double sigma2;   // Gaussian: SpreadDuration^2
    double ihalfN;   // 1/(2N): the envelope phase w t/(2N) as a product (stage fields of Tsit5)
};

struct LaserP {
    int n;
    double delay;
    MonoP c[2];
};

inline int make_mono(const etraj_mono_laser& d, MonoP& m) {
    if (d.kind != ETRAJ_LASER_COS2 && d.kind != ETRAJ_LASER_COS4 && d.kind != ETRAJ_LASER_GAUSSIAN) return -1;
    if (!(d.wave_len > 0) || !(d.cyc_num > 0) || !(d.peak_int >= 0)) return -1;
    m.kind = d.kind;
    m.w = 45.563352525 / d.wave_len;                                   // AngFreq (Cos2Laser.jl:81)
    double period = 2 * kPi / m.w;                                     // Period  (:82)
    m.F0 = sqrt(d.peak_int / (1.0 + d.ellip * d.ellip) / 3.50944521e16); // LaserF0 (:85)
    m.A0 = m.F0 / m.w;                                                 // LaserA0 (:86)
    m.N = d.cyc_num;
    m.cep = d.cep;
    m.tshift = d.t_shift;
    m.ellip = d.ellip;
    m.azi_nz = d.azi != 0;
    m.cazi = cos(d.azi);
    m.sazi = sin(d.azi);
    m.k = 200 / period;
    m.halfw = m.N * kPi / m.w;
    m.coefN = (d.kind == ETRAJ_LASER_COS2 ? 1.0 : 2.0) / m.N;
    m.ihalfN = 1.0 / (2 * m.N);
    double sd = d.cyc_num * period;                                    // SpreadDuration (GaussianLaser.jl:120)
    m.sigma2 = sd * sd;
    return 0;
}

inline int make_laser(const etraj_laser_desc& d, LaserP& L) {
    if (d.n_colors != 1 && d.n_colors != 2) return -1;
    L.n = d.n_colors;
    L.delay = d.delay21;
    for (int i = 0; i < 2; ++i) L.c[i] = MonoP{};
    for (int i = 0; i < L.n; ++i)
        if (make_mono(d.color[i], L.c[i])) return -1;
    return 0;
}

__host__ __device__ inline double cos_window(const MonoP& m, double t) {
    return (0.5 + 0.5 * tanh(m.k * (t + m.halfw))) * (0.5 + 0.5 * tanh(-m.k * (t - m.halfw)));
}

// UnitEnvelope (Cos2Laser.jl:89-95, Cos4Laser.jl:89-95, GaussianLaser.jl:138-144)
__host__ __device__ inline double mono_envelope(const MonoP& m, double t) {
    t -= m.tshift;
    if (m.kind == ETRAJ_LASER_GAUSSIAN) return exp(-t * t / 2 / m.sigma2);
    double c = cos(m.w * t / (2 * m.N));
    double cn = (m.kind == ETRAJ_LASER_COS2) ? c * c : (c * c) * (c * c);
    return cn * cos_window(m, t);
}

// LaserFx/LaserFy (Cos2Laser.jl:125-152, Cos4Laser.jl:125-152, GaussianLaser.jl:176-203)
__host__ __device__ inline void mono_field(const MonoP& m, double t, double& Fx, double& Fy) {
    t -= m.tshift;
    double a, b, g;
    double sw, cw;
    sincos(m.w * t + m.cep, &sw, &cw);
    if (m.kind == ETRAJ_LASER_GAUSSIAN) {
        g = m.A0 * exp(-t * t / 2 / m.sigma2);
        a = t / m.sigma2 * cw + m.w * sw;
        b = t / m.sigma2 * sw - m.w * cw;
    } else {
        double sh, ch;
        sincos(m.w * t / (2 * m.N), &sh, &ch);
        double env = (m.kind == ETRAJ_LASER_COS2) ? ch : ch * ch * ch;
        g = m.F0 * env * cos_window(m, t);
        a = ch * sw + m.coefN * sh * cw;
        b = -ch * cw + m.coefN * sh * sw;
    }
    if (!m.azi_nz) {
        Fx = g * a;
        Fy = g * b * m.ellip;
    } else {
        Fx = g * (a * m.cazi + b * m.ellip * m.sazi);
        Fy = g * (a * -m.sazi + b * m.ellip * m.cazi);
    }
}

// LaserAx/LaserAy (Cos2Laser.jl:97-124 and siblings)
__host__ __device__ inline void mono_potential(const MonoP& m, double t, double& Ax, double& Ay) {
    double g = m.A0 * mono_envelope(m, t);
    t -= m.tshift;
    double sw, cw;
    sincos(m.w * t + m.cep, &sw, &cw);
    if (!m.azi_nz) {
        Ax = g * cw;
        Ay = g * sw * m.ellip;
    } else {
        Ax = g * (cw * m.cazi + sw * m.ellip * m.sazi);
        Ay = g * (cw * -m.sazi + sw * m.ellip * m.cazi);
    }
}

// BichromaticLaser.jl:59-85: F1(t) + F2(t - delay)
__host__ __device__ inline void laser_field(const LaserP& L, double t, double& Fx, double& Fy) {
    mono_field(L.c[0], t, Fx, Fy);
    if (L.n == 2) {
        double fx, fy;
        mono_field(L.c[1], t - L.delay, fx, fy);
        Fx += fx;
        Fy += fy;
    }
}

// ---------------------------------------------------------------------------
// The laser at several times t + d near one base time t (the Tsit5 stage
// times t + c_j dt). The carrier and envelope phases at t + d follow from
// their sin/cos at t by angle addition with the sin/cos of the increments
// (Taylor series while |increment| <= 1/2, the full sin/cos beyond), and each
// half of the cut-off window, 1/2 + tanh(x)/2, is the logistic function
// 1/(1 + exp(-2x)) of the same x (zero below x = -19.1, where the tanh form
// is exactly zero in double precision). Same field as mono_field up to
// rounding, at about a third of the instructions.
// ---------------------------------------------------------------------------
struct MonoBase {
    double tp;        // t - t_shift
    double sw, cw;    // sin/cos(w tp + cep)
    double sh, ch;    // sin/cos(w tp / 2N) (cos^n envelopes)
};
struct LaserBase {
    MonoBase c[2];
};

// sin and cos of |x| <= sincos_taylor_bound(KS) from their Taylor series through
// x^(2 KS + 1) and x^(2 KS + 2): KS = 7 for |x| <= 1/2, 5 for |x| <= 1/4, 4 for
// |x| <= 1/8, 3 for |x| <= 1/20, 2 for |x| <= 1/100 (in each case the first omitted
// term is below 1e-17, and below an ulp of the value).
constexpr double sincos_taylor_bound(int KS) {
    return KS >= 7 ? 0.5 : KS >= 5 ? 0.25 : KS == 4 ? 0.125 : KS == 3 ? 0.05 : 0.01;
}
constexpr double inv_factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;   // exact in double up to 18!
    return 1.0 / f;
}
template <int KS>
__host__ __device__ inline void sincos_taylor(double x, double& s, double& c) {
    const double x2 = x * x;
    double ps = ((KS & 1) ? -1.0 : 1.0) * inv_factorial(2 * KS + 1);
#pragma unroll
    for (int k = KS - 1; k >= 1; --k) ps = fma(ps, x2, ((k & 1) ? -1.0 : 1.0) * inv_factorial(2 * k + 1));
    s = fma(ps * x2, x, x);
    double pc = (((KS + 1) & 1) ? -1.0 : 1.0) * inv_factorial(2 * KS + 2);
#pragma unroll
    for (int k = KS; k >= 1; --k) pc = fma(pc, x2, ((k & 1) ? -1.0 : 1.0) * inv_factorial(2 * k));
    c = fma(pc, x2, 1.0);
}

// 1/x for x >= 1: on the device the hardware reciprocal refined by two Newton
// steps (within an ulp of the quotient), on the host the division.
__host__ __device__ inline double recip_ge1(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}

// 1/2 + tanh(x)/2 as 1/(1 + exp(-2x)); beyond |x| = 19.1 the result is exactly 0 or 1 in double
// (exp(-38.2) < 2^-54), so the exp is skipped there. (On the device the lanes near a pulse edge
// evaluate it while the rest of their wavefront waits: the reciprocal takes 5 instructions where
// the division takes 10.)
__host__ __device__ inline double half_tanh_step(double x) {
    return x < -19.1 ? 0.0 : (x > 19.1 ? 1.0 : recip_ge1(1.0 + exp(-2.0 * x)));
}

__host__ __device__ inline MonoBase mono_base(const MonoP& m, double t) {
    MonoBase b;
    b.tp = t - m.tshift;
    sincos(m.w * b.tp + m.cep, &b.sw, &b.cw);
    b.sh = 0.0;
    b.ch = 1.0;
    if (m.kind != ETRAJ_LASER_GAUSSIAN) sincos(m.w * b.tp * m.ihalfN, &b.sh, &b.ch);
    return b;
}

// SIMPLE: the caller guarantees a cos^2 / cos^4 pulse with azimuth 0 (simple_laser), so that the
// Gaussian and rotated forms, and their parameters, drop out of the code. at (may be null): the
// carrier and envelope sin/cos at t + d (at->tp = 1), or at->tp = 2 (t + d past the window) or 0
// (before it) where the window skipped them. KX: the Taylor order (sincos_taylor) of the carrier
// increment, for callers whose increments are known to be small (a Tsit5 stage at c dt, c < 1,
// whose step passes the |w dt| <= 1/2 test of the last stage); beyond its range the full sin/cos
// is taken. SIMPLE (N >= 2 cycles): the envelope increment x/(2N) is at most 1/8 where the
// carrier's is at most 1/2 and takes the order-4 series.
template <bool SIMPLE = false, int KX = 7>
__host__ __device__ inline void mono_field_at(const MonoP& m, const MonoBase& b, double d, double& Fx, double& Fy,
                                              MonoBase* at = nullptr) {
    const double t = b.tp + d;
    const double x = m.w * d;
    const bool gauss = !SIMPLE && m.kind == ETRAJ_LASER_GAUSSIAN;
    const double win = gauss ? 1.0 : half_tanh_step(m.k * (t + m.halfw)) * half_tanh_step(-m.k * (t - m.halfw));
    if (win == 0.0) {   // outside the pulse's window the field is exactly zero: no trigonometry
        Fx = 0.0;
        Fy = 0.0;
        if (at) at->tp = t > 0.0 ? 2.0 : 0.0;
        return;
    }
    const double xh = gauss ? 0.0 : x * m.ihalfN;
    constexpr bool env_short = SIMPLE;   // N >= 2 (simple_laser)
    double sw, cw, sh = 0.0, ch = 1.0;
    if (fabs(x) <= sincos_taylor_bound(KX) && fabs(xh) <= (env_short ? 0.125 : 0.5)) {
        double sd, cd;
        sincos_taylor<KX>(x, sd, cd);
        sw = b.sw * cd + b.cw * sd;
        cw = b.cw * cd - b.sw * sd;
        if (!gauss) {
            if (env_short) sincos_taylor<4>(xh, sd, cd);
            else sincos_taylor<7>(xh, sd, cd);
            sh = b.sh * cd + b.ch * sd;
            ch = b.ch * cd - b.sh * sd;
        }
    } else {
        sincos(m.w * t + m.cep, &sw, &cw);
        if (!gauss) sincos(m.w * t * m.ihalfN, &sh, &ch);
    }
    if (at) {
        at->tp = 1.0;
        at->sw = sw;
        at->cw = cw;
        at->sh = sh;
        at->ch = ch;
    }
    double a, bb, g;
    if (gauss) {
        g = m.A0 * exp(-t * t / 2 / m.sigma2);
        a = t / m.sigma2 * cw + m.w * sw;
        bb = t / m.sigma2 * sw - m.w * cw;
    } else {
        const double env = (m.kind == ETRAJ_LASER_COS2) ? ch : ch * ch * ch;
        g = m.F0 * env * win;
        a = ch * sw + m.coefN * sh * cw;
        bb = -ch * cw + m.coefN * sh * sw;
    }
    if (SIMPLE || !m.azi_nz) {
        Fx = g * a;
        Fy = g * bb * m.ellip;
    } else {
        Fx = g * (a * m.cazi + bb * m.ellip * m.sazi);
        Fy = g * (a * -m.sazi + bb * m.ellip * m.cazi);
    }
}

// The simple laser's field at b.tp + d from the phases `a6` at the later time b.tp + d + back
// (a6.tp == 1, as mono_field_at<true> gives them), turned back by w back and w back/(2N) with
// the order-KB and order-KH Taylor series (|w back| <= sincos_taylor_bound(KB), and its
// envelope part within KH's bound); otherwise mono_field_at<true> from the base b.
template <int KB, int KH>
__host__ __device__ inline void simple_field_back(const MonoP& m, const MonoBase& b, double d, const MonoBase& a6,
                                                  double back, double& Fx, double& Fy) {
    const double xb = m.w * back;
    if (!(a6.tp == 1.0 && fabs(xb) <= sincos_taylor_bound(KB))) {
        mono_field_at<true>(m, b, d, Fx, Fy);
        return;
    }
    const double t = b.tp + d;
    const double win = half_tanh_step(m.k * (t + m.halfw)) * half_tanh_step(-m.k * (t - m.halfw));
    if (win == 0.0) {
        Fx = 0.0;
        Fy = 0.0;
        return;
    }
    double sd, cd;
    sincos_taylor<KB>(xb, sd, cd);
    const double sw = a6.sw * cd - a6.cw * sd, cw = a6.cw * cd + a6.sw * sd;
    sincos_taylor<KH>(xb * m.ihalfN, sd, cd);
    const double sh = a6.sh * cd - a6.ch * sd, ch = a6.ch * cd + a6.sh * sd;
    const double env = (m.kind == ETRAJ_LASER_COS2) ? ch : ch * ch * ch;
    const double g = m.F0 * env * win;
    Fx = g * (ch * sw + m.coefN * sh * cw);
    Fy = g * (-ch * cw + m.coefN * sh * sw) * m.ellip;
}

// One colour, cos^2 or cos^4 envelope of at least two cycles, azimuth 0: the laser the SIMPLE
// forms below take.
__host__ __device__ inline bool simple_laser(const LaserP& L) {
    return L.n == 1 && L.c[0].kind != ETRAJ_LASER_GAUSSIAN && !L.c[0].azi_nz && L.c[0].ihalfN <= 0.25;
}

// SIMPLE: simple_laser(L) holds, so the second colour's parameters drop out of the code.
template <bool SIMPLE = false>
__host__ __device__ inline LaserBase laser_base(const LaserP& L, double t) {
    LaserBase b;
    b.c[0] = mono_base(L.c[0], t);
    b.c[1] = b.c[0];
    if (!SIMPLE && L.n == 2) b.c[1] = mono_base(L.c[1], t - L.delay);
    return b;
}

// the field at t + d, t the base time of `b`
template <bool SIMPLE = false>
__host__ __device__ inline void laser_field_at(const LaserP& L, const LaserBase& b, double d, double& Fx, double& Fy) {
    mono_field_at<SIMPLE>(L.c[0], b.c[0], d, Fx, Fy);
    if (!SIMPLE && L.n == 2) {
        double fx, fy;
        mono_field_at<false>(L.c[1], b.c[1], d, fx, fy);
        Fx += fx;
        Fy += fy;
    }
}

__host__ __device__ inline void laser_potential(const LaserP& L, double t, double& Ax, double& Ay) {
    mono_potential(L.c[0], t, Ax, Ay);
    if (L.n == 2) {
        double ax, ay;
        mono_potential(L.c[1], t - L.delay, ax, ay);
        Ax += ax;
        Ay += ay;
    }
}

// ---------------------------------------------------------------------------
// Lasers at complex time (the SFA-SPA sampler) and the field's time derivative
// (the SPANE sampler).
// ---------------------------------------------------------------------------
struct Cx {
    double re, im;
};
__host__ __device__ inline Cx cx(double r, double i = 0.0) { return Cx{r, i}; }
__host__ __device__ inline Cx operator+(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline Cx operator-(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline Cx operator*(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline Cx operator*(double s, Cx a) { return Cx{s * a.re, s * a.im}; }
__host__ __device__ inline Cx operator/(Cx a, double s) { return Cx{a.re / s, a.im / s}; }
// cos(a + ib) = cos a cosh b - i sin a sinh b;  sin(a + ib) = sin a cosh b + i cos a sinh b
__host__ __device__ inline void csincos(Cx z, Cx& s, Cx& c) {
    double sa, ca;
    sincos(z.re, &sa, &ca);
    const double chb = cosh(z.im), shb = sinh(z.im);
    s = Cx{sa * chb, ca * shb};
    c = Cx{ca * chb, -sa * shb};
}
__host__ __device__ inline Cx cexp_(Cx z) {
    double sb, cb;
    sincos(z.im, &sb, &cb);
    const double e = exp(z.re);
    return Cx{e * cb, e * sb};
}

// LaserFx/Fy/Ax/Ay (Cos2Laser.jl:97-152 and siblings) at complex t; the
// cut-off window takes real(t) (Cos2Laser.jl:93), as in the reference.
__host__ __device__ inline void mono_field_pot_c(const MonoP& m, Cx t, Cx& Fx, Cx& Fy, Cx& Ax, Cx& Ay) {
    t.re -= m.tshift;
    Cx sw, cw;
    csincos(Cx{m.w * t.re + m.cep, m.w * t.im}, sw, cw);
    Cx g, gA, a, b;
    if (m.kind == ETRAJ_LASER_GAUSSIAN) {
        const Cx t2 = t * t;
        gA = m.A0 * cexp_(Cx{-t2.re / 2 / m.sigma2, -t2.im / 2 / m.sigma2});
        g = gA;
        const Cx ts = t / m.sigma2;
        a = ts * cw + m.w * sw;
        b = ts * sw - m.w * cw;
    } else {
        const double win = cos_window(m, t.re);
        Cx sh, ch;
        csincos(Cx{m.w * t.re / (2 * m.N), m.w * t.im / (2 * m.N)}, sh, ch);
        const Cx c2 = ch * ch;
        const bool cos2 = m.kind == ETRAJ_LASER_COS2;
        gA = (m.A0 * win) * (cos2 ? c2 : c2 * c2);
        g = (m.F0 * win) * (cos2 ? ch : c2 * ch);
        a = ch * sw + m.coefN * (sh * cw);
        b = m.coefN * (sh * sw) - ch * cw;
    }
    if (!m.azi_nz) {
        Fx = g * a;
        Fy = m.ellip * (g * b);
        Ax = gA * cw;
        Ay = m.ellip * (gA * sw);
    } else {
        Fx = g * (m.cazi * a + (m.ellip * m.sazi) * b);
        Fy = g * ((-m.sazi) * a + (m.ellip * m.cazi) * b);
        Ax = gA * (m.cazi * cw + (m.ellip * m.sazi) * sw);
        Ay = gA * ((-m.sazi) * cw + (m.ellip * m.cazi) * sw);
    }
}

// mono_field_pot_c at t = tr + i tau with the parts that depend on tr alone
// computed once (one batch of the SPA sampler has one tr): sin / cos of the
// real parts of the carrier and envelope phases and the cut-off window (which
// takes real(t)). The arithmetic is that of mono_field_pot_c.
struct MonoTr {
    double sa, ca;     // sin, cos of w (tr - tshift) + cep
    double sae, cae;   // sin, cos of w (tr - tshift) / (2N)  (cos^n envelopes)
    double win;        // cos_window(tr - tshift)
};
struct LaserTr {
    MonoTr m[2];       // m[1] at tr - delay (bichromatic)
};
__host__ __device__ inline MonoTr mono_tr(const MonoP& m, double tr) {
    const double tre = tr - m.tshift;
    MonoTr r;
    sincos(m.w * tre + m.cep, &r.sa, &r.ca);
    r.sae = r.cae = r.win = 0.0;
    if (m.kind != ETRAJ_LASER_GAUSSIAN) {
        sincos(m.w * tre / (2 * m.N), &r.sae, &r.cae);
        r.win = cos_window(m, tre);
    }
    return r;
}
__host__ __device__ inline LaserTr laser_tr(const LaserP& L, double tr) {
    LaserTr r;
    r.m[0] = mono_tr(L.c[0], tr);
    r.m[1] = L.n == 2 ? mono_tr(L.c[1], tr - L.delay) : r.m[0];
    return r;
}
// cosh and sinh of x from one exp: cosh = (e + 1/e)/2, and sinh = (e - 1/e)/2
// for |x| >= 1/2, where the difference loses at most a bit and a half; below
// that the Taylor series of sinh through x^15 (first omitted term < 1e-19).
__host__ __device__ inline void cosh_sinh(double x, double& ch, double& sh) {
    const double ax = fabs(x);
    const double e = exp(ax);
    const double ie = recip_ge1(e);
    ch = 0.5 * (e + ie);
    const double x2 = x * x;
    double ps = 1.0 / 1307674368000.0;
    ps = fma(ps, x2, 1.0 / 6227020800.0);
    ps = fma(ps, x2, 1.0 / 39916800.0);
    ps = fma(ps, x2, 1.0 / 362880.0);
    ps = fma(ps, x2, 1.0 / 5040.0);
    ps = fma(ps, x2, 1.0 / 120.0);
    ps = fma(ps, x2, 1.0 / 6.0);
    const double small = fma(ps * x2, x, x);
    const double big = copysign(0.5 * (e - ie), x);
    sh = ax < 0.5 ? small : big;
}

// cosh and sinh of |x| <= 1/2 from their Taylor series through x^16 and x^15 (the
// first omitted terms are below 1e-19), no exp.
__host__ __device__ inline void cosh_sinh_small(double x, double& ch, double& sh) {
    const double x2 = x * x;
    double ps = inv_factorial(15);
#pragma unroll
    for (int k = 6; k >= 1; --k) ps = fma(ps, x2, inv_factorial(2 * k + 1));
    sh = fma(ps * x2, x, x);
    double pc = inv_factorial(16);
#pragma unroll
    for (int k = 7; k >= 1; --k) pc = fma(pc, x2, inv_factorial(2 * k));
    ch = fma(pc, x2, 1.0);
}

__host__ __device__ inline void mono_field_pot_tau(const MonoP& m, const MonoTr& r, double trm, double tau, Cx& Fx,
                                                   Cx& Fy, Cx& Ax, Cx& Ay) {
    const double wi = m.w * tau;
    double chb, shb;
    cosh_sinh(wi, chb, shb);
    const Cx sw{r.sa * chb, r.ca * shb}, cw{r.ca * chb, -r.sa * shb};
    Cx g, gA, a, b;
    if (m.kind == ETRAJ_LASER_GAUSSIAN) {
        const Cx t{trm - m.tshift, tau};
        const Cx t2 = t * t;
        gA = m.A0 * cexp_(Cx{-t2.re / 2 / m.sigma2, -t2.im / 2 / m.sigma2});
        g = gA;
        const Cx ts = t / m.sigma2;
        a = ts * cw + m.w * sw;
        b = ts * sw - m.w * cw;
    } else {
        const double ei = m.w * tau * m.ihalfN;
        double che, she;
        if (fabs(ei) <= 0.5) cosh_sinh_small(ei, che, she);   // the envelope's small argument
        else cosh_sinh(ei, che, she);
        const Cx sh{r.sae * che, r.cae * she}, ch{r.cae * che, -r.sae * she};
        const Cx c2 = ch * ch;
        const bool cos2 = m.kind == ETRAJ_LASER_COS2;
        gA = (m.A0 * r.win) * (cos2 ? c2 : c2 * c2);
        g = (m.F0 * r.win) * (cos2 ? ch : c2 * ch);
        a = ch * sw + m.coefN * (sh * cw);
        b = m.coefN * (sh * sw) - ch * cw;
    }
    if (!m.azi_nz) {
        Fx = g * a;
        Fy = m.ellip * (g * b);
        Ax = gA * cw;
        Ay = m.ellip * (gA * sw);
    } else {
        Fx = g * (m.cazi * a + (m.ellip * m.sazi) * b);
        Fy = g * ((-m.sazi) * a + (m.ellip * m.cazi) * b);
        Ax = gA * (m.cazi * cw + (m.ellip * m.sazi) * sw);
        Ay = gA * ((-m.sazi) * cw + (m.ellip * m.cazi) * sw);
    }
}
// laser_field_pot_c(L, tr + i tau) from laser_tr(L, tr)
__host__ __device__ inline void laser_field_pot_tau(const LaserP& L, const LaserTr& r, double tr, double tau, Cx& Fx,
                                                    Cx& Fy, Cx& Ax, Cx& Ay) {
    mono_field_pot_tau(L.c[0], r.m[0], tr, tau, Fx, Fy, Ax, Ay);
    if (L.n == 2) {
        Cx fx, fy, ax, ay;
        mono_field_pot_tau(L.c[1], r.m[1], tr - L.delay, tau, fx, fy, ax, ay);
        Fx = Fx + fx; Fy = Fy + fy; Ax = Ax + ax; Ay = Ay + ay;
    }
}

// Not inlined: the SPA sampler calls it from several loops, and one shared
// body keeps that kernel's register count within an occupancy of 2+.
__host__ __device__ inline __attribute__((noinline)) void laser_field_pot_c(const LaserP& L, Cx t, Cx& Fx, Cx& Fy, Cx& Ax,
                                                                     Cx& Ay) {
    mono_field_pot_c(L.c[0], t, Fx, Fy, Ax, Ay);
    if (L.n == 2) {
        Cx fx, fy, ax, ay;
        mono_field_pot_c(L.c[1], Cx{t.re - L.delay, t.im}, fx, fy, ax, ay);
        Fx = Fx + fx; Fy = Fy + fy; Ax = Ax + ax; Ay = Ay + ay;
    }
}

// (value, d/dt) pairs: the field and its full time derivative, window
// included (SPANESampler.jl:164-165 differentiates the whole closed form).
struct Dv {
    double v, d;
};
__host__ __device__ inline Dv operator+(Dv a, Dv b) { return Dv{a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dv operator*(Dv a, Dv b) { return Dv{a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dv operator*(double s, Dv a) { return Dv{s * a.v, s * a.d}; }
__host__ __device__ inline void dsincos(Dv x, Dv& s, Dv& c) {
    double sv, cv;
    sincos(x.v, &sv, &cv);
    s = Dv{sv, cv * x.d};
    c = Dv{cv, -sv * x.d};
}
__host__ __device__ inline Dv dtanh_(Dv x) {
    const double th = tanh(x.v);
    return Dv{th, (1 - th * th) * x.d};
}

__host__ __device__ inline void mono_field_deriv(const MonoP& m, double t0, double& Fx, double& Fy, double& dFx,
                                                 double& dFy) {
    const Dv t{t0 - m.tshift, 1.0};
    Dv sw, cw;
    dsincos(m.w * t + Dv{m.cep, 0.0}, sw, cw);
    Dv g, a, b;
    if (m.kind == ETRAJ_LASER_GAUSSIAN) {
        const double e = exp(-t.v * t.v / 2 / m.sigma2);
        g = Dv{m.A0 * e, m.A0 * e * (-t.v / m.sigma2)};
        const Dv ts = (1.0 / m.sigma2) * t;
        a = ts * cw + m.w * sw;
        b = ts * sw + (-m.w) * cw;
    } else {
        const Dv w1 = Dv{0.5, 0.0} + 0.5 * dtanh_(m.k * (t + Dv{m.halfw, 0.0}));
        const Dv w2 = Dv{0.5, 0.0} + 0.5 * dtanh_((-m.k) * (t + Dv{-m.halfw, 0.0}));
        Dv sh, ch;
        dsincos((m.w / (2 * m.N)) * t, sh, ch);
        const Dv env = (m.kind == ETRAJ_LASER_COS2) ? ch : ch * ch * ch;
        g = (m.F0 * env) * w1 * w2;
        a = ch * sw + m.coefN * (sh * cw);
        b = m.coefN * (sh * sw) + (-1.0) * (ch * cw);
    }
    Dv fx, fy;
    if (!m.azi_nz) {
        fx = g * a;
        fy = m.ellip * (g * b);
    } else {
        fx = g * (m.cazi * a + (m.ellip * m.sazi) * b);
        fy = g * ((-m.sazi) * a + (m.ellip * m.cazi) * b);
    }
    Fx = fx.v; Fy = fy.v; dFx = fx.d; dFy = fy.d;
}

__host__ __device__ inline void laser_field_deriv(const LaserP& L, double t, double& Fx, double& Fy, double& dFx,
                                                  double& dFy) {
    mono_field_deriv(L.c[0], t, Fx, Fy, dFx, dFy);
    if (L.n == 2) {
        double a, b, c, d;
        mono_field_deriv(L.c[1], t - L.delay, a, b, c, d);
        Fx += a; Fy += b; dFx += c; dFy += d;
    }
}

// ---------------------------------------------------------------------------
// Targets (HydrogenLikeAtom.jl:75-76, SAEAtom.jl:84-96)
// ---------------------------------------------------------------------------
struct TargetP {
    int kind;
    int Z;
    double Zd, Ip, C, sc;
    double a1, b1, a2, b2, a3, b3;
};

inline int make_target(const etraj_target_desc& d, TargetP& T) {
    if (d.kind != ETRAJ_TARGET_HYDROGEN_LIKE && d.kind != ETRAJ_TARGET_SAE) return -1;
    if (!(d.Ip > 0) || !(d.soft_core >= 0)) return -1;
    T.kind = d.kind;
    T.Z = d.nucl_charge;
    T.Zd = (double)d.nucl_charge;
    T.Ip = d.Ip;
    T.C = d.asymp_coeff;
    T.sc = d.soft_core;
    T.a1 = d.a1; T.b1 = d.b1; T.a2 = d.a2; T.b2 = d.b2; T.a3 = d.a3; T.b3 = d.b3;
    return 0;
}

// Kernel-internal target kind: a hydrogen-like target whose soft core is not
// negligible next to |r|^2 in FP32 (the molecular a = 1); same physics as
// ETRAJ_TARGET_HYDROGEN_LIKE, selected once per wavefront.
constexpr int kTargetHSoft = 3;
constexpr bool is_hydrogen_like(int tgt) { return tgt == ETRAJ_TARGET_HYDROGEN_LIKE || tgt == kTargetHSoft; }

template <typename R>
__host__ __device__ inline R rsqrt_t(R x);
template <>
__host__ __device__ inline double rsqrt_t<double>(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return 1.0 / sqrt(x);
#endif
}
// FP32: the hardware reciprocal square root (v_rsq_f32, 1 ulp); arguments are
// r^2 + soft_core >= 1e-10, far from the denormal range.
template <>
__host__ __device__ inline float rsqrt_t<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// Potential V(r) and force scale s with F_target = s * r_vec, evaluated in
// real type R, for target kind TGT. `sae_quirk_r2` reproduces SAEAtom.jl:136
// (2D SCTS divides by r^2+a instead of sqrt(r^2+a)) for the phase potential.
template <int TGT, typename R>
__host__ __device__ inline void target_vs_k(const TargetP& T, R r2n, bool want_v, bool sae_quirk_r2, R& V, R& s) {
    if (is_hydrogen_like(TGT)) {
        R ir = rsqrt_t<R>(r2n + (R)T.sc);
        R zr = (R)T.Zd * ir;
        s = -(zr * ir * ir);
        V = -zr;
    } else {
        R r = sqrt(r2n);
        R rs2 = r2n + (R)T.sc;
        R irs = rsqrt_t<R>(rs2);
        R e1 = exp(-(R)T.b1 * r), e2 = exp(-(R)T.b2 * r), e3 = exp(-(R)T.b3 * r);
        s = -((irs * irs * irs) * ((R)T.Zd + (R)T.a1 * (1 + (R)T.b1 * r) * e1 + (R)T.a3 * (1 + (R)T.b3 * r) * e3)
              + (R)T.a2 * (R)T.b2 * irs * e2);
        if (want_v) {
            R num = (R)T.Zd + (R)T.a1 * e1 + (R)T.a2 * r * e2 + (R)T.a3 * e3;
            V = sae_quirk_r2 ? -num / rs2 : -num * irs;
        } else {
            V = 0;
        }
    }
}

template <typename R>
__host__ __device__ inline void target_vs(const TargetP& T, R r2n, bool want_v, bool sae_quirk_r2, R& V, R& s) {
    if (T.kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        target_vs_k<ETRAJ_TARGET_HYDROGEN_LIKE, R>(T, r2n, want_v, sae_quirk_r2, V, s);
    else
        target_vs_k<ETRAJ_TARGET_SAE, R>(T, r2n, want_v, sae_quirk_r2, V, s);
}

// ---------------------------------------------------------------------------
// Equations of motion (TrajectoryFunction, HydrogenLikeAtom.jl:78-159,
// SAEAtom.jl:98-186): r' = p, p' = F_target(r) - F_laser(t) (z: target
// only); QTMC phase' = -(Ip + p^2/2 + V); SCTS phase' = -(p^2/2 + V + r.F_target).
// ---------------------------------------------------------------------------
template <int DIM, typename R>
struct State {
    R r[DIM];
    R p[DIM];
};

template <int DIM, int PHASE, int TGT, typename R>
__host__ __device__ inline void eom(const TargetP& T, bool sae_quirk, const State<DIM, R>& u, R Fx, R Fy,
                                    State<DIM, R>& du, R& dphi, R& r2out) {
    if (sizeof(R) == 4 && is_hydrogen_like(TGT)) {
        // FP32 form: the soft core a is folded into r2 = |r|^2 + a once, so
        // V = -Z/sqrt(r2) and r.F_target = -Z (r2 - a)/r2^(3/2); the SCTS
        // integrand -(p^2/2 + V + r.F) is Z/sqrt(r2) (2 - a/r2) - p^2/2. For
        // the atomic default a = 1e-10 the a/r2 term is below FP32 resolution
        // wherever the FP32 pass integrates (r2 >= the re-step radius^2) and is
        // skipped; kTargetHSoft (a molecular a = 1) keeps it.
        R r2 = fma(u.r[1], u.r[1], (R)T.sc);
        if (DIM == 3) r2 = fma(u.r[DIM - 1], u.r[DIM - 1], r2);
        r2 = fma(u.r[0], u.r[0], r2);
        r2out = r2;
        const R ir = rsqrt_t<R>(r2);
        const R zr = (R)T.Zd * ir;
        const R sm = zr * (ir * ir);   // -s; the sign goes into the FMAs' operand modifiers
#pragma unroll
        for (int i = 0; i < DIM; ++i) du.r[i] = u.p[i];
        du.p[0] = fma(-sm, u.r[0], -Fx);
        du.p[1] = fma(-sm, u.r[1], -Fy);
        if (DIM == 3) du.p[DIM - 1] = -(sm * u.r[DIM - 1]);
        R k2 = u.p[1] * u.p[1];
        if (DIM == 3) k2 = fma(u.p[DIM - 1], u.p[DIM - 1], k2);
        k2 = fma(u.p[0], u.p[0], k2);
        if (PHASE == ETRAJ_PHASE_QTMC) dphi = fma(k2, (R)-0.5, zr - (R)T.Ip);
        else if (PHASE == ETRAJ_PHASE_SCTS) {
            if (TGT == kTargetHSoft) dphi = fma(k2, (R)-0.5, zr * fma(-(R)T.sc, ir * ir, (R)2));
            else dphi = fma(k2, (R)-0.5, zr + zr);
        }
        else dphi = 0;
        return;
    }
    R r2n = u.r[0] * u.r[0] + u.r[1] * u.r[1];
    if (DIM == 3) r2n += u.r[DIM - 1] * u.r[DIM - 1];
    r2out = r2n;
    R V, s;
    target_vs_k<TGT, R>(T, r2n, PHASE != ETRAJ_PHASE_CTMC, sae_quirk, V, s);
#pragma unroll
    for (int i = 0; i < DIM; ++i) du.r[i] = u.p[i];
    du.p[0] = s * u.r[0] - Fx;
    du.p[1] = s * u.r[1] - Fy;
    if (DIM == 3) du.p[DIM - 1] = s * u.r[DIM - 1];
    R k2 = u.p[0] * u.p[0] + u.p[1] * u.p[1];
    if (DIM == 3) k2 += u.p[DIM - 1] * u.p[DIM - 1];
    if (PHASE == ETRAJ_PHASE_QTMC) {
        dphi = -((R)T.Ip + k2 * (R)0.5 + V);
    } else if (PHASE == ETRAJ_PHASE_SCTS) {
        // r . F_target = s * |r|^2 (without the soft core)
        dphi = -(k2 * (R)0.5 + V + s * r2n);
    } else {
        dphi = 0;
    }
}

// ---------------------------------------------------------------------------
// One classical RK4 step of size h from (r, p), stage fields F0 at t, Fm at
// t + h/2 (stages 2 and 3) and F1 at t + h, in the form the step loops use.
// Since r' = p, the r-slope of stage j is the stage's momentum p_j, and the
// step is
//   r += h/6 Sr,  Sr = p_1 + 2 p_2 + 2 p_3 + p_4      (p_1 = p),
//   p += h/6 Sp,  Sp = k_1 + 2 k_2 + 2 k_3 + k_4      (k_j the stage forces),
// with stage states (r + h/2 p_1, p + h/2 k_1), (r + h/2 p_2, p + h/2 k_2),
// (r + h p_3, p + h k_3). The sums are formed stage by stage (progressive
// FMAs), so the caller adds h/6 Sr to its state with one FMA per component.
// ---------------------------------------------------------------------------

// Force and phase-rate parts of the hydrogen-like EOM (HydrogenLikeAtom.jl:78-159,
// V = -Z/sqrt(|r|^2 + a)) at one stage: k = -Z r/r2^(3/2) - F (z: no field),
// and pa, with the phase rate = cA pa - |p|^2/2 + c0:
//   QTMC: pa = Z/sqrt(r2), cA = 1, c0 = -Ip;
//   SCTS: pa = Z/sqrt(r2) (2 - a/r2), cA = 1 (SOFT), or pa = Z/sqrt(r2), cA = 2:
//         -(p^2/2 + V + r.F_target) with r.F_target = -Z (r2 - a)/r2^(3/2);
//         the a/r2 term is dropped (not SOFT) only where it is below the
//         arithmetic's resolution (FP32 with an atomic a = 1e-10).
// UNITZ: Z = 1 (Z/sqrt(r2) is 1/sqrt(r2)). r2 = |r|^2 + a is returned.
template <typename R, int DIM, int PHASE, bool SOFT, bool UNITZ>
__host__ __device__ __forceinline__ void hl_stage(R Z, R a, const R* yr, R Fx, R Fy, R* k, R& pa, R& r2) {
    R s2 = fma(yr[1], yr[1], a);
    if (DIM == 3) s2 = fma(yr[DIM - 1], yr[DIM - 1], s2);
    r2 = fma(yr[0], yr[0], s2);
    const R ir = rsqrt_t<R>(r2);
    const R zr = UNITZ ? ir : Z * ir;
    const R ir2 = ir * ir;
    const R sm = zr * ir2;
    k[0] = fma(-sm, yr[0], -Fx);
    k[1] = fma(-sm, yr[1], -Fy);
    if (DIM == 3) k[DIM - 1] = -(sm * yr[DIM - 1]);
    if (PHASE != ETRAJ_PHASE_CTMC) pa = (PHASE == ETRAJ_PHASE_SCTS && SOFT) ? zr * fma(-a, ir2, (R)2) : zr;
}

// acc + |p|^2 (one FMA per component)
template <typename R, int DIM>
__host__ __device__ __forceinline__ R add_p2(const R* p, R acc) {
#pragma unroll
    for (int i = DIM - 1; i >= 0; --i) acc = fma(p[i], p[i], acc);
    return acc;
}

// The step sums of the hydrogen-like EOM: Sr, Sp as above; SA = sum_j w_j pa_j
// and SQ = sum_j w_j q_j (w = 1, 2, 2, 1), so that the phase increment over
// the step is h/6 (cA SA - SQ/2) + h c0. r2first: |r|^2 + a at the step's start.
template <typename R, int DIM, int PHASE, bool SOFT, bool UNITZ>
// hh = h / 2 (passed in, so that a wave-uniform caller can keep it in a scalar register).
__host__ __device__ __forceinline__ void hl_rk4_sums(R Z, R a, const R* r, const R* p, R h, R hh, R F0x, R F0y, R Fmx,
                                                     R Fmy, R F1x, R F1y, R* Sr, R* Sp, R& SA, R& SQ, R& r2first) {
    constexpr bool PH = PHASE != ETRAJ_PHASE_CTMC;
    R k[DIM], yr[DIM], yp[DIM], pa = 0, r2, Q14 = 0, Q23 = 0;   // |p|^2 sums of stages 1 + 4 and 2 + 3
    hl_stage<R, DIM, PHASE, SOFT, UNITZ>(Z, a, r, F0x, F0y, k, pa, r2first);
    if (PH) Q14 = add_p2<R, DIM>(p, (R)0);
    SA = pa;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = p[i];
        Sp[i] = k[i];
        yr[i] = fma(hh, p[i], r[i]);
        yp[i] = fma(hh, k[i], p[i]);
    }
    hl_stage<R, DIM, PHASE, SOFT, UNITZ>(Z, a, yr, Fmx, Fmy, k, pa, r2);
    if (PH) Q23 = add_p2<R, DIM>(yp, (R)0);
    SA = fma((R)2, pa, SA);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = fma((R)2, yp[i], Sr[i]);
        Sp[i] = fma((R)2, k[i], Sp[i]);
        yr[i] = fma(hh, yp[i], r[i]);
        yp[i] = fma(hh, k[i], p[i]);
    }
    hl_stage<R, DIM, PHASE, SOFT, UNITZ>(Z, a, yr, Fmx, Fmy, k, pa, r2);
    if (PH) Q23 = add_p2<R, DIM>(yp, Q23);
    SA = fma((R)2, pa, SA);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = fma((R)2, yp[i], Sr[i]);
        Sp[i] = fma((R)2, k[i], Sp[i]);
        yr[i] = fma(h, yp[i], r[i]);
        yp[i] = fma(h, k[i], p[i]);
    }
    hl_stage<R, DIM, PHASE, SOFT, UNITZ>(Z, a, yr, F1x, F1y, k, pa, r2);
    if (PH) Q14 = add_p2<R, DIM>(yp, Q14);
    SA = SA + pa;
    SQ = fma((R)2, Q23, Q14);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = Sr[i] + yp[i];
        Sp[i] = Sp[i] + k[i];
    }
}

// The phase-increment constants of hl_rk4_sums for a step of size h:
// increment = cA SA + cQ SQ + c0 (cA = h/6 or h/3, cQ = -h/12, c0 = -h Ip for QTMC).
template <typename R>
struct HlPhase {
    R cA, cQ, c0;
};
template <typename R, int PHASE, bool SOFT>
__host__ __device__ inline HlPhase<R> hl_phase(double h, double Ip) {
    HlPhase<R> c;
    const bool twice = PHASE == ETRAJ_PHASE_SCTS && !SOFT;
    c.cA = (R)(twice ? h / 3 : h / 6);
    c.cQ = (R)(-h / 12);
    c.c0 = (R)(PHASE == ETRAJ_PHASE_QTMC ? -h * Ip : 0.0);
    return c;
}

// The same sums for any target through its full EOM (eom): Sr, Sp as above and
// SPh = sum_j w_j phase'_j (phase increment h/6 SPh).
template <typename R, int DIM, int PHASE, int TGT>
__host__ __device__ __forceinline__ void gen_rk4_sums(const TargetP& T, bool sae_quirk, const R* r, const R* p, R h,
                                                      R F0x, R F0y, R Fmx, R Fmy, R F1x, R F1y, R* Sr, R* Sp,
                                                      R& SPh, R& r2first) {
    const R hh = h * (R)0.5;
    State<DIM, R> u, y, k;
    R ph, r2;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        u.r[i] = r[i];
        u.p[i] = p[i];
    }
    eom<DIM, PHASE, TGT, R>(T, sae_quirk, u, F0x, F0y, k, ph, r2first);
    SPh = ph;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = p[i];
        Sp[i] = k.p[i];
        y.r[i] = fma(hh, p[i], r[i]);
        y.p[i] = fma(hh, k.p[i], p[i]);
    }
    eom<DIM, PHASE, TGT, R>(T, sae_quirk, y, Fmx, Fmy, k, ph, r2);
    SPh = fma((R)2, ph, SPh);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = fma((R)2, y.p[i], Sr[i]);
        Sp[i] = fma((R)2, k.p[i], Sp[i]);
        y.r[i] = fma(hh, y.p[i], r[i]);
        y.p[i] = fma(hh, k.p[i], p[i]);
    }
    eom<DIM, PHASE, TGT, R>(T, sae_quirk, y, Fmx, Fmy, k, ph, r2);
    SPh = fma((R)2, ph, SPh);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = fma((R)2, y.p[i], Sr[i]);
        Sp[i] = fma((R)2, k.p[i], Sp[i]);
        y.r[i] = fma(h, y.p[i], r[i]);
        y.p[i] = fma(h, k.p[i], p[i]);
    }
    eom<DIM, PHASE, TGT, R>(T, sae_quirk, y, F1x, F1y, k, ph, r2);
    SPh = SPh + ph;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        Sr[i] = Sr[i] + y.p[i];
        Sp[i] = Sp[i] + k.p[i];
    }
}

// ---------------------------------------------------------------------------
// Post-processing after traj_t_final (TrajectorySimulation.jl:356-396,
// :451-496): E_inf, SCTS asymptotic phase, Kepler asymptotic momentum.
// Always FP64.
// ---------------------------------------------------------------------------
struct Final {
    double E, phase, pinf[3];
    bool ionized;
};

__host__ __device__ inline double target_potential(const TargetP& T, double x, double y, double z) {
    double r2 = x * x + y * y + z * z;
    double V, s;
    target_vs<double>(T, r2, true, false, V, s);
    return V;
}

template <int DIM, int PHASE>
__host__ __device__ inline Final kepler_final(const TargetP& T, int quirks, const double* r, const double* p, double phase, double t0) {
    Final f;
    double x = r[0], y = r[1], z = DIM == 3 ? r[2] : 0.0;
    double px = p[0], py = p[1], pz = DIM == 3 ? p[2] : 0.0;
    double E = (px * px + py * py + pz * pz) / 2 + target_potential(T, x, y, z);
    double Lx = y * pz - z * py, Ly = z * px - x * pz, Lz = x * py - y * px;
    double L2 = Lx * Lx + Ly * Ly + Lz * Lz;
    if (E > 0 && PHASE == ETRAJ_PHASE_SCTS) {
        phase += T.Ip * t0;                                   // Ip*tr (:364)
        double pinf = sqrt(2 * E);
        double ib = 1.0 / pinf;                               // (2E)^(-1/2)
        double g = sqrt(1 + 2 * E * L2);
        double rp = (DIM == 2) ? x * px + y * py
                               : ((quirks & ETRAJ_QUIRK_SCTS3D_XPY) ? x * py + y * py + z * pz : x * px + y * py + z * pz);
        phase -= T.Zd * ib * (log(g) + asinh(rp / (g * ib)));  // (:368 / :463)
    }
    f.E = E;
    f.phase = phase;
    f.ionized = E >= 0;
    if (f.ionized) {
        double pinf = sqrt(2 * E);
        double ir = 1.0 / sqrt(x * x + y * y + z * z);
        double ax = (py * Lz - pz * Ly) - T.Zd * x * ir;    // a = p x L - Z r/|r| (:373)
        double ay = (pz * Lx - px * Lz) - T.Zd * y * ir;
        double az = (px * Ly - py * Lx) - T.Zd * z * ir;
        double lax = Ly * az - Lz * ay, lay = Lz * ax - Lx * az, laz = Lx * ay - Ly * ax;
        double c = pinf / (1 + pinf * pinf * L2);            // (:374)
        f.pinf[0] = c * (pinf * lax - ax);
        f.pinf[1] = c * (pinf * lay - ay);
        f.pinf[2] = c * (pinf * laz - az);
    } else {
        f.pinf[0] = f.pinf[1] = f.pinf[2] = 0.0;
    }
    return f;
}

// round(Int, (p+pmax)/(pmax/num*2)) with ties to even, 1-based; 0 when the
// value is out of the representable range (TrajectorySimulation.jl:375-384).
__host__ __device__ inline int64_t bin_index(double p, double pmax, int num) {
    if (num <= 1) return 1;
    double q = (p + pmax) / (pmax / num * 2);
    if (!(fabs(q) < 9.0e15)) return 0;
    return (int64_t)rint(q);
}

}  // namespace etraj
