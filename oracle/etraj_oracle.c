/*
 * etraj_oracle.c -- CPU FP64 restatement of eTraj.jl's trajectory-simulation
 * path, used ONLY as test infrastructure: by tests/, by the smoke check and by
 * bench.py's cpu_baseline leg. The product (libetraj_hip.so) never links,
 * loads or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the eTraj.jl source tree). Compiled with -ffp-contract=off so every
 * operation is a plain IEEE FP64 operation in the order written.
 *
 * What pins each part (tests under tests/):
 *   philox4x64_10        numpy.random.Philox (same generator, independent code)
 *   laser_*              dA/dt = -F by 8th-order central differences (Lasers_test.jl:40-54),
 *                        property identities (Lasers_test.jl:29-39), zero outside the window
 *   target_*             V(1,1,1) closed forms and grad V = -F (Targets_test.jl:31-36,59-64)
 *   adk_gen_batch        the reference's survivor-count KATs (ElectronSamplers_test.jl:53-130)
 *   rk4 / tsit5          tableau order conditions, convergence order, Z=0 closed form
 *   kepler_post          p_inf invariance along a field-free Coulomb orbit, |p_inf|=sqrt(2E)
 *   scts asymptote       Phi(t)+S_C(t)+E t constant after the pulse
 *   bin_index            hand-built edge cases (ties-to-even, index 0, num == 1)
 * Parity unpinned: the exact Tsit5 step sequence of OrdinaryDiffEq 6.89 (the
 * reference cannot run here), so Tsit5 output is pinned only by the
 * mathematical checks above, never by reference vectors.
 */
#include "etraj_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================== */
/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based RNG.      */
/* ========================================================================== */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

void oracle_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* stream tags (counter word 3) */
#define TAG_T0 1ULL
#define TAG_K  2ULL

static inline double u01_open_right(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }       /* [0,1) */
static inline double u01_open_left(uint64_t x) { return ((double)(x >> 11) + 1.0) * 0x1.0p-53; } /* (0,1] */

/* ========================================================================== */
/* Lasers                                                                      */
/* ========================================================================== */
typedef struct {
    int kind;
    double F0, A0, w, N, cep, dt_shift, ellip, azi, k, sigma;
} mono_t;

typedef struct {
    int n;
    double delay;
    mono_t c[2];
} laser_t;

/* LaserF0/AngFreq/Period etc.: Cos2Laser.jl:81-86, Cos4Laser.jl (same lines),
 * GaussianLaser.jl:119-136 (SpreadDuration = spread_cyc_num * Period). */
static int mono_init(const etraj_mono_laser* d, mono_t* m) {
    if (d->kind != ETRAJ_LASER_COS2 && d->kind != ETRAJ_LASER_COS4 && d->kind != ETRAJ_LASER_GAUSSIAN) return -1;
    if (!(d->wave_len > 0) || !(d->cyc_num > 0)) return -1;
    m->kind = d->kind;
    m->w = 45.563352525 / d->wave_len;
    double period = 2 * M_PI / m->w;
    m->F0 = sqrt(d->peak_int / (1.0 + d->ellip * d->ellip) / 3.50944521e16);
    m->A0 = m->F0 / m->w;
    m->N = d->cyc_num;
    m->cep = d->cep;
    m->dt_shift = d->t_shift;
    m->ellip = d->ellip;
    m->azi = d->azi;
    m->k = 200 / period;
    m->sigma = d->cyc_num * period;
    return 0;
}

static int laser_init(const etraj_laser_desc* d, laser_t* L) {
    if (d->n_colors != 1 && d->n_colors != 2) return -1;
    L->n = d->n_colors;
    L->delay = d->delay21;
    for (int i = 0; i < L->n; ++i)
        if (mono_init(&d->color[i], &L->c[i])) return -1;
    return 0;
}

/* smooth cut-off window of the cos^n envelopes (Cos2Laser.jl:93) */
static inline double cosn_window(const mono_t* m, double t) {
    return (0.5 + 0.5 * tanh(m->k * (t + m->N * M_PI / m->w))) * (0.5 + 0.5 * tanh(-m->k * (t - m->N * M_PI / m->w)));
}

/* UnitEnvelope: Cos2Laser.jl:89-95, Cos4Laser.jl:89-95, GaussianLaser.jl:138-144 */
static double mono_envelope(const mono_t* m, double t) {
    t -= m->dt_shift;
    if (m->kind == ETRAJ_LASER_GAUSSIAN) return exp(-t * t / 2 / (m->sigma * m->sigma));
    double c = cos(m->w * t / (2 * m->N));
    double cn = (m->kind == ETRAJ_LASER_COS2) ? c * c : (c * c) * (c * c);
    return cn * cosn_window(m, t);
}

/* LaserFx/LaserFy: Cos2Laser.jl:125-152, Cos4Laser.jl:125-152, GaussianLaser.jl:176-203 */
static void mono_F(const mono_t* m, double t, double* Fx, double* Fy) {
    t -= m->dt_shift;
    const double w = m->w, N = m->N, phi = m->cep, e = m->ellip, az = m->azi;
    if (m->kind == ETRAJ_LASER_GAUSSIAN) {
        double s2 = m->sigma * m->sigma;
        double g = m->A0 * exp(-t * t / 2 / s2);
        double cw = cos(w * t + phi), sw = sin(w * t + phi);
        double a = t / s2 * cw + w * sw;   /* x-like component */
        double b = t / s2 * sw - w * cw;   /* y-like component */
        if (az == 0) {
            *Fx = g * a;
            *Fy = g * b * e;
        } else {
            *Fx = g * (a * cos(az) + b * e * sin(az));
            *Fy = g * (a * -sin(az) + b * e * cos(az));
        }
        return;
    }
    double ch = cos(w * t / (2 * N)), sh = sin(w * t / (2 * N));
    double cw = cos(w * t + phi), sw = sin(w * t + phi);
    double win = cosn_window(m, t);
    double pre, a, b;
    if (m->kind == ETRAJ_LASER_COS2) {
        pre = m->F0 * ch * win;
        a = ch * sw + 1 / N * sh * cw;        /* Cos2Laser.jl:130 */
        b = -ch * cw + 1 / N * sh * sw;       /* Cos2Laser.jl:144 */
    } else {
        pre = m->F0 * (ch * ch * ch) * win;
        a = ch * sw + 2 / N * sh * cw;        /* Cos4Laser.jl:130 */
        b = -ch * cw + 2 / N * sh * sw;       /* Cos4Laser.jl:144 */
    }
    if (az == 0) {
        *Fx = pre * a;
        *Fy = pre * b * e;
    } else {
        *Fx = pre * (a * cos(az) + b * e * sin(az));
        *Fy = pre * (a * -sin(az) + b * e * cos(az));
    }
}

/* LaserAx/LaserAy: Cos2Laser.jl:97-124, Cos4Laser.jl:97-124, GaussianLaser.jl:146-175 */
static void mono_A(const mono_t* m, double t, double* Ax, double* Ay) {
    double env = mono_envelope(m, t);
    t -= m->dt_shift;
    double cw = cos(m->w * t + m->cep), sw = sin(m->w * t + m->cep);
    double g = m->A0 * env;
    if (m->azi == 0) {
        *Ax = g * cw;
        *Ay = g * sw * m->ellip;
    } else {
        *Ax = g * (cw * cos(m->azi) + sw * m->ellip * sin(m->azi));
        *Ay = g * (cw * -sin(m->azi) + sw * m->ellip * cos(m->azi));
    }
}

static inline void laser_F(const laser_t* L, double t, double* Fx, double* Fy) {
    mono_F(&L->c[0], t, Fx, Fy);
    if (L->n == 2) { /* BichromaticLaser.jl:73-85 */
        double fx2, fy2;
        mono_F(&L->c[1], t - L->delay, &fx2, &fy2);
        *Fx += fx2;
        *Fy += fy2;
    }
}

static inline void laser_A(const laser_t* L, double t, double* Ax, double* Ay) {
    mono_A(&L->c[0], t, Ax, Ay);
    if (L->n == 2) { /* BichromaticLaser.jl:59-71 */
        double ax2, ay2;
        mono_A(&L->c[1], t - L->delay, &ax2, &ay2);
        *Ax += ax2;
        *Ay += ay2;
    }
}

int oracle_laser_field(const etraj_laser_desc* d, const double* t, int64_t n,
                       double* Fx, double* Fy, double* Ax, double* Ay) {
    laser_t L;
    if (laser_init(d, &L)) return ETRAJ_EINVAL;
    for (int64_t i = 0; i < n; ++i) {
        double fx, fy, ax, ay;
        laser_F(&L, t[i], &fx, &fy);
        laser_A(&L, t[i], &ax, &ay);
        if (Fx) Fx[i] = fx;
        if (Fy) Fy[i] = fy;
        if (Ax) Ax[i] = ax;
        if (Ay) Ay[i] = ay;
    }
    return ETRAJ_OK;
}

double oracle_unit_envelope(const etraj_laser_desc* d, double t) {
    laser_t L;
    if (laser_init(d, &L)) return NAN;
    return mono_envelope(&L.c[0], t);
}

/* ========================================================================== */
/* Targets                                                                     */
/* ========================================================================== */
typedef struct {
    int kind, Z, l, m;
    double Ip, C, sc, a1, b1, a2, b2, a3, b3;
    double qth, qph;   /* QuantizationAxisOrientaion (theta, phi) */
} target_t;

static int target_init(const etraj_target_desc* d, target_t* T) {
    if (d->kind != ETRAJ_TARGET_HYDROGEN_LIKE && d->kind != ETRAJ_TARGET_SAE) return -1;
    if (!(d->Ip > 0) || d->soft_core < 0) return -1;
    T->kind = d->kind; T->Z = d->nucl_charge; T->l = d->l; T->m = d->m;
    T->Ip = d->Ip; T->C = d->asymp_coeff; T->sc = d->soft_core;
    T->a1 = d->a1; T->b1 = d->b1; T->a2 = d->a2; T->b2 = d->b2; T->a3 = d->a3; T->b3 = d->b3;
    T->qth = d->quan_ax_theta; T->qph = d->quan_ax_phi;
    return 0;
}

/* TargetPotential: HydrogenLikeAtom.jl:75, SAEAtom.jl:84-89 */
static inline double target_V(const target_t* T, double x, double y, double z) {
    if (T->kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        return -T->Z * pow(x * x + y * y + z * z + T->sc, -0.5);
    double r = sqrt(x * x + y * y + z * z);
    return -(T->Z + T->a1 * exp(-T->b1 * r) + T->a2 * r * exp(-T->b2 * r) + T->a3 * exp(-T->b3 * r)) / sqrt(r * r + T->sc);
}

/* TargetForce: HydrogenLikeAtom.jl:76, SAEAtom.jl:91-96 (returns the scalar s with F = s * (x,y,z)) */
static inline double target_Fscale(const target_t* T, double x, double y, double z) {
    if (T->kind == ETRAJ_TARGET_HYDROGEN_LIKE)
        return -T->Z * pow(x * x + y * y + z * z + T->sc, -1.5);
    double r = sqrt(x * x + y * y + z * z);
    return -(pow(r * r + T->sc, -1.5) * (T->Z + T->a1 * (1 + T->b1 * r) * exp(-T->b1 * r) + T->a3 * (1 + T->b3 * r) * exp(-T->b3 * r))
             + T->a2 * T->b2 / sqrt(r * r + T->sc) * exp(-T->b2 * r));
}

int oracle_target_eval(const etraj_target_desc* d, const double* xyz, int64_t n, double* V, double* F) {
    target_t T;
    if (target_init(d, &T)) return ETRAJ_EINVAL;
    for (int64_t i = 0; i < n; ++i) {
        double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (V) V[i] = target_V(&T, x, y, z);
        if (F) {
            double s = target_Fscale(&T, x, y, z);
            F[3 * i] = s * x; F[3 * i + 1] = s * y; F[3 * i + 2] = s * z;
        }
    }
    return ETRAJ_OK;
}

/* ========================================================================== */
/* Equations of motion: TrajectoryFunction                                     */
/*   HydrogenLikeAtom.jl:78-159, SAEAtom.jl:98-186                             */
/* state (2D) x y px py [phi];  (3D) x y z px py pz [phi]                      */
/* ========================================================================== */
typedef struct {
    const laser_t* L;
    const target_t* T;
    int dim, phase, quirks;
} eom_t;

static void eom_rhs(const eom_t* E, double t, const double* u, double* du) {
    const target_t* T = E->T;
    double Fx, Fy;
    laser_F(E->L, t, &Fx, &Fy);
    double x = u[0], y = u[1], z = 0, px, py, pz = 0;
    if (E->dim == 2) { px = u[2]; py = u[3]; }
    else { z = u[2]; px = u[3]; py = u[4]; pz = u[5]; }
    double tFx, tFy, tFz = 0, V = 0;
    if (T->kind == ETRAJ_TARGET_HYDROGEN_LIKE) {
        double r2 = (E->dim == 2) ? x * x + y * y + T->sc : x * x + y * y + z * z + T->sc;
        double s = -T->Z * pow(r2, -1.5);
        tFx = s * x; tFy = s * y; tFz = s * z;
        if (E->phase != ETRAJ_PHASE_CTMC) V = -T->Z * pow(r2, -0.5);
    } else {
        double r = (E->dim == 2) ? sqrt(x * x + y * y) : sqrt(x * x + y * y + z * z);
        double s = -(pow(r * r + T->sc, -1.5) * (T->Z + T->a1 * (1 + T->b1 * r) * exp(-T->b1 * r) + T->a3 * (1 + T->b3 * r) * exp(-T->b3 * r))
                     + T->a2 * T->b2 / sqrt(r * r + T->sc) * exp(-T->b2 * r));
        tFx = s * x; tFy = s * y; tFz = s * z;
        if (E->phase != ETRAJ_PHASE_CTMC) {
            double num = T->Z + T->a1 * exp(-T->b1 * r) + T->a2 * r * exp(-T->b2 * r) + T->a3 * exp(-T->b3 * r);
            /* SAEAtom.jl:136 divides by (r^2+sc) in 2D SCTS; :124/:166/:181 use sqrt */
            if (E->dim == 2 && E->phase == ETRAJ_PHASE_SCTS && (E->quirks & ETRAJ_QUIRK_SAE2D_SCTS))
                V = -num / (r * r + T->sc);
            else
                V = -num / sqrt(r * r + T->sc);
        }
    }
    if (E->dim == 2) {
        du[0] = px; du[1] = py; du[2] = tFx - Fx; du[3] = tFy - Fy;
        double k2 = (px * px + py * py) / 2;
        if (E->phase == ETRAJ_PHASE_QTMC) du[4] = -(T->Ip + k2 + V);                 /* :101 */
        else if (E->phase == ETRAJ_PHASE_SCTS) du[4] = -(k2 + V + (x * tFx + y * tFy)); /* :112 */
    } else {
        du[0] = px; du[1] = py; du[2] = pz; du[3] = tFx - Fx; du[4] = tFy - Fy; du[5] = tFz;
        double k2 = (px * px + py * py + pz * pz) / 2;
        if (E->phase == ETRAJ_PHASE_QTMC) du[6] = -(T->Ip + k2 + V);                                /* :140 */
        else if (E->phase == ETRAJ_PHASE_SCTS) du[6] = -(k2 + V + (x * tFx + y * tFy + z * tFz));     /* :154 */
    }
}

/* ========================================================================== */
/* Integrators                                                                 */
/* ========================================================================== */

/* Fixed-step classical RK4, the engine's extension of the reference's ODE
 * solve (TrajectorySimulation.jl:342 uses Tsit5): n = ceil((tf-t0)/dt) equal
 * steps of h = (tf-t0)/n; stage times t0 + j*(h/2) for integer j. */
static int integrate_rk4(const eom_t* E, int nv, double t0, double tf, double dt, double* u, int64_t* nsteps) {
    double span = tf - t0;
    int64_t n = (int64_t)ceil(span / dt);
    if (n < 1) n = 1;
    double h = span / (double)n;
    double hh = h * 0.5;
    double h6 = h / 6.0;
    double k1[7], k2[7], k3[7], k4[7], y[7];
    for (int64_t s = 0; s < n; ++s) {
        double t = t0 + (double)(2 * s) * hh;
        double tm = t0 + (double)(2 * s + 1) * hh;
        double tn = t0 + (double)(2 * s + 2) * hh;
        eom_rhs(E, t, u, k1);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + hh * k1[i];
        eom_rhs(E, tm, y, k2);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + hh * k2[i];
        eom_rhs(E, tm, y, k3);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + h * k3[i];
        eom_rhs(E, tn, y, k4);
        for (int i = 0; i < nv; ++i) u[i] = u[i] + h6 * (k1[i] + 2.0 * (k2[i] + k3[i]) + k4[i]);
    }
    *nsteps = n;
    for (int i = 0; i < nv; ++i)
        if (isnan(u[i])) return 1;
    return 0;
}

/* Tsit5 (Tsitouras 2011, Comput. Math. Appl. 62, 770) coefficients. */
static const double TS_c2 = 0.161, TS_c3 = 0.327, TS_c4 = 0.9, TS_c5 = 0.9800255409045097;
static const double TS_a21 = 0.161;
static const double TS_a31 = -0.008480655492356989, TS_a32 = 0.335480655492357;
static const double TS_a41 = 2.897153057105493, TS_a42 = -6.359448489975075, TS_a43 = 4.3622954328695815;
static const double TS_a51 = 5.325864828439257, TS_a52 = -11.748883564062828, TS_a53 = 7.4955393428898365, TS_a54 = -0.09249506636175525;
static const double TS_a61 = 5.86145544294642, TS_a62 = -12.92096931784711, TS_a63 = 8.159367898576159, TS_a64 = -0.071584973281401, TS_a65 = -0.028269050394068383;
static const double TS_a71 = 0.09646076681806523, TS_a72 = 0.01, TS_a73 = 0.4798896504144996, TS_a74 = 1.379008574103742, TS_a75 = -3.290069515436081, TS_a76 = 2.324710524099774;
static const double TS_bt1 = -0.00178001105222577714, TS_bt2 = -0.0008164344596567469, TS_bt3 = 0.007880878010261995,
                    TS_bt4 = -0.1447110071732629, TS_bt5 = 0.5823571654525552, TS_bt6 = -0.45808210592918697, TS_bt7 = 0.015151515151515152;

void oracle_tsit5_tableau(double* c, double* a, double* b, double* btilde) {
    /* c[7], a[7*7] row-major strictly lower, b[7] (= a7 row, FSAL), btilde[7] */
    const double cc[7] = {0, TS_c2, TS_c3, TS_c4, TS_c5, 1, 1};
    double aa[49] = {0};
    aa[1 * 7 + 0] = TS_a21;
    aa[2 * 7 + 0] = TS_a31; aa[2 * 7 + 1] = TS_a32;
    aa[3 * 7 + 0] = TS_a41; aa[3 * 7 + 1] = TS_a42; aa[3 * 7 + 2] = TS_a43;
    aa[4 * 7 + 0] = TS_a51; aa[4 * 7 + 1] = TS_a52; aa[4 * 7 + 2] = TS_a53; aa[4 * 7 + 3] = TS_a54;
    aa[5 * 7 + 0] = TS_a61; aa[5 * 7 + 1] = TS_a62; aa[5 * 7 + 2] = TS_a63; aa[5 * 7 + 3] = TS_a64; aa[5 * 7 + 4] = TS_a65;
    aa[6 * 7 + 0] = TS_a71; aa[6 * 7 + 1] = TS_a72; aa[6 * 7 + 2] = TS_a73; aa[6 * 7 + 3] = TS_a74; aa[6 * 7 + 4] = TS_a75; aa[6 * 7 + 5] = TS_a76;
    const double bt[7] = {TS_bt1, TS_bt2, TS_bt3, TS_bt4, TS_bt5, TS_bt6, TS_bt7};
    for (int i = 0; i < 7; ++i) {
        c[i] = cc[i];
        b[i] = aa[6 * 7 + i];
        btilde[i] = bt[i];
    }
    memcpy(a, aa, sizeof aa);
}

static inline double rms_norm(const double* v, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s / n);
}

/* Adaptive Tsit5 with OrdinaryDiffEq's defaults for an out-of-place
 * explicit method (restated, parity unpinned): Hairer initial dt,
 * PI controller beta1 = 7/50, beta2 = 2/25, qmin = 1/5, qmax = 10,
 * gamma = 9/10, qoldinit = 1e-4, RMS error norm, maxiters, NaN check.
 * Called as solve(..., Tsit5(), adaptive=true, reltol=traj_rtol)
 * (TrajectorySimulation.jl:342). Returns 0 on success, 1 on failure. */
static int integrate_tsit5(const eom_t* E, int nv, double t0, double tf, double rtol, double atol,
                           int64_t maxiters, double* u, int64_t* nsteps) {
    const double qmin = 0.2, qmax = 10.0, gamma = 0.9, beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qoldinit = 1e-4;
    const double dtmax = tf - t0;
    const double dtmin = fmax(fabs(t0), fabs(tf)) * 0x1.0p-52;   /* ~eps(max(|t0|,|tf|)) */
    double k1[7], k2[7], k3[7], k4[7], k5[7], k6[7], k7[7], y[7] = {0}, un[7], sk[7], tmp[7];
    double t = t0;
    /* ode_determine_initdt (Hairer & Wanner, Solving ODEs I, II.4) */
    double dtmin_init = nextafter(fmax(0.0, fabs(t) * 0x1.0p-52), INFINITY);
    double smalldt = fmax(dtmin_init, 1e-6);
    for (int i = 0; i < nv; ++i) sk[i] = atol + fabs(u[i]) * rtol;
    for (int i = 0; i < nv; ++i) tmp[i] = u[i] / sk[i];
    double d0 = rms_norm(tmp, nv);
    eom_rhs(E, t, u, k1);
    for (int i = 0; i < nv; ++i) tmp[i] = k1[i] / sk[i];
    double d1 = rms_norm(tmp, nv);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
    dt0 = fmin(dt0, dtmax);
    for (int i = 0; i < nv; ++i) y[i] = u[i] + dt0 * k1[i];
    eom_rhs(E, t + dt0, y, k2);
    for (int i = 0; i < nv; ++i) tmp[i] = (k2[i] - k1[i]) / sk[i];
    double d2 = rms_norm(tmp, nv) / dt0;
    double mx = fmax(d1, d2);
    double dt1 = (mx <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(10.0, -(2 + log10(mx)) / 5.0);
    double dt = fmax(dtmin_init, fmin(fmin(100 * dt0, dt1), dtmax));

    double qold = qoldinit, q11 = 1.0;
    int64_t iter = 0, accepted = 0;
    while (t < tf) {
        if (tf - t < dt) dt = tf - t;
        if (++iter > maxiters) { *nsteps = accepted; return 1; }
        for (int i = 0; i < nv; ++i) y[i] = u[i] + dt * (TS_a21 * k1[i]);
        eom_rhs(E, t + TS_c2 * dt, y, k2);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + dt * (TS_a31 * k1[i] + TS_a32 * k2[i]);
        eom_rhs(E, t + TS_c3 * dt, y, k3);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + dt * (TS_a41 * k1[i] + TS_a42 * k2[i] + TS_a43 * k3[i]);
        eom_rhs(E, t + TS_c4 * dt, y, k4);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + dt * (TS_a51 * k1[i] + TS_a52 * k2[i] + TS_a53 * k3[i] + TS_a54 * k4[i]);
        eom_rhs(E, t + TS_c5 * dt, y, k5);
        for (int i = 0; i < nv; ++i) y[i] = u[i] + dt * (TS_a61 * k1[i] + TS_a62 * k2[i] + TS_a63 * k3[i] + TS_a64 * k4[i] + TS_a65 * k5[i]);
        eom_rhs(E, t + dt, y, k6);
        for (int i = 0; i < nv; ++i)
            un[i] = u[i] + dt * (TS_a71 * k1[i] + TS_a72 * k2[i] + TS_a73 * k3[i] + TS_a74 * k4[i] + TS_a75 * k5[i] + TS_a76 * k6[i]);
        eom_rhs(E, t + dt, un, k7);
        for (int i = 0; i < nv; ++i) {
            double ut = dt * (TS_bt1 * k1[i] + TS_bt2 * k2[i] + TS_bt3 * k3[i] + TS_bt4 * k4[i] + TS_bt5 * k5[i] + TS_bt6 * k6[i] + TS_bt7 * k7[i]);
            tmp[i] = ut / (atol + fmax(fabs(u[i]), fabs(un[i])) * rtol);
        }
        double EEst = rms_norm(tmp, nv);
        if (isnan(EEst)) { *nsteps = accepted; return 1; }
        double q;
        if (EEst == 0) {
            q = 1.0 / qmax;
        } else {
            q11 = pow(EEst, beta1);
            q = q11 / pow(qold, beta2);
            q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
        }
        if (EEst <= 1.0) {
            t = (tf - t == dt) ? tf : t + dt;
            memcpy(u, un, sizeof(double) * nv);
            memcpy(k1, k7, sizeof(double) * nv);
            ++accepted;
            if (q == 1.0) q = 1.0;  /* qsteady_min = qsteady_max = 1 */
            qold = fmax(EEst, qoldinit);
            dt = fmin(dt / q, dtmax);
        } else {
            dt = dt / fmin(1.0 / qmin, q11 / gamma);
        }
        if (fabs(dt) < dtmin && t < tf) { *nsteps = accepted; return 1; }
    }
    *nsteps = accepted;
    return 0;
}

/* ========================================================================== */
/* Post-processing: Kepler asymptote, SCTS post-pulse phase, binning           */
/*   TrajectorySimulation.jl:356-396 (2D) / :451-496 (3D)                      */
/* ========================================================================== */

/* round(Int, (p + pmax) / (pmax / num * 2)), RoundNearest = ties to even;
 * returns the 1-based Julia index or 0 when not representable. */
int64_t oracle_bin_index(double p, double pmax, int32_t num) {
    if (num <= 1) return 1;
    double q = (p + pmax) / (pmax / num * 2);
    if (!(fabs(q) < 9.0e15)) return 0;
    return (int64_t)nearbyint(q);
}

typedef struct {
    double rec[12];
    int status; /* 0 E>=0, 1 failed, 2 bound */
} post_t;

static void post_process(const target_t* T, int dim, int phase_method, int quirks,
                         const double* u, double t0, post_t* out) {
    double x = u[0], y = u[1], z = 0, px, py, pz = 0, phase = 0;
    if (dim == 2) { px = u[2]; py = u[3]; if (phase_method != ETRAJ_PHASE_CTMC) phase = u[4]; }
    else { z = u[2]; px = u[3]; py = u[4]; pz = u[5]; if (phase_method != ETRAJ_PHASE_CTMC) phase = u[6]; }
    double rec0[7] = {x, y, z, px, py, pz, phase};
    memcpy(out->rec, rec0, sizeof rec0);
    double E = (px * px + py * py + pz * pz) / 2 + target_V(T, x, y, z);
    double Lx = y * pz - z * py, Ly = z * px - x * pz, Lz = x * py - y * px;
    double L2 = Lx * Lx + Ly * Ly + Lz * Lz;
    double Z = (double)T->Z;
    if (E > 0 && phase_method == ETRAJ_PHASE_SCTS) {
        phase += T->Ip * t0;
        double sqrtb = pow(2 * E, -0.5);
        double g = sqrt(1 + 2 * E * L2);
        double rp;
        if (dim == 2) rp = x * px + y * py;
        else rp = (quirks & ETRAJ_QUIRK_SCTS3D_XPY) ? x * py + y * py + z * pz : x * px + y * py + z * pz;
        phase -= Z * sqrtb * (log(g) + asinh(rp / (g * sqrtb)));
    }
    out->rec[7] = E;
    out->rec[11] = phase;
    if (E >= 0) {
        double pinf = sqrt(2 * E);
        double r = sqrt(x * x + y * y + z * z);
        /* a = p x L - Z r/|r| */
        double ax = (py * Lz - pz * Ly) - Z * x / r;
        double ay = (pz * Lx - px * Lz) - Z * y / r;
        double az = (px * Ly - py * Lx) - Z * z / r;
        /* L x a */
        double lax = Ly * az - Lz * ay, lay = Lz * ax - Lx * az, laz = Lx * ay - Ly * ax;
        double f = pinf / (1 + pinf * pinf * L2);
        out->rec[8] = f * (pinf * lax - ax);
        out->rec[9] = f * (pinf * lay - ay);
        out->rec[10] = f * (pinf * laz - az);
        out->status = 0;
    } else {
        out->rec[8] = out->rec[9] = out->rec[10] = NAN;
        out->status = 2;
    }
}

/* ========================================================================== */
/* ADK sampler: ADKSampler.jl:19-396 (atoms, l = 0)                            */
/* ========================================================================== */
typedef struct {
    const etraj_sampler_desc* sd;
    laser_t L;
    target_t T;
    int prefix;        /* resolved ETRAJ_PREFIX_* bits */
    double* tsamp;     /* t_samples, length sample_t_num */
    double tstep;      /* step(t_samples) = (max-min)/length  (ADKSampler.jl:267) */
    double c_pre;      /* c (ADKSampler.jl:233) */
    double Ip;         /* the sampler's Ip: IonPotential(mol, ridx) of a molecule, else the target's */
} sampler_t;

/* i-th point (0-based) of range(lo, hi; length=n): (lo*(n-1-i) + hi*i)/(n-1),
 * the correctly rounded rational Julia's range produces for integer endpoints. */
double oracle_grid_point(double lo, double hi, int n, int i) {
    if (n <= 1) return lo;
    return (lo * (double)(n - 1 - i) + hi * (double)i) / (double)(n - 1);
}

static int cmp_double(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* t_samples: range(lo,hi;length=n) (ADKSampler.jl:136), sorted seeded
 * uniforms (:141, MONTECARLO), or one jittered point per stratum (STRATIFIED). */
int oracle_batch_times(const etraj_sampler_desc* sd, double* t) {
    int n = sd->sample_t_num;
    if (n <= 0) return ETRAJ_EINVAL;
    double lo = sd->sample_t_lo, hi = sd->sample_t_hi;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        for (int i = 0; i < n; ++i) t[i] = oracle_grid_point(lo, hi, n, i);
    } else if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
        for (int i = 0; i < n; ++i) {
            uint64_t c[4] = {(uint64_t)i, 0, 0, TAG_T0}, k[2] = {sd->seed, 0}, o[4];
            oracle_philox4x64_10(c, k, o);
            t[i] = u01_open_right(o[0]) * (hi - lo) + lo;
        }
        qsort(t, n, sizeof(double), cmp_double);
    } else if (sd->sampling == ETRAJ_SAMPLING_STRATIFIED) {
        double w = (hi - lo) / (double)n;
        for (int i = 0; i < n; ++i) {
            uint64_t c[4] = {(uint64_t)i, 0, 0, TAG_T0}, k[2] = {sd->seed, 0}, o[4];
            oracle_philox4x64_10(c, k, o);
            t[i] = lo + ((double)i + u01_open_right(o[0])) * w;
        }
    } else {
        return ETRAJ_EINVAL;
    }
    return ETRAJ_OK;
}

int64_t oracle_batch_max_size(const etraj_sampler_desc* sd) {
    if (sd->sampling == ETRAJ_SAMPLING_STEP)
        return (int64_t)sd->ss_kd_num * (sd->dimension == 3 ? (int64_t)sd->ss_kz_num : 1);
    return sd->mc_kt_num;
}

static int resolve_prefix(int bits, const laser_t* L) {
    if ((bits & ETRAJ_PREFIX_PRE) && (bits & ETRAJ_PREFIX_PRECC)) return -1; /* ADKSampler.jl:66-68 */
    if ((bits & ETRAJ_PREFIX_PRECC) && L->n != 1) {                          /* :72-77, Q13 */
        bits &= ~ETRAJ_PREFIX_PRECC;
        bits |= ETRAJ_PREFIX_PRE;
    }
    return bits;
}

static int sampler_init(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd, sampler_t* S) {
    memset(S, 0, sizeof *S);
    S->sd = sd;
    if (laser_init(ld, &S->L) || target_init(td, &S->T)) return ETRAJ_EINVAL;
    if (sd->init_method < ETRAJ_INIT_ADK || sd->init_method > ETRAJ_INIT_WFAT) return ETRAJ_EUNSUPPORTED;
    if (sd->init_method != ETRAJ_INIT_WFAT && sd->asymp_coeff && !(sd->asymp_lmax >= 0 && sd->asymp_lmax <= 15))
        return ETRAJ_EINVAL;
    if (sd->init_method == ETRAJ_INIT_WFAT && (sd->wfat_na < 1 || sd->wfat_nb < 1 || !sd->wfat_coeff)) return ETRAJ_EINVAL;
    if (sd->dimension != 2 && sd->dimension != 3) return ETRAJ_EINVAL;
    if (sd->phase_method < 0 || sd->phase_method > 2) return ETRAJ_EINVAL;
    if (sd->sample_t_num <= 0 || sd->cutoff_limit < 0) return ETRAJ_EINVAL;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        if (sd->ss_kd_num <= 0) return ETRAJ_EINVAL;
        if (sd->dimension == 2 && !(sd->ss_kd_max > 0)) return ETRAJ_EINVAL;
        if (sd->dimension == 3 && (sd->ss_kz_num <= 0 || !(sd->ss_kd_max > 0 || sd->ss_kz_max > 0))) return ETRAJ_EINVAL;
    } else {
        if (sd->mc_kt_num <= 0 || !(sd->sample_t_lo < sd->sample_t_hi)) return ETRAJ_EINVAL;
    }
    int p = resolve_prefix(sd->rate_prefix, &S->L);
    if (p < 0) return ETRAJ_EINVAL;
    S->prefix = p;
    S->tsamp = (double*)malloc(sizeof(double) * sd->sample_t_num);
    if (!S->tsamp) return ETRAJ_ENOMEM;
    oracle_batch_times(sd, S->tsamp);
    double tmin = S->tsamp[0], tmax = S->tsamp[0];
    for (int i = 1; i < sd->sample_t_num; ++i) {
        tmin = fmin(tmin, S->tsamp[i]);
        tmax = fmax(tmax, S->tsamp[i]);
    }
    S->tstep = (tmax - tmin) / sd->sample_t_num;
    S->Ip = sd->orbit_ip > 0 ? sd->orbit_ip : S->T.Ip;
    double kappa = sqrt(2 * S->Ip), n = S->T.Z / kappa;
    S->c_pre = pow(2, n / 2 + 1) * pow(kappa, 2 * n + 0.5) * tgamma(n / 2 + 1);
    return ETRAJ_OK;
}

static void sampler_free(sampler_t* S) { free(S->tsamp); S->tsamp = NULL; }

static inline double gauss_pdf(double x, double s) { return exp(-0.5 * (x / s) * (x / s)) / (s * sqrt(2 * M_PI)); }

/* ---- prefactor beyond l = 0: sum over (l, m, m') of C_lm D^l_{m' m}(alpha, beta, gamma)
 * Y_{l m'}(k') (ADKSampler.jl:188-265, SPASampler.jl:204-278, SPANESampler.jl:208-266) ---- */
typedef struct {
    int on;               /* 0: the scalar C Y_00 of an atom with l = 0 */
    int L, atom_l, atom_m;
    const double* C;      /* molecule: (re, im) of C_lm at [l, m+l], column-major (L+1) x (2L+1) */
    double atomC;         /* atom: its asymptotic coefficient at (atom_l, atom_m) */
    double alpha, beta, gamma, kappa, ex, ey;
} prefac_t;

/* Per batch: the Euler angles FF -> target frame (obtain_FF_MF_Euler of MolRotation(mol), or of
 * (phi, theta, 0) of an atom's quantization axis, ADKSampler.jl:198-204) and F/|F| at tr. */
static void prefac_init(const sampler_t* S, double Fx, double Fy, prefac_t* P) {
    const etraj_sampler_desc* sd = S->sd;
    memset(P, 0, sizeof *P);
    double rot[3];
    if (sd->asymp_coeff) {
        P->L = sd->asymp_lmax;
        P->C = sd->asymp_coeff;
        for (int k = 0; k < 3; ++k) rot[k] = sd->mol_rot[k];
    } else if (S->T.l != 0) {
        P->L = P->atom_l = S->T.l;
        P->atom_m = S->T.m;
        P->atomC = S->T.C;
        rot[0] = S->T.qph; rot[1] = S->T.qth; rot[2] = 0.0;
    } else {
        return;
    }
    P->on = 1;
    double eul[3];
    oracle_ff_mf_euler(rot, Fx, Fy, eul);
    P->alpha = eul[0]; P->beta = eul[1]; P->gamma = eul[2];
    P->kappa = sqrt(2 * S->Ip);
    double F = hypot(Fx, Fy);
    P->ex = Fx / F;
    P->ey = Fy / F;
}

/* Y_{lm}(x, y, z) of gen_sph_harm_funcs (shared_methods.jl:15-24) at complex (x, y, z):
 * (-1)^m sqrt((2l+1)/4pi (l-m)!/(l+m)!) / (-i kappa)^l (l+m)! sum_k (-1)^k /
 * (2^(2k+m) (k+m)! k! (l-m-2k)!) (x+iy)^(k+m) (x-iy)^k z^(l-m-2k). */
/* z^n for an integer n >= 0, as Julia's ^(::Complex, ::Integer) (z^0 = 1, also for z = 0) */
static double complex cipow(double complex z, int n) {
    double complex r = 1;
    for (int i = 0; i < n; ++i) r *= z;
    return r;
}

static double complex sph_harm_c(int l, int m, double kappa, double complex x, double complex y, double complex z) {
    double complex s = 0;
    for (int k = (m < 0 ? -m : 0); k <= (l - m) / 2; ++k) {
        double c = ((k & 1) ? -1.0 : 1.0) /
                   (pow(2.0, 2 * k + m) * tgamma(k + m + 1.0) * tgamma(k + 1.0) * tgamma(l - m - 2 * k + 1.0));
        s += c * cipow(x + I * y, k + m) * cipow(x - I * y, k) * cipow(z, l - m - 2 * k);
    }
    double nrm = ((m & 1) ? -1.0 : 1.0) * sqrt((2 * l + 1) / (4 * M_PI) * tgamma(l - m + 1.0) / tgamma(l + m + 1.0));
    return nrm / cipow(-I * kappa, l) * tgamma(l + m + 1.0) * s;
}

void oracle_sph_harm_c(int l, int m, double kappa, const double xyz[6], double out[2]) {
    double complex v = sph_harm_c(l, m, kappa, xyz[0] + I * xyz[1], xyz[2] + I * xyz[3], xyz[4] + I * xyz[5]);
    out[0] = creal(v);
    out[1] = cimag(v);
}

/* D^l_{m' m}(alpha, beta, gamma) = e^{-i m' alpha} d^l_{m' m}(beta) e^{-i m gamma} (WignerD.jl's wignerDjmn) */
static double complex wigner_D(int l, int mp, int m, double a, double b, double g) {
    return cexp(-I * (double)mp * a) * oracle_wigner_d(l, mp, m, b) * cexp(-I * (double)m * g);
}

/* pre_numerator summed over its (l, m, m') list. k' are the FF components of kts
 * (obtain_xyz_FF_LF: x' = z, y' = (Fy, -Fx, 0)/F, z' = (Fx, Fy, 0)/F) taken with
 * LinearAlgebra.dot, which conjugates its first argument: k'_i = conj(kts) . e_i. */
static double complex prefac_numer(const prefac_t* P, const double complex kts[3]) {
    double complex c0 = conj(kts[0]), c1 = conj(kts[1]), c2 = conj(kts[2]);
    double complex nx = c2, ny = c0 * P->ey - c1 * P->ex, nz = c0 * P->ex + c1 * P->ey;
    double complex sum = 0;
    for (int l = (P->C ? 0 : P->atom_l); l <= P->L; ++l)
        for (int m = -l; m <= l; ++m) {
            double complex C;
            if (P->C) {
                const double* c = P->C + 2 * (l + (int64_t)(P->L + 1) * (m + l));
                C = c[0] + I * c[1];
                if (C == 0) continue;                              /* :258-259 */
            } else {
                if (m != P->atom_m) continue;
                C = P->atomC;
            }
            for (int mp = -l; mp <= l; ++mp)
                sum += C * wigner_D(l, mp, m, P->alpha, P->beta, P->gamma) * sph_harm_c(l, mp, P->kappa, nx, ny, nz);
        }
    return sum;
}

/* gen_electron_batch(sp::ADKSampler, batch_id) (ADKSampler.jl:167-396). For an
 * atom with l = 0 the Wigner-D factor is 1 and Y_00 = 1/sqrt(4 pi), so the
 * prefactor is real and positive; otherwise the sum of prefac_numer. */
static int64_t adk_gen_batch(const sampler_t* S, int b /*0-based*/, double* out, int64_t cap) {
    const etraj_sampler_desc* sd = S->sd;
    const int dim = sd->dimension, ph = sd->phase_method;
    const int rows = (dim == 3) ? (ph == ETRAJ_PHASE_CTMC ? 8 : 9) : (ph == ETRAJ_PHASE_CTMC ? 6 : 7);
    const double tr = S->tsamp[b];
    double Fxtr, Fytr;
    laser_F(&S->L, tr, &Fxtr, &Fytr);
    const double Ftr = hypot(Fxtr, Fytr);
    const double phi = atan2(-Fytr, -Fxtr);                       /* :173 */
    const double Ip = S->Ip, Z = (double)S->T.Z;
    double gam = 0.0;                                              /* :183-187 */
    if (S->L.n == 1) gam = S->L.c[0].w * sqrt(2 * Ip) / (S->L.c[0].F0 * mono_envelope(&S->L.c[0], tr));
    const double kappa = sqrt(2 * Ip), n = Z / kappa;
    const double e0 = 2.71828182845904523;
    const double c_cc = pow(2, 3 * n / 2 + 1) * pow(kappa, 5 * n + 0.5) * pow(Ftr, -n) * pow(1 + 2 * gam / e0, -n); /* :235 */
    const double Y00 = sqrt(1.0 / (4 * M_PI));                     /* shared_methods.jl:21, l = m = 0 */
    const double numer = S->T.C * 1.0 * Y00;                       /* pre_numerator :255, D^0_00 = 1 */
    prefac_t PF;
    prefac_init(S, Fxtr, Fytr, &PF);                               /* l > 0 or a molecule: :188-265 */
    const double jac = Ftr;                                        /* :266 */
    double dkdt;                                                   /* :268-276 */
    int64_t nk;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        double kd_step = sd->ss_kd_num > 1 ? (2 * fabs(sd->ss_kd_max)) / sd->ss_kd_num : 1.0;
        double kz_step = (dim == 3 && sd->ss_kz_num > 1) ? (2 * fabs(sd->ss_kz_max)) / sd->ss_kz_num : 1.0;
        dkdt = S->tstep * kd_step * kz_step;
        nk = (int64_t)sd->ss_kd_num * (dim == 3 ? sd->ss_kz_num : 1);
    } else if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
        double kd_span = sd->mc_kd_max > 0 ? 2 * sd->mc_kd_max : 1.0;
        double kz_span = sd->mc_kz_max > 0 ? 2 * sd->mc_kz_max : 1.0;   /* no dimension check in :274 */
        dkdt = S->tstep * kd_span * kz_span / sd->mc_kt_num;
        nk = sd->mc_kt_num;
    } else {
        dkdt = (sd->sample_t_hi - sd->sample_t_lo) / sd->sample_t_num / sd->mc_kt_num; /* divided by pdf below */
        nk = sd->mc_kt_num;
    }
    int64_t count = 0;
    const int kzn = (dim == 3) ? sd->ss_kz_num : 1;
    for (int64_t j = 0; j < nk; ++j) {
        double kd0, kz0 = 0, w = dkdt;
        if (sd->sampling == ETRAJ_SAMPLING_STEP) {
            int64_t ikd = j / kzn, ikz = j % kzn;     /* loop order ikd outer, ikz inner (:318-319) */
            double kdm = fabs(sd->ss_kd_max), kzm = fabs(sd->ss_kz_max);
            kd0 = oracle_grid_point(-kdm, kdm, sd->ss_kd_num, (int)ikd);
            if (dim == 3) kz0 = oracle_grid_point(-kzm, kzm, kzn, (int)ikz);
        } else {
            uint64_t c[4] = {(uint64_t)j, (uint64_t)b, 0, TAG_K}, k[2] = {sd->seed, 0}, o[4];
            oracle_philox4x64_10(c, k, o);
            if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
                /* gen_rand_pt_1d / gen_rand_pt_2dsq (shared_methods.jl:4-12) */
                kd0 = (u01_open_right(o[0]) - 0.5) * 2 * sd->mc_kd_max;
                if (dim == 3) {
                    kz0 = (u01_open_right(o[1]) - 0.5) * 2 * sd->mc_kz_max;
                    if (sd->mc_kd_max == 0) kd0 = 0.0;
                    if (sd->mc_kz_max == 0) kz0 = 0.0;
                }
            } else {
                double r = sqrt(-2.0 * log(u01_open_left(o[0])));
                double a = 2 * M_PI * u01_open_right(o[1]);
                kd0 = sd->kd_sigma * (r * cos(a));
                w = w / gauss_pdf(kd0, sd->kd_sigma);
                if (dim == 3) {
                    kz0 = sd->kz_sigma * (r * sin(a));
                    w = w / gauss_pdf(kz0, sd->kz_sigma);
                }
            }
        }
        if (fabs(kd0) < 1e-4) continue;                              /* :312,:321 (Q5) */
        double kx0 = kd0 * -sin(phi), ky0 = kd0 * cos(phi);           /* :324-325 */
        double r0 = (Ip + (kd0 * kd0 + kz0 * kz0) / 2) / Ftr;         /* :326 */
        double x0 = r0 * cos(phi), y0 = r0 * sin(phi);
        double complex amp = sqrt(w) * exp(-pow(kd0 * kd0 + kz0 * kz0 + 2 * Ip, 1.5) / (3 * Ftr));  /* ADKAmpExp :198 */
        if (S->prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
            /* pre/pre_cc take (kx,ky,kz) and ts = tr + i ti, ti = sqrt(kappa^2+kd^2+kz^2)/F (:238-265, :283-291) */
            double cpre = (S->prefix & ETRAJ_PREFIX_PRE) ? S->c_pre : c_cc;
            double complex num = numer;
            if (PF.on) {
                double ti = sqrt(kappa * kappa + kd0 * kd0 + kz0 * kz0) / Ftr;   /* :237 */
                double complex kts[3] = {kx0 - I * ti * Fxtr, ky0 - I * ti * Fytr, kz0};   /* k_ts :238 */
                num = prefac_numer(&PF, kts);
            }
            amp *= cpre * num / pow((kx0 * kx0 + ky0 * ky0 + kz0 * kz0 + 2 * Ip) * (Ftr * Ftr), (n + 1) / 4);
        }
        if (S->prefix & ETRAJ_PREFIX_JAC) amp *= sqrt(jac);
        double rate = creal(amp) * creal(amp) + cimag(amp) * cimag(amp);   /* abs2 :331 */
        double ph0 = carg(amp);                                              /* angle(amp) :337 */
        if (isnan(rate) || rate < sd->cutoff_limit) continue;
        if (count >= cap) return -1;
        double* col = out + count * rows;
        if (dim == 3) {
            double v[8] = {x0, y0, 0.0, kx0, ky0, kz0, tr, rate};
            memcpy(col, v, sizeof v);
            if (ph != ETRAJ_PHASE_CTMC) col[8] = ph0;
        } else {
            double v[6] = {x0, y0, kx0, ky0, tr, rate};
            memcpy(col, v, sizeof v);
            if (ph != ETRAJ_PHASE_CTMC) col[6] = ph0;
        }
        ++count;
    }
    return count;
}

/* ---- WFAT (WFATSampler.jl:132-249), channel (0,0), tabulated G00 ---------- */

/* params(RotZYZ(R)) of Rotations.jl for a proper rotation R (row-major):
 * R = Rz(a) Ry(b) Rz(c) with b in [0, pi]. */
static void zyz_params(const double R[3][3], double* a, double* b, double* c) {
    double t1 = atan2(R[1][2], R[0][2]);
    double ct = cos(t1), st = sin(t1);
    *a = t1;
    *b = atan2(ct * R[0][2] + st * R[1][2], R[2][2]);
    *c = atan2(-st * R[0][0] + ct * R[1][0], -st * R[0][1] + ct * R[1][1]);
}

static void rot_zyz(double a, double b, double c, double M[3][3]) {
    double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cc = cos(c), sc = sin(c);
    double Rz1[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    double Rz2[3][3] = {{cc, -sc, 0}, {sc, cc, 0}, {0, 0, 1}};
    double T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0;
            for (int k = 0; k < 3; ++k) T[i][j] += Rz1[i][k] * Ry[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            M[i][j] = 0;
            for (int k = 0; k < 3; ++k) M[i][j] += T[i][k] * Rz2[k][j];
        }
}

/* obtain_FF_MF_Euler (shared_methods.jl:31-38): R = MF^-1 FF with
 * FF = [0 Fy/F Fx/F; 0 -Fx/F Fy/F; 1 0 0] and MF = RotZYZ(mol_rot). */
void oracle_ff_mf_euler(const double mol_rot[3], double Fx, double Fy, double out[3]) {
    double F = hypot(Fx, Fy);
    double FF[3][3] = {{0, Fy / F, Fx / F}, {0, -Fx / F, Fy / F}, {1, 0, 0}};
    double MF[3][3], R[3][3];
    rot_zyz(mol_rot[0], mol_rot[1], mol_rot[2], MF);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += MF[k][i] * FF[k][j];   /* MF^T FF */
        }
    zyz_params(R, &out[0], &out[1], &out[2]);
}

/* Wigner small-d d^l_{m m'}(beta) = <l m| exp(-i beta J_y) |l m'> (Condon-Shortley
 * J_y), by Wigner's formula summed over s:
 * (-1)^(m-m'+s) sqrt((l+m)!(l-m)!(l+m')!(l-m')!) / ((l+m'-s)! s! (m-m'+s)! (l-m-s)!)
 *   cos(beta/2)^(2l+m'-m-2s) sin(beta/2)^(m-m'+2s). The convention of
 * WignerD.jl's wignerdjmn, which GenericMolecule.jl:257 calls. */
double oracle_wigner_d(int l, int m, int mp, double beta) {
    double c = cos(beta / 2), sn = sin(beta / 2), sum = 0;
    double pre = sqrt(tgamma(l + m + 1.0) * tgamma(l - m + 1.0) * tgamma(l + mp + 1.0) * tgamma(l - mp + 1.0));
    int s0 = mp - m > 0 ? mp - m : 0, s1 = (l + mp < l - m) ? l + mp : l - m;
    for (int k = s0; k <= s1; ++k) {
        double den = tgamma(l + mp - k + 1.0) * tgamma(k + 1.0) * tgamma(m - mp + k + 1.0) * tgamma(l - m - k + 1.0);
        double t = pre / den * pow(c, 2 * l + mp - m - 2 * k) * pow(sn, m - mp + 2 * k);
        sum += ((m - mp + k) & 1) ? -t : t;
    }
    return sum;
}

/* The structure factor G_{nxi m}(beta, gamma):
 * ETRAJ_WFAT_TABLE: channel (0, 0) only, sum_{a,b} c_ab P_a(cos beta) e^{-i b gamma};
 * ETRAJ_WFAT_INTDATA: MolWFATStructureFactor_G (GenericMolecule.jl:236-260),
 *   sum_{l=|m|..L} sum_{m'=-l..l} I[nxi, m, l, m'] d^l_{m m'}(beta) e^{-i m' gamma}
 *   times exp(-kappa mu_z), mu_z = (RotZYZ(gamma, beta, 0) mu)_z
 *   = -sin(beta) mu_x + cos(beta) mu_z; I is Julia's column-major array. */
static double complex wfat_G(const etraj_sampler_desc* sd, double kappa, double beta, double gamma, int nxi, int m) {
    double complex g = 0;
    if (sd->wfat_form == ETRAJ_WFAT_INTDATA) {
        const int D0 = sd->wfat_dims[0], D1 = sd->wfat_dims[1], D2 = sd->wfat_dims[2];
        const int mM = (D1 - 1) / 2, L = D2 - 1;
        if (nxi < 0 || nxi >= D0 || abs(m) > mM) return 0;
        for (int l = abs(m); l <= L; ++l)
            for (int mp = -l; mp <= l; ++mp) {
                const double* c = sd->wfat_coeff + 2 * (nxi + (int64_t)D0 * ((m + mM) + (int64_t)D1 * (l + (int64_t)D2 * (mp + l))));
                g += (c[0] + I * c[1]) * oracle_wigner_d(l, m, mp, beta) * cexp(-I * (double)mp * gamma);
            }
        double muz = -sin(beta) * sd->wfat_mu[0] + cos(beta) * sd->wfat_mu[2];
        return g * exp(-kappa * muz);
    }
    if (nxi != 0 || m != 0) return 0;
    double x = cos(beta);
    double pm1 = 0, p = 1;   /* P_{a-1}, P_a */
    for (int a = 0; a < sd->wfat_na; ++a) {
        if (a == 1) { pm1 = 1; p = x; }
        else if (a > 1) { double pn = ((2 * a - 1) * x * p - (a - 1) * pm1) / a; pm1 = p; p = pn; }
        for (int b = 0; b < sd->wfat_nb; ++b) {
            const double* c = sd->wfat_coeff + 2 * (a * sd->wfat_nb + b);
            g += (c[0] + I * c[1]) * p * cexp(-I * (double)b * gamma);
        }
    }
    return g;
}

void oracle_wfat_G(const etraj_sampler_desc* sd, double kappa, double beta, double gamma, int nxi, int m,
                   double out[2]) {
    double complex g = wfat_G(sd, kappa, beta, gamma, nxi, m);
    out[0] = creal(g);
    out[1] = cimag(g);
}

static int64_t wfat_gen_batch(const sampler_t* S, int b, double* out, int64_t cap) {
    const etraj_sampler_desc* sd = S->sd;
    const int dim = sd->dimension, ph = sd->phase_method;
    const int rows = (dim == 3) ? (ph == ETRAJ_PHASE_CTMC ? 8 : 9) : (ph == ETRAJ_PHASE_CTMC ? 6 : 7);
    const double tr = S->tsamp[b];
    double Fxtr, Fytr;
    laser_F(&S->L, tr, &Fxtr, &Fytr);
    const double Ftr = hypot(Fxtr, Fytr);
    const double phi = atan2(-Fytr, -Fxtr);                          /* :137 */
    const double Ip = S->Ip, Z = (double)S->T.Z;   /* :139 */
    double eul[3];
    oracle_ff_mf_euler(sd->mol_rot, Fxtr, Fytr, eul);                 /* :143-145 */
    const double kappa = sqrt(2 * Ip), n = Z / kappa;
    /* G_data[nxi, m] for every channel (:150-155); the table form has (0, 0) only */
    const int D0 = sd->wfat_form == ETRAJ_WFAT_INTDATA ? sd->wfat_dims[0] : 1;
    const int D1 = sd->wfat_form == ETRAJ_WFAT_INTDATA ? sd->wfat_dims[1] : 1, mM = (D1 - 1) / 2;
    double* G2 = malloc(sizeof(double) * D0 * D1);
    for (int nxi = 0; nxi < D0; ++nxi)
        for (int m = -mM; m <= mM; ++m) {
            double complex G = wfat_G(sd, kappa, eul[1], eul[2], nxi, m);
            G2[nxi * D1 + m + mM] = creal(G) * creal(G) + cimag(G) * cimag(G);
        }
    double dkdt;                                                      /* :159-168 */
    int64_t nk;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        double kd_step = sd->ss_kd_num > 1 ? (2 * fabs(sd->ss_kd_max)) / sd->ss_kd_num : 1.0;
        double kz_step = (dim == 3 && sd->ss_kz_num > 1) ? (2 * fabs(sd->ss_kz_max)) / sd->ss_kz_num : 1.0;
        dkdt = S->tstep * kd_step * kz_step;
        nk = (int64_t)sd->ss_kd_num * (dim == 3 ? sd->ss_kz_num : 1);
    } else if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
        double kd_span = sd->mc_kd_max > 0 ? 2 * sd->mc_kd_max : 1.0;
        double kz_span = sd->mc_kz_max > 0 ? 2 * sd->mc_kz_max : 1.0;
        dkdt = S->tstep * kd_span * kz_span / sd->mc_kt_num;
        nk = sd->mc_kt_num;
    } else {
        dkdt = (sd->sample_t_hi - sd->sample_t_lo) / sd->sample_t_num / sd->mc_kt_num;
        nk = sd->mc_kt_num;
    }
    const int kzn = (dim == 3) ? sd->ss_kz_num : 1;
    int64_t count = 0;
    for (int64_t j = 0; j < nk; ++j) {
        double kd0, kz0 = 0, w = dkdt;
        if (sd->sampling == ETRAJ_SAMPLING_STEP) {
            int64_t ikd = j / kzn, ikz = j % kzn;
            kd0 = oracle_grid_point(-fabs(sd->ss_kd_max), fabs(sd->ss_kd_max), sd->ss_kd_num, (int)ikd);
            if (dim == 3) kz0 = oracle_grid_point(-fabs(sd->ss_kz_max), fabs(sd->ss_kz_max), kzn, (int)ikz);
        } else {
            uint64_t c[4] = {(uint64_t)j, (uint64_t)b, 0, TAG_K}, k[2] = {sd->seed, 0}, o[4];
            oracle_philox4x64_10(c, k, o);
            if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
                kd0 = (u01_open_right(o[0]) - 0.5) * 2 * sd->mc_kd_max;
                if (dim == 3) {
                    kz0 = (u01_open_right(o[1]) - 0.5) * 2 * sd->mc_kz_max;
                    if (sd->mc_kd_max == 0) kd0 = 0.0;
                    if (sd->mc_kz_max == 0) kz0 = 0.0;
                }
            } else {
                double r = sqrt(-2.0 * log(u01_open_left(o[0])));
                double a = 2 * M_PI * u01_open_right(o[1]);
                kd0 = sd->kd_sigma * (r * cos(a));
                w = w / gauss_pdf(kd0, sd->kd_sigma);
                if (dim == 3) {
                    kz0 = sd->kz_sigma * (r * sin(a));
                    w = w / gauss_pdf(kz0, sd->kz_sigma);
                }
            }
        }
        if (kd0 * kd0 + kz0 * kz0 < 1e-4 * 1e-4) continue;           /* :188 */
        double kx0 = kd0 * -sin(phi), ky0 = kd0 * cos(phi);
        double r0 = (Ip + (kd0 * kd0 + kz0 * kz0) / 2) / Ftr;
        /* rate = sum_{nxi, m} |G|^2 W(nxi, |m|, kd, kz) dkdt (:157, :169) */
        double rate = 0;
        for (int nxi = 0; nxi < D0; ++nxi)
            for (int m = -mM; m <= mM; ++m) {
                int am = abs(m);
                double W = (pow(kappa, am + 2) / 2 / pow(Ftr, am + 1) / tgamma(am + 1.0)) *
                           pow(4 * kappa * kappa / Ftr, 2 * n - 2 * nxi - am - 1) *
                           pow(kd0 * kd0 + kz0 * kz0, am) *
                           exp(-2 * pow(kappa * kappa + kd0 * kd0 + kz0 * kz0, 1.5) / (3 * Ftr));
                rate += G2[nxi * D1 + m + mM] * W;
            }
        rate *= w;
        if (isnan(rate) || rate < sd->cutoff_limit) continue;
        if (count >= cap) {
            free(G2);
            return -1;
        }
        double* col = out + count * rows;
        if (dim == 3) {
            double v[8] = {r0 * cos(phi), r0 * sin(phi), 0.0, kx0, ky0, kz0, tr, rate};
            memcpy(col, v, sizeof v);
            if (ph != ETRAJ_PHASE_CTMC) col[8] = 0.0;
        } else {
            double v[6] = {r0 * cos(phi), r0 * sin(phi), kx0, ky0, tr, rate};
            memcpy(col, v, sizeof v);
            if (ph != ETRAJ_PHASE_CTMC) col[6] = 0.0;
        }
        ++count;
    }
    free(G2);
    return count;
}

/* ---- lasers at complex time and dF/dt (SPASampler.jl, SPANESampler.jl) ---- */

/* Complex-time LaserAx/Ay/Fx/Fy: the closed forms of Cos2Laser.jl:97-152
 * (and Cos4, Gaussian) with complex t; the cut-off window takes real(t)
 * (Cos2Laser.jl:93), so it is constant along t = tr + i tau. */
static void mono_FA_c(const mono_t* m, double complex t, double complex* Fx, double complex* Fy,
                      double complex* Ax, double complex* Ay) {
    t -= m->dt_shift;
    const double w = m->w, N = m->N, e = m->ellip, az = m->azi;
    const double complex cw = ccos(w * t + m->cep), sw = csin(w * t + m->cep);
    double complex g, gA, a, b;
    if (m->kind == ETRAJ_LASER_GAUSSIAN) {
        const double s2 = m->sigma * m->sigma;
        gA = m->A0 * cexp(-t * t / 2 / s2);
        g = gA;
        a = t / s2 * cw + w * sw;
        b = t / s2 * sw - w * cw;
    } else {
        const double win = cosn_window(m, creal(t));
        const double complex ch = ccos(w * t / (2 * N)), sh = csin(w * t / (2 * N));
        const double cN = (m->kind == ETRAJ_LASER_COS2) ? 1 / N : 2 / N;
        const double complex env = (m->kind == ETRAJ_LASER_COS2) ? ch * ch : (ch * ch) * (ch * ch);
        gA = m->A0 * env * win;
        g = m->F0 * ((m->kind == ETRAJ_LASER_COS2) ? ch : ch * ch * ch) * win;
        a = ch * sw + cN * sh * cw;
        b = -ch * cw + cN * sh * sw;
    }
    if (az == 0) {
        *Fx = g * a;
        *Fy = g * b * e;
        *Ax = gA * cw;
        *Ay = gA * sw * e;
    } else {
        *Fx = g * (a * cos(az) + b * e * sin(az));
        *Fy = g * (a * -sin(az) + b * e * cos(az));
        *Ax = gA * (cw * cos(az) + sw * e * sin(az));
        *Ay = gA * (cw * -sin(az) + sw * e * cos(az));
    }
}

static void laser_FA_c(const laser_t* L, double complex t, double complex* Fx, double complex* Fy,
                       double complex* Ax, double complex* Ay) {
    mono_FA_c(&L->c[0], t, Fx, Fy, Ax, Ay);
    if (L->n == 2) {
        double complex f2x, f2y, a2x, a2y;
        mono_FA_c(&L->c[1], t - L->delay, &f2x, &f2y, &a2x, &a2y);
        *Fx += f2x; *Fy += f2y; *Ax += a2x; *Ay += a2y;
    }
}

int oracle_laser_complex(const etraj_laser_desc* d, double t_re, double t_im, double out[8]) {
    laser_t L;
    if (laser_init(d, &L)) return ETRAJ_EINVAL;
    double complex Fx, Fy, Ax, Ay;
    laser_FA_c(&L, t_re + I * t_im, &Fx, &Fy, &Ax, &Ay);
    double v[8] = {creal(Fx), cimag(Fx), creal(Fy), cimag(Fy), creal(Ax), cimag(Ax), creal(Ay), cimag(Ay)};
    memcpy(out, v, sizeof v);
    return ETRAJ_OK;
}

/* Forward-mode derivative (value, d/dt) of the field: what ForwardDiff's
 * derivative(Fx, tr) gives in SPANESampler.jl:164-165 (the tanh window is
 * differentiated too). */
typedef struct { double v, d; } dual_t;
static inline dual_t dl(double v, double d) { dual_t r = {v, d}; return r; }
static inline dual_t dmul(dual_t a, dual_t b) { return dl(a.v * b.v, a.d * b.v + a.v * b.d); }
static inline dual_t dadd(dual_t a, dual_t b) { return dl(a.v + b.v, a.d + b.d); }
static inline dual_t dscale(double s, dual_t a) { return dl(s * a.v, s * a.d); }
static inline dual_t dcos(dual_t a) { return dl(cos(a.v), -sin(a.v) * a.d); }
static inline dual_t dsin(dual_t a) { return dl(sin(a.v), cos(a.v) * a.d); }
static inline dual_t dexp(dual_t a) { double e = exp(a.v); return dl(e, e * a.d); }
static inline dual_t dtanh(dual_t a) { double th = tanh(a.v); return dl(th, (1 - th * th) * a.d); }

static void mono_F_dual(const mono_t* m, double t0, double* Fx, double* Fy, double* dFx, double* dFy) {
    const dual_t t = dl(t0 - m->dt_shift, 1.0);
    const double w = m->w, N = m->N, e = m->ellip, az = m->azi;
    const dual_t wt = dadd(dscale(w, t), dl(m->cep, 0));
    const dual_t cw = dcos(wt), sw = dsin(wt);
    dual_t g, a, b;
    if (m->kind == ETRAJ_LASER_GAUSSIAN) {
        const double s2 = m->sigma * m->sigma;
        g = dscale(m->A0, dexp(dscale(-1.0 / (2 * s2), dmul(t, t))));
        a = dadd(dmul(dscale(1 / s2, t), cw), dscale(w, sw));
        b = dadd(dmul(dscale(1 / s2, t), sw), dscale(-w, cw));
    } else {
        const double hw = N * M_PI / w;
        const dual_t w1 = dadd(dl(0.5, 0), dscale(0.5, dtanh(dscale(m->k, dadd(t, dl(hw, 0))))));
        const dual_t w2 = dadd(dl(0.5, 0), dscale(0.5, dtanh(dscale(-m->k, dadd(t, dl(-hw, 0))))));
        const dual_t half = dscale(w / (2 * N), t);
        const dual_t ch = dcos(half), sh = dsin(half);
        const double cN = (m->kind == ETRAJ_LASER_COS2) ? 1 / N : 2 / N;
        const dual_t env = (m->kind == ETRAJ_LASER_COS2) ? ch : dmul(dmul(ch, ch), ch);
        g = dmul(dmul(dscale(m->F0, env), w1), w2);
        a = dadd(dmul(ch, sw), dscale(cN, dmul(sh, cw)));
        b = dadd(dscale(-1, dmul(ch, cw)), dscale(cN, dmul(sh, sw)));
    }
    dual_t fx, fy;
    if (az == 0) {
        fx = dmul(g, a);
        fy = dscale(e, dmul(g, b));
    } else {
        fx = dmul(g, dadd(dscale(cos(az), a), dscale(e * sin(az), b)));
        fy = dmul(g, dadd(dscale(-sin(az), a), dscale(e * cos(az), b)));
    }
    *Fx = fx.v; *Fy = fy.v; *dFx = fx.d; *dFy = fy.d;
}

static void laser_F_dual(const laser_t* L, double t, double* Fx, double* Fy, double* dFx, double* dFy) {
    mono_F_dual(&L->c[0], t, Fx, Fy, dFx, dFy);
    if (L->n == 2) {
        double a, b, c, d;
        mono_F_dual(&L->c[1], t - L->delay, &a, &b, &c, &d);
        *Fx += a; *Fy += b; *dFx += c; *dFy += d;
    }
}

int oracle_laser_field_deriv(const etraj_laser_desc* d, double t, double out[4]) {
    laser_t L;
    if (laser_init(d, &L)) return ETRAJ_EINVAL;
    laser_F_dual(&L, t, &out[0], &out[1], &out[2], &out[3]);
    return ETRAJ_OK;
}

/* dkdt of one sample (ADKSampler.jl:267-276 and its copies in SPA/SPANE/WFAT) */
static double sample_dkdt(const sampler_t* S, int64_t* nk) {
    const etraj_sampler_desc* sd = S->sd;
    const int dim = sd->dimension;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        double kd_step = sd->ss_kd_num > 1 ? (2 * fabs(sd->ss_kd_max)) / sd->ss_kd_num : 1.0;
        double kz_step = (dim == 3 && sd->ss_kz_num > 1) ? (2 * fabs(sd->ss_kz_max)) / sd->ss_kz_num : 1.0;
        *nk = (int64_t)sd->ss_kd_num * (dim == 3 ? sd->ss_kz_num : 1);
        return S->tstep * kd_step * kz_step;
    }
    *nk = sd->mc_kt_num;
    if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
        double kd_span = sd->mc_kd_max > 0 ? 2 * sd->mc_kd_max : 1.0;
        double kz_span = sd->mc_kz_max > 0 ? 2 * sd->mc_kz_max : 1.0;
        return S->tstep * kd_span * kz_span / sd->mc_kt_num;
    }
    return (sd->sample_t_hi - sd->sample_t_lo) / sd->sample_t_num / sd->mc_kt_num;
}

/* (kd, kz) of sample j of batch b and its weight factor w (dkdt, over the
 * pdf for stratified sampling): the loops of ADKSampler.jl:315-380. */
static void sample_k(const sampler_t* S, int b, int64_t j, double dkdt, double* kd, double* kz, double* w) {
    const etraj_sampler_desc* sd = S->sd;
    const int dim = sd->dimension;
    *kz = 0.0;
    *w = dkdt;
    if (sd->sampling == ETRAJ_SAMPLING_STEP) {
        const int kzn = (dim == 3) ? sd->ss_kz_num : 1;
        int64_t ikd = j / kzn, ikz = j % kzn;
        *kd = oracle_grid_point(-fabs(sd->ss_kd_max), fabs(sd->ss_kd_max), sd->ss_kd_num, (int)ikd);
        if (dim == 3) *kz = oracle_grid_point(-fabs(sd->ss_kz_max), fabs(sd->ss_kz_max), kzn, (int)ikz);
        return;
    }
    uint64_t c[4] = {(uint64_t)j, (uint64_t)b, 0, TAG_K}, k[2] = {sd->seed, 0}, o[4];
    oracle_philox4x64_10(c, k, o);
    if (sd->sampling == ETRAJ_SAMPLING_MONTECARLO) {
        *kd = (u01_open_right(o[0]) - 0.5) * 2 * sd->mc_kd_max;
        if (dim == 3) {
            *kz = (u01_open_right(o[1]) - 0.5) * 2 * sd->mc_kz_max;
            if (sd->mc_kd_max == 0) *kd = 0.0;
            if (sd->mc_kz_max == 0) *kz = 0.0;
        }
        return;
    }
    double r = sqrt(-2.0 * log(u01_open_left(o[0])));
    double a = 2 * M_PI * u01_open_right(o[1]);
    *kd = sd->kd_sigma * (r * cos(a));
    *w = *w / gauss_pdf(*kd, sd->kd_sigma);
    if (dim == 3) {
        *kz = sd->kz_sigma * (r * sin(a));
        *w = *w / gauss_pdf(*kz, sd->kz_sigma);
    }
}

static int put_column(const etraj_sampler_desc* sd, double* out, int64_t* count, int64_t cap, double x0, double y0,
                      double kx0, double ky0, double kz0, double tr, double rate, double phase) {
    const int dim = sd->dimension, ph = sd->phase_method;
    const int rows = (dim == 3) ? (ph == ETRAJ_PHASE_CTMC ? 8 : 9) : (ph == ETRAJ_PHASE_CTMC ? 6 : 7);
    if (*count >= cap) return -1;
    double* col = out + *count * rows;
    if (dim == 3) {
        double v[9] = {x0, y0, 0.0, kx0, ky0, kz0, tr, rate, phase};
        memcpy(col, v, sizeof(double) * rows);
    } else {
        double v[7] = {x0, y0, kx0, ky0, tr, rate, phase};
        memcpy(col, v, sizeof(double) * rows);
    }
    ++*count;
    return 0;
}

/* gen_electron_batch(sp::SPANESampler, batch_id) for an atom with l = 0
 * (SPANESampler.jl:162-420): the non-adiabatic expansion of the SFA saddle
 * point to first order in dF/dt. For l = 0 the amplitude is real and
 * positive, so angle(amp) = 0. */
static int64_t spane_gen_batch(const sampler_t* S, int b, double* out, int64_t cap) {
    const etraj_sampler_desc* sd = S->sd;
    const double tr = S->tsamp[b];
    double Fx, Fy, dFx, dFy;
    laser_F_dual(&S->L, tr, &Fx, &Fy, &dFx, &dFy);                    /* :164-167 */
    const double Ftr = hypot(Fx, Fy), phi = atan2(-Fy, -Fx);           /* :168-169 */
    const double Ip = S->Ip, Z = (double)S->T.Z;
    double gam = 0.0;                                                  /* :178-182 */
    if (S->L.n == 1) gam = S->L.c[0].w * sqrt(2 * Ip) / (S->L.c[0].F0 * mono_envelope(&S->L.c[0], tr));
    const double kappa = sqrt(2 * Ip), n = Z / kappa;
    const double e0 = 2.71828182845904523;
    const double c_cc = pow(2, 3 * n / 2 + 1) * pow(kappa, 5 * n + 0.5) * pow(Ftr, -n) * pow(1 + 2 * gam / e0, -n);
    const double numer = S->T.C * sqrt(1.0 / (4 * M_PI));           /* C D^0_00 Y_00 */
    prefac_t PF;
    prefac_init(S, Fx, Fy, &PF);                                       /* l > 0 or a molecule: :188-257 */
    const double cross = Fx * dFy - dFx * Fy;                          /* :238 */
    int64_t nk;
    const double dkdt = sample_dkdt(S, &nk);
    const double f2min = Ftr * Ftr / (sd->sampling == ETRAJ_SAMPLING_STEP ? 9.0 : 25.0);   /* :332, :378 (Q7) */
    int64_t count = 0;
    for (int64_t j = 0; j < nk; ++j) {
        double kd0, kz0, w;
        sample_k(S, b, j, dkdt, &kd0, &kz0, &w);
        if (fabs(kd0) < 1e-4) continue;                                /* :323 */
        const double kx0 = kd0 * -sin(phi), ky0 = kd0 * cos(phi);
        const double k2 = kx0 * kx0 + ky0 * ky0 + kz0 * kz0;
        const double f2 = Ftr * Ftr - (kx0 * dFx + ky0 * dFy);         /* F2eff :239 */
        const double r0 = Ftr / 2 * (k2 + 2 * Ip) / f2;                 /* :240 */
        if (f2 <= f2min) continue;
        double complex amp = sqrt(w) * exp(-pow(kd0 * kd0 + kz0 * kz0 + 2 * Ip, 1.5) / (3 * sqrt(f2)));  /* :275 */
        if (S->prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
            const double cpre = (S->prefix & ETRAJ_PREFIX_PRE) ? S->c_pre : c_cc;
            double complex num = numer;
            if (PF.on) {   /* ti = sqrt((kappa^2+kd^2+kz^2)/F2eff), k_ts of :241-242 */
                const double ti = sqrt((kappa * kappa + kd0 * kd0 + kz0 * kz0) / f2);
                double complex kts[3] = {kx0 - I * ti * Fx + ti * ti / 2 * dFx, ky0 - I * ti * Fy + ti * ti / 2 * dFy,
                                         kz0};
                num = prefac_numer(&PF, kts);
            }
            amp *= cpre * num / pow((k2 + 2 * Ip) * f2, (n + 1) / 4);   /* :262-269 */
        }
        if (S->prefix & ETRAJ_PREFIX_JAC)
            amp *= sqrt(fabs(Ftr + sqrt(kd0 * kd0 + kz0 * kz0) / (Ftr * Ftr) * cross));   /* :270 */
        const double rate = creal(amp) * creal(amp) + cimag(amp) * cimag(amp);
        if (isnan(rate) || rate / w < sd->cutoff_limit) continue;      /* :337 (rate/dkdt, Q7) */
        if (put_column(sd, out, &count, cap, r0 * cos(phi), r0 * sin(phi), kx0, ky0, kz0, tr, rate, carg(amp)))
            return -1;
    }
    return count;
}

/* Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n). */
static void gauss_legendre(int n, double* x, double* wt) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int k = 1; k <= n; ++k) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * k - 1) * z * p2 - (k - 1) * p3) / k;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        wt[i] = wt[n - 1 - i] = 2 / ((1 - z * z) * pp * pp);
    }
}

/* The SFA-SPA ingredients at complex time tr + i tau (SPASampler.jl:230-292). */
typedef struct {
    const laser_t* L;
    double tr, Ip, kz;
    double px, py;       /* canonical momentum (set after the saddle point) */
} spa_ctx;

/* px_/py_ of SPASampler.jl:231-232 */
static void spa_p(double kd, double complex Ax, double complex Ay, double* px, double* py) {
    const double rho = sqrt(cimag(Ax) * cimag(Ax) + cimag(Ay) * cimag(Ay));
    *px = kd * cimag(Ay) / rho - creal(Ax);
    *py = -kd * cimag(Ax) / rho - creal(Ay);
}

/* Saddle-point equation (SPASampler.jl:233-238) solved for ti > 0. With
 * px, py of :231-232 the bracket (px + A)^2 + (py + A)^2 is real and equals
 * kd^2 - |Im A(ts)|^2, so f(ti) = (kd^2 + kz^2 - |Im A|^2)/2 + Ip, and
 * f'(ti) = Im Ax Re Fx + Im Ay Re Fy (dA/dti = -i F). Newton from the
 * reference's initial guess asinh(w/F sqrt(k^2 + 2Ip))/w (:240); converged if
 * |f| <= 1e-8 (NLsolve's ftol) and ti > 0 (:241-246); 0 otherwise. */
static double spa_solve(const laser_t* L, double tr, double Ftr, double omega, double Ip, double kd, double kz) {
    double ti = asinh(omega / Ftr * sqrt(kd * kd + kz * kz + 2 * Ip)) / omega;
    double f = 0;
    for (int it = 0; it < 100; ++it) {
        double complex Fx, Fy, Ax, Ay;
        laser_FA_c(L, tr + I * ti, &Fx, &Fy, &Ax, &Ay);
        const double ia = cimag(Ax), ib = cimag(Ay);
        f = (kd * kd + kz * kz - (ia * ia + ib * ib)) / 2 + Ip;
        const double fp = ia * creal(Fx) + ib * creal(Fy);
        const double step = f / fp;
        if (!isfinite(step)) return 0.0;
        ti -= step;
        if (fabs(step) <= 1e-15 * fabs(ti)) break;
    }
    {
        double complex Fx, Fy, Ax, Ay;
        laser_FA_c(L, tr + I * ti, &Fx, &Fy, &Ax, &Ay);
        f = (kd * kd + kz * kz - (cimag(Ax) * cimag(Ax) + cimag(Ay) * cimag(Ay))) / 2 + Ip;
    }
    if (!(fabs(f) <= 1e-8) || !(ti > 0)) return 0.0;
    return ti;
}

/* Adaptive Gauss-Legendre (16 against 32 points, bisection) of three complex
 * integrands over tau in [0, ti]: x0 = int Im Ax, y0 = int Im Ay (:350-351)
 * and G = int g(tr + i tau) with g = ((px+Ax)^2 + (py+Ay)^2 + kz^2)/2 + Ip,
 * so that S_tun = i G (:279: S_tun = -int_{ts}^{tr} g dt along the segment). */
typedef struct { double x0, y0; double complex G; } spa_ints;
static double gl16x[16], gl16w[16], gl32x[32], gl32w[32];
static pthread_once_t gl_once = PTHREAD_ONCE_INIT;
static void gl_init(void) { gauss_legendre(16, gl16x, gl16w); gauss_legendre(32, gl32x, gl32w); }

static spa_ints spa_rule(const spa_ctx* C, double a, double b, int n, const double* x, const double* wt) {
    spa_ints r = {0, 0, 0};
    const double h = (b - a) / 2, m = (a + b) / 2;
    for (int i = 0; i < n; ++i) {
        double complex Fx, Fy, Ax, Ay;
        laser_FA_c(C->L, C->tr + I * (m + h * x[i]), &Fx, &Fy, &Ax, &Ay);
        const double complex kx = C->px + Ax, ky = C->py + Ay;
        r.x0 += wt[i] * cimag(Ax);
        r.y0 += wt[i] * cimag(Ay);
        r.G += wt[i] * ((kx * kx + ky * ky + C->kz * C->kz) / 2 + C->Ip);
    }
    r.x0 *= h; r.y0 *= h; r.G *= h;
    return r;
}

static spa_ints spa_integrate(const spa_ctx* C, double a, double b, int depth) {
    spa_ints lo = spa_rule(C, a, b, 16, gl16x, gl16w), hi = spa_rule(C, a, b, 32, gl32x, gl32w);
    const double err = fabs(hi.x0 - lo.x0) + fabs(hi.y0 - lo.y0) + cabs(hi.G - lo.G);
    const double mag = fabs(hi.x0) + fabs(hi.y0) + cabs(hi.G);
    if (err <= 1e-14 * mag || depth >= 30) return hi;
    spa_ints l = spa_integrate(C, a, (a + b) / 2, depth + 1), r = spa_integrate(C, (a + b) / 2, b, depth + 1);
    spa_ints s = {l.x0 + r.x0, l.y0 + r.y0, l.G + r.G};
    return s;
}

typedef struct {
    double ti, px, py, x0, y0, kx0, ky0;
    double complex G;      /* S_tun = i G */
    double complex kdotF;  /* k(ts) . F(ts) */
    double complex ktx, kty;   /* k_ts = (px + Ax(ts), py + Ay(ts)) (:250) */
    double jac;
} spa_point_t;

/* One sample's saddle point and tunnelling integrals; returns 0 if the
 * saddle-point solve fails (the reference's `ti == 0 && continue`). */
static int spa_point(const sampler_t* S, double tr, double Ftr, double omega, double kd, double kz, int want_jac,
                     spa_point_t* P) {
    const laser_t* L = &S->L;
    const double Ip = S->Ip;
    pthread_once(&gl_once, gl_init);
    const double ti = spa_solve(L, tr, Ftr, omega, Ip, kd, kz);
    if (ti == 0) return 0;
    double complex Fx, Fy, Ax, Ay;
    laser_FA_c(L, tr + I * ti, &Fx, &Fy, &Ax, &Ay);
    P->ti = ti;
    spa_p(kd, Ax, Ay, &P->px, &P->py);
    spa_ctx C = {L, tr, Ip, kz, P->px, P->py};
    spa_ints in = spa_integrate(&C, 0.0, ti, 0);
    P->x0 = in.x0;
    P->y0 = in.y0;
    P->G = in.G;
    double Axr, Ayr;
    laser_A(L, tr, &Axr, &Ayr);
    P->kx0 = P->px + Axr;                                             /* :352-353 */
    P->ky0 = P->py + Ayr;
    P->kdotF = (P->px + Ax) * Fx + (P->py + Ay) * Fy;                 /* :272-277 */
    P->ktx = P->px + Ax;
    P->kty = P->py + Ay;
    P->jac = 0;
    if (want_jac) {                                                   /* :280-292 */
        const double dtr = 0.01, dkd = 0.01;
        const double t_tp = spa_solve(L, tr + dtr, Ftr, omega, Ip, kd, kz);
        const double t_tm = spa_solve(L, tr - dtr, Ftr, omega, Ip, kd, kz);
        const double t_kp = spa_solve(L, tr, Ftr, omega, Ip, kd + dkd, kz);
        const double t_km = spa_solve(L, tr, Ftr, omega, Ip, kd - dkd, kz);
        double complex f1, f2, a1, b1, a2, b2;
        double pxa, pya, pxb, pyb;
        laser_FA_c(L, tr + dtr + I * t_tp, &f1, &f2, &a1, &b1);
        laser_FA_c(L, tr - dtr + I * t_tm, &f1, &f2, &a2, &b2);
        spa_p(kd, a1, b1, &pxa, &pya);
        spa_p(kd, a2, b2, &pxb, &pyb);
        const double dpxdtr = pxa - pxb, dpydtr = pya - pyb;
        laser_FA_c(L, tr + I * t_kp, &f1, &f2, &a1, &b1);
        laser_FA_c(L, tr + I * t_km, &f1, &f2, &a2, &b2);
        spa_p(kd + dkd, a1, b1, &pxa, &pya);
        spa_p(kd - dkd, a2, b2, &pxb, &pyb);
        const double dpxdkd = pxa - pxb, dpydkd = pya - pyb;
        P->jac = fabs(dpxdtr * dpydkd - dpxdkd * dpydtr) / (4 * dtr * dkd);
    }
    return 1;
}

/* omega of SPASampler.jl:160-166 (the initial guess of the saddle point) */
static double spa_omega(const laser_t* L) {
    if (L->n == 1) return L->c[0].w;
    return (L->c[0].w + L->c[1].w) / 2;
}

int oracle_spa_point(const etraj_laser_desc* ld, const etraj_target_desc* td, double tr, double kd, double kz,
                     double out[12]) {
    sampler_t S;
    memset(&S, 0, sizeof S);
    if (laser_init(ld, &S.L) || target_init(td, &S.T)) return ETRAJ_EINVAL;
    S.Ip = S.T.Ip;
    double Fx, Fy;
    laser_F(&S.L, tr, &Fx, &Fy);
    spa_point_t P;
    if (!spa_point(&S, tr, hypot(Fx, Fy), spa_omega(&S.L), kd, kz, 1, &P)) return 1;
    double v[12] = {P.ti, P.px, P.py, P.x0, P.y0, P.kx0, P.ky0, creal(P.G), cimag(P.G), creal(P.kdotF),
                    cimag(P.kdotF), P.jac};
    memcpy(out, v, sizeof v);
    return ETRAJ_OK;
}

/* gen_electron_batch(sp::SPASampler, batch_id) for an atom with l = 0
 * (SPASampler.jl:150-431). */
static int64_t spa_gen_batch(const sampler_t* S, int b, double* out, int64_t cap) {
    const etraj_sampler_desc* sd = S->sd;
    const double tr = S->tsamp[b];
    double Fx, Fy;
    laser_F(&S->L, tr, &Fx, &Fy);
    const double Ftr = hypot(Fx, Fy);
    const double omega = spa_omega(&S->L);
    if (Ftr <= 1e-3) return 0;                                        /* :167-169 (Q8) */
    const double Ip = S->Ip, Z = (double)S->T.Z;
    double gam = 0.0;                                                 /* :177-181 */
    if (S->L.n == 1) gam = S->L.c[0].w * sqrt(2 * Ip) / (S->L.c[0].F0 * mono_envelope(&S->L.c[0], tr));
    const double kappa = sqrt(2 * Ip), n = Z / kappa;
    const double e0 = 2.71828182845904523;
    const double c_cc = pow(2, 3 * n / 2 + 1) * pow(kappa, 5 * n + 0.5) * pow(Ftr, -n) * pow(1 + 2 * gam / e0, -n);
    const double numer = S->T.C * sqrt(1.0 / (4 * M_PI));
    prefac_t PF;
    prefac_init(S, Fx, Fy, &PF);                                      /* l > 0 or a molecule: :184-270 */
    int64_t nk;
    const double dkdt = sample_dkdt(S, &nk);
    const int jac = (S->prefix & ETRAJ_PREFIX_JAC) != 0;
    int64_t count = 0;
    for (int64_t j = 0; j < nk; ++j) {
        double kd0, kz0, w;
        sample_k(S, b, j, dkdt, &kd0, &kz0, &w);
        if (fabs(kd0) < 1e-4) continue;                               /* :345 */
        spa_point_t P;
        if (!spa_point(S, tr, Ftr, omega, kd0, kz0, jac, &P)) continue;
        double complex amp = sqrt(w) * cexp(-P.G);                    /* sqrt(dkdt) exp(i S_tun) */
        if (S->prefix & (ETRAJ_PREFIX_PRE | ETRAJ_PREFIX_PRECC)) {
            const double cpre = (S->prefix & ETRAJ_PREFIX_PRE) ? S->c_pre : c_cc;
            double complex num = numer;
            if (PF.on) {
                double complex kts[3] = {P.ktx, P.kty, kz0};
                num = prefac_numer(&PF, kts);
            }
            amp *= cpre * num / cpow(P.kdotF, (n + 1) / 2);           /* :271-278 */
        }
        if (jac) amp *= sqrt(P.jac);
        const double rate = creal(amp) * creal(amp) + cimag(amp) * cimag(amp);
        if (isnan(rate) || rate < sd->cutoff_limit) continue;         /* :357 */
        if (put_column(sd, out, &count, cap, P.x0, P.y0, P.kx0, P.ky0, kz0, tr, rate, carg(amp))) return -1;
    }
    return count;
}

static int64_t gen_batch(const sampler_t* S, int b, double* out, int64_t cap) {
    switch (S->sd->init_method) {
        case ETRAJ_INIT_WFAT: return wfat_gen_batch(S, b, out, cap);
        case ETRAJ_INIT_SPA: return spa_gen_batch(S, b, out, cap);
        case ETRAJ_INIT_SPANE: return spane_gen_batch(S, b, out, cap);
        default: return adk_gen_batch(S, b, out, cap);
    }
}

int oracle_gen_electron_batch(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                              int32_t batch_id, double* out, int64_t capacity, int64_t* count) {
    sampler_t S;
    int rc = sampler_init(ld, td, sd, &S);
    if (rc) { sampler_free(&S); return rc; }
    if (batch_id < 1 || batch_id > sd->sample_t_num) { sampler_free(&S); return ETRAJ_EINVAL; }
    int64_t c = gen_batch(&S, batch_id - 1, out, capacity);
    sampler_free(&S);
    if (c < 0) return ETRAJ_ECAPACITY;
    *count = c;
    return ETRAJ_OK;
}

/* ========================================================================== */
/* Integration of a sampled batch                                              */
/* ========================================================================== */
typedef struct {
    laser_t L;
    target_t T;
    const etraj_run_desc* rd;
    eom_t E;
    int nv, rows;
} runner_t;

static int runner_init(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd, int32_t rows, runner_t* R) {
    if (laser_init(ld, &R->L) || target_init(td, &R->T)) return ETRAJ_EINVAL;
    if (rd->dimension != 2 && rd->dimension != 3) return ETRAJ_EINVAL;
    if (rd->phase_method < 0 || rd->phase_method > 2) return ETRAJ_EINVAL;
    int exp_rows = (rd->dimension == 3) ? (rd->phase_method == ETRAJ_PHASE_CTMC ? 8 : 9) : (rd->phase_method == ETRAJ_PHASE_CTMC ? 6 : 7);
    if (rows != exp_rows) return ETRAJ_EINVAL;
    if (rd->integrator == ETRAJ_INTEG_RK4 && !(rd->dt > 0)) return ETRAJ_EINVAL;
    if (rd->integrator == ETRAJ_INTEG_TSIT5 && !(rd->rtol > 0)) return ETRAJ_EINVAL;
    R->rd = rd;
    R->E.L = &R->L; R->E.T = &R->T; R->E.dim = rd->dimension; R->E.phase = rd->phase_method; R->E.quirks = rd->quirks;
    R->nv = 2 * rd->dimension + (rd->phase_method == ETRAJ_PHASE_CTMC ? 0 : 1);
    R->rows = rows;
    return ETRAJ_OK;
}

/* one electron: TrajectorySimulation.jl:337-342 (remake u0, tspan) + :344-374 */
static void run_one(const runner_t* R, const double* col, post_t* P, int64_t* nsteps) {
    const etraj_run_desc* rd = R->rd;
    double u[7];
    int dim = rd->dimension;
    double t0, rate;
    for (int i = 0; i < 2 * dim; ++i) u[i] = col[i];
    t0 = col[2 * dim];
    rate = col[2 * dim + 1];
    (void)rate;
    if (rd->phase_method != ETRAJ_PHASE_CTMC) u[2 * dim] = col[2 * dim + 2];
    int fail;
    if (rd->integrator == ETRAJ_INTEG_RK4)
        fail = integrate_rk4(&R->E, R->nv, t0, rd->traj_t_final, rd->dt, u, nsteps);
    else
        fail = integrate_tsit5(&R->E, R->nv, t0, rd->traj_t_final, rd->rtol, rd->atol > 0 ? rd->atol : 1e-6,
                               rd->maxiters > 0 ? rd->maxiters : 1000000, u, nsteps);
    post_process(&R->T, dim, rd->phase_method, rd->quirks, u, t0, P);
    if (fail) P->status = 1;
}

int oracle_integrate_records(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                             const double* init, int64_t n, int32_t rows, double* records, int32_t* status) {
    runner_t R;
    int rc = runner_init(ld, td, rd, rows, &R);
    if (rc) return rc;
    for (int64_t i = 0; i < n; ++i) {
        post_t P;
        int64_t ns;
        run_one(&R, init + i * rows, &P, &ns);
        memcpy(records + 12 * i, P.rec, sizeof P.rec);
        status[i] = P.status;
    }
    return ETRAJ_OK;
}

/* binning of one electron (TrajectorySimulation.jl:370-396) */
typedef struct {
    double* pmd;      /* CTMC: n doubles; coherent: 2n (re, im) */
    double ion_prob, ion_unc;
    int64_t failed, bound, eff, steps;
} acc_t;

static void bin_one(const runner_t* R, const post_t* P, double prob, acc_t* A) {
    const etraj_run_desc* rd = R->rd;
    if (P->status == 1) { A->failed++; return; }
    if (P->status == 2) { A->bound++; return; }
    A->ion_prob += prob;
    int dim = rd->dimension;
    int64_t idx[3] = {1, 1, 1};
    int inb = 1;
    for (int d = 0; d < dim; ++d) {
        idx[d] = oracle_bin_index(P->rec[8 + d], rd->final_p_max[d], rd->final_p_num[d]);
        if (idx[d] < 1 || idx[d] > rd->final_p_num[d]) inb = 0;
    }
    if (!inb) { A->ion_unc += prob; return; }
    int64_t lin = (idx[0] - 1) + (int64_t)rd->final_p_num[0] * ((idx[1] - 1) + (dim == 3 ? (int64_t)rd->final_p_num[1] * (idx[2] - 1) : 0));
    if (rd->phase_method == ETRAJ_PHASE_CTMC) {
        A->pmd[lin] += prob;
    } else {
        double s = sqrt(prob);
        double complex a = s * cexp(I * P->rec[11]);
        A->pmd[2 * lin] += creal(a);
        A->pmd[2 * lin + 1] += cimag(a);
    }
}

int oracle_integrate_and_bin(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                             const double* init, int64_t n, int32_t rows, double* pmd, double* ion_prob,
                             double* ion_prob_uncollected, int64_t* n_failed) {
    runner_t R;
    int rc = runner_init(ld, td, rd, rows, &R);
    if (rc) return rc;
    acc_t A;
    memset(&A, 0, sizeof A);
    A.pmd = pmd;
    for (int64_t i = 0; i < n; ++i) {
        post_t P;
        int64_t ns;
        const double* col = init + i * rows;
        run_one(&R, col, &P, &ns);
        bin_one(&R, &P, col[2 * rd->dimension + 1], &A);
    }
    *ion_prob += A.ion_prob;
    *ion_prob_uncollected += A.ion_unc;
    *n_failed = A.failed;
    return ETRAJ_OK;
}

/* ========================================================================== */
/* Whole job: Threads.@threads over batches with thread-private PMDs and one   */
/* final sum (TrajectorySimulation.jl:154-157, :262-267, :521-540).            */
/* ========================================================================== */
typedef struct {
    const sampler_t* S;
    const runner_t* R;
    int b_begin, b_end, b_stride;
    int tid, nthreads;
    int64_t pmd_len;
    int64_t* next;   /* non-NULL: batches are taken one at a time from this shared counter */
    acc_t A;
    int rc;
} worker_t;

static void* worker_main(void* arg) {
    worker_t* W = (worker_t*)arg;
    int64_t cap = oracle_batch_max_size(W->S->sd);
    int rows = W->R->rows;
    double* buf = (double*)malloc(sizeof(double) * rows * (cap > 0 ? cap : 1));
    if (!buf) { W->rc = ETRAJ_ENOMEM; return NULL; }
    /* static, contiguous partition of the selected batches over threads, or (next != NULL)
       dynamic: each thread takes the next batch from a shared counter, as Julia 1.11's
       @threads :dynamic schedule does */
    int64_t nsel = (W->b_end - W->b_begin + W->b_stride - 1) / W->b_stride;
    int64_t lo = nsel * W->tid / W->nthreads, hi = nsel * (W->tid + 1) / W->nthreads;
    if (W->next) { lo = 0; hi = nsel; }
    for (int64_t s = lo; s < hi; ++s) {
        if (W->next) {
            s = __atomic_fetch_add(W->next, 1, __ATOMIC_RELAXED);
            if (s >= nsel) break;
        }
        int b = W->b_begin + (int)(s * W->b_stride);
        int64_t c = gen_batch(W->S, b, buf, cap);
        if (c <= 0) continue;
        for (int64_t i = 0; i < c; ++i) {
            post_t P;
            int64_t ns = 0;
            const double* col = buf + i * rows;
            run_one(W->R, col, &P, &ns);
            W->A.steps += ns;
            bin_one(W->R, &P, col[2 * W->R->rd->dimension + 1], &W->A);
        }
        W->A.eff += c;
    }
    free(buf);
    return NULL;
}

static int run_job_impl(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                        const etraj_run_desc* rd, int nthreads, int32_t batch_begin, int32_t batch_end,
                        int32_t batch_stride, int dynamic, double* pmd_out, etraj_stats* st) {
    sampler_t S;
    runner_t R;
    int rc = sampler_init(ld, td, sd, &S);
    if (rc) { sampler_free(&S); return rc; }
    if (sd->dimension != rd->dimension || sd->phase_method != rd->phase_method) { sampler_free(&S); return ETRAJ_EINVAL; }
    int rows = (sd->dimension == 3) ? (sd->phase_method == ETRAJ_PHASE_CTMC ? 8 : 9) : (sd->phase_method == ETRAJ_PHASE_CTMC ? 6 : 7);
    rc = runner_init(ld, td, rd, rows, &R);
    if (rc) { sampler_free(&S); return rc; }
    if (batch_end <= 0 || batch_end > sd->sample_t_num) batch_end = sd->sample_t_num;
    if (batch_begin < 0) batch_begin = 0;
    if (batch_stride <= 0) batch_stride = 1;
    if (nthreads <= 0) nthreads = 1;
    int64_t npix = (int64_t)rd->final_p_num[0] * rd->final_p_num[1] * (rd->dimension == 3 ? rd->final_p_num[2] : 1);
    int coh = rd->phase_method != ETRAJ_PHASE_CTMC;
    int64_t len = coh ? 2 * npix : npix;
    worker_t* W = (worker_t*)calloc(nthreads, sizeof(worker_t));
    pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
    int64_t next = 0;
    if (!W || !th) { free(W); free(th); sampler_free(&S); return ETRAJ_ENOMEM; }
    for (int i = 0; i < nthreads; ++i) {
        W[i].S = &S; W[i].R = &R;
        W[i].next = dynamic ? &next : NULL;
        W[i].b_begin = batch_begin; W[i].b_end = batch_end; W[i].b_stride = batch_stride;
        W[i].tid = i; W[i].nthreads = nthreads; W[i].pmd_len = len;
        W[i].A.pmd = (double*)calloc(len, sizeof(double));
        if (!W[i].A.pmd) rc = ETRAJ_ENOMEM;
    }
    if (!rc) {
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, worker_main, &W[i]);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    etraj_stats s;
    memset(&s, 0, sizeof s);
    if (!rc) {
        memset(pmd_out, 0, sizeof(double) * npix);
        double* sum = (double*)calloc(len, sizeof(double));
        for (int i = 0; i < nthreads; ++i) {
            if (W[i].rc) rc = W[i].rc;
            for (int64_t k = 0; k < len; ++k) sum[k] += W[i].A.pmd[k];
            s.ion_prob += W[i].A.ion_prob;
            s.ion_prob_uncollected += W[i].A.ion_unc;
            s.num_effective_traj += W[i].A.eff;
            s.n_failed += W[i].A.failed;
            s.n_bound += W[i].A.bound;
            s.n_steps += W[i].A.steps;
        }
        for (int64_t k = 0; k < npix; ++k)
            pmd_out[k] = coh ? sum[2 * k] * sum[2 * k] + sum[2 * k + 1] * sum[2 * k + 1] : sum[k];
        free(sum);
    }
    for (int i = 0; i < nthreads; ++i) free(W[i].A.pmd);
    free(W);
    free(th);
    sampler_free(&S);
    if (st) *st = s;
    return rc;
}

int oracle_run_job(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                   const etraj_run_desc* rd, int nthreads, int32_t batch_begin, int32_t batch_end, int32_t batch_stride,
                   double* pmd_out, etraj_stats* st) {
    return run_job_impl(ld, td, sd, rd, nthreads, batch_begin, batch_end, batch_stride, 0, pmd_out, st);
}

/* the same with dynamic scheduling of the batches over the threads (timing runs: the PMD's
   summation order then depends on the thread timing) */
int oracle_run_job_dynamic(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                           const etraj_run_desc* rd, int nthreads, int32_t batch_begin, int32_t batch_end,
                           int32_t batch_stride, double* pmd_out, etraj_stats* st) {
    return run_job_impl(ld, td, sd, rd, nthreads, batch_begin, batch_end, batch_stride, 1, pmd_out, st);
}

/* amplitude-level job output (before abs2), for coherent-sum checks */
int oracle_run_job_amplitude(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_sampler_desc* sd,
                             const etraj_run_desc* rd, int nthreads, double* amp_out, etraj_stats* st) {
    sampler_t S;
    runner_t R;
    int rc = sampler_init(ld, td, sd, &S);
    if (rc) { sampler_free(&S); return rc; }
    int rows = (sd->dimension == 3) ? (sd->phase_method == ETRAJ_PHASE_CTMC ? 8 : 9) : (sd->phase_method == ETRAJ_PHASE_CTMC ? 6 : 7);
    rc = runner_init(ld, td, rd, rows, &R);
    if (rc) { sampler_free(&S); return rc; }
    int64_t npix = (int64_t)rd->final_p_num[0] * rd->final_p_num[1] * (rd->dimension == 3 ? rd->final_p_num[2] : 1);
    int64_t len = rd->phase_method != ETRAJ_PHASE_CTMC ? 2 * npix : npix;
    worker_t W;
    memset(&W, 0, sizeof W);
    W.S = &S; W.R = &R; W.b_begin = 0; W.b_end = sd->sample_t_num; W.b_stride = 1; W.tid = 0; W.nthreads = 1;
    W.A.pmd = amp_out;
    memset(amp_out, 0, sizeof(double) * len);
    (void)nthreads;
    worker_main(&W);
    if (st) {
        memset(st, 0, sizeof *st);
        st->ion_prob = W.A.ion_prob; st->ion_prob_uncollected = W.A.ion_unc;
        st->num_effective_traj = W.A.eff; st->n_failed = W.A.failed; st->n_bound = W.A.bound; st->n_steps = W.A.steps;
    }
    sampler_free(&S);
    return W.rc;
}

/* single-trajectory integrator access for the convergence / invariance tests:
 * integrates u (nv values) from t0 to tf, returns steps in *nsteps */
int oracle_integrate_state(const etraj_laser_desc* ld, const etraj_target_desc* td, const etraj_run_desc* rd,
                           double t0, double* u, int64_t* nsteps) {
    runner_t R;
    int rows = (rd->dimension == 3) ? (rd->phase_method == ETRAJ_PHASE_CTMC ? 8 : 9) : (rd->phase_method == ETRAJ_PHASE_CTMC ? 6 : 7);
    int rc = runner_init(ld, td, rd, rows, &R);
    if (rc) return rc;
    int fail = (rd->integrator == ETRAJ_INTEG_RK4)
                   ? integrate_rk4(&R.E, R.nv, t0, rd->traj_t_final, rd->dt, u, nsteps)
                   : integrate_tsit5(&R.E, R.nv, t0, rd->traj_t_final, rd->rtol, rd->atol > 0 ? rd->atol : 1e-6,
                                     rd->maxiters > 0 ? rd->maxiters : 1000000, u, nsteps);
    return fail ? 1 : 0;
}

/* Kepler / SCTS post-processing of a given final state (for invariance tests) */
int oracle_post_process(const etraj_target_desc* td, int32_t dim, int32_t phase_method, int32_t quirks,
                        const double* u, double t0, double* rec, int32_t* status) {
    target_t T;
    if (target_init(td, &T)) return ETRAJ_EINVAL;
    post_t P;
    post_process(&T, dim, phase_method, quirks, u, t0, &P);
    memcpy(rec, P.rec, sizeof P.rec);
    *status = P.status;
    return ETRAJ_OK;
}

/* the RHS itself (for direct checks) */
int oracle_eom_rhs(const etraj_laser_desc* ld, const etraj_target_desc* td, int32_t dim, int32_t phase_method,
                   int32_t quirks, double t, const double* u, double* du) {
    laser_t L;
    target_t T;
    if (laser_init(ld, &L) || target_init(td, &T)) return ETRAJ_EINVAL;
    eom_t E = {&L, &T, dim, phase_method, quirks};
    eom_rhs(&E, t, u, du);
    return ETRAJ_OK;
}
