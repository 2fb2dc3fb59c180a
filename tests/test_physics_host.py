"""Host-compiled checks of the device math that re-arranges a formula of the
reference for speed (etraj.jl_amd/csrc/etraj_physics.hpp), against the plain
form evaluated with the C library:

  * laser_field_at (the laser at t + d from sin/cos at t by angle addition, the
    cut-off window as logistic exps) equals laser_field(t + d)
    (Cos2Laser.jl:125-152 and siblings) to 1e-14 of the peak field, for every
    envelope, ellipticity, azimuth, CEP, time shift, two colours, increments up
    to 20 a.u. (past the Taylor range of the rotation) and times over the pulse
    and its edges;
  * the one-colour cos^n stage fields (mono_field_at<true, KX>, shorter Taylor
    series for Tsit5's stages 2 and 3 and for the envelope of N >= 2 cycles;
    simple_field_back for stages 4 and 5, turned back from stage 6's phases)
    equal laser_field(t + d) to 1e-14 of the peak field, and each Taylor order
    agrees with the C library's sin/cos to 2 ulp over its range;
  * laser_field_pot_tau (the samplers' complex-time laser from its real-time
    parts) equals laser_field_pot_c to 1e-13 of the largest component;
  * cosh_sinh (both from one exp, sinh by its Taylor series below 1/2) equals
    cosh / sinh to 3 ulp over |x| up to 30, and cosh_sinh_small (both by their
    Taylor series) to 3 ulp over |x| <= 1/2.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r"""
#include "etraj_physics.hpp"
#include <cstdio>
#include <random>
using namespace etraj;
int main() {
    std::mt19937_64 g(1);
    std::uniform_real_distribution<double> U(0, 1);
    double worst_field = 0;
    for (int kind = 1; kind <= 3; ++kind)
    for (int two = 0; two < 2; ++two)
    for (int az = 0; az < 2; ++az) {
        etraj_laser_desc d{};
        d.n_colors = two ? 2 : 1;
        d.delay21 = 13.7;
        for (int c = 0; c < 2; ++c) {
            d.color[c].kind = kind;
            d.color[c].peak_int = 2e14 * (c ? 0.5 : 1);
            d.color[c].wave_len = c ? 400 : 800;
            d.color[c].cyc_num = c ? 4 : 2;
            d.color[c].ellip = c ? -1 : 0.7;
            d.color[c].azi = az ? 0.3 : 0;
            d.color[c].cep = 0.2;
            d.color[c].t_shift = c ? 5 : -3;
        }
        LaserP L;
        if (make_laser(d, L)) return 2;
        const double F0 = L.c[0].F0 + (two ? L.c[1].F0 : 0.0);
        for (int i = 0; i < 100000; ++i) {
            const double t = -300 + 600 * U(g);
            const double dt = (i % 3 == 0) ? 20 * U(g) : 0.5 * U(g);
            const LaserBase b = laser_base(L, t);
            const double cs[5] = {0.161, 0.327, 0.9, 0.9800255409045097, 1.0};
            for (int j = 0; j < 5; ++j) {
                double fx, fy, rx, ry;
                laser_field_at(L, b, cs[j] * dt, fx, fy);
                laser_field(L, t + cs[j] * dt, rx, ry);
                const double e = fmax(fabs(fx - rx), fabs(fy - ry)) / F0;
                if (e > worst_field) worst_field = e;
            }
        }
    }
    // the simple-laser stage fields (mono_field_at<true, KX>, the orders the Tsit5 stages use)
    double worst_simple = 0;
    for (int kind = 1; kind <= 2; ++kind)
    for (int cyc = 2; cyc <= 4; ++cyc) {
        etraj_laser_desc d{};
        d.n_colors = 1;
        d.color[0].kind = kind;
        d.color[0].peak_int = 2e14;
        d.color[0].wave_len = 800;
        d.color[0].cyc_num = cyc == 4 ? 6 : cyc == 3 ? 2.5 : cyc;
        d.color[0].ellip = 0.7;
        d.color[0].cep = 0.2;
        d.color[0].t_shift = -3;
        LaserP L;
        if (make_laser(d, L) || !simple_laser(L)) return 3;
        for (int i = 0; i < 100000; ++i) {
            const double t = -500 + 1000 * U(g);
            const double dt = (i % 3 == 0) ? 20 * U(g) : (i % 3 == 1) ? 14 * U(g) : 0.5 * U(g);
            const MonoBase b = mono_base(L.c[0], t);
            const double cs[5] = {0.161, 0.327, 0.9, 0.9800255409045097, 1.0};
            for (int j = 0; j < 5; ++j) {
                double fx, fy, rx, ry;
                MonoBase a6;
                double f6x, f6y;
                mono_field_at<true>(L.c[0], b, dt, f6x, f6y, &a6);
                if (j == 0) mono_field_at<true, 4>(L.c[0], b, cs[j] * dt, fx, fy);
                else if (j == 1) mono_field_at<true, 5>(L.c[0], b, cs[j] * dt, fx, fy);
                else if (j == 2) simple_field_back<3, 3>(L.c[0], b, cs[j] * dt, a6, (1.0 - cs[j]) * dt, fx, fy);
                else if (j == 3) simple_field_back<2, 2>(L.c[0], b, cs[j] * dt, a6, (1.0 - cs[j]) * dt, fx, fy);
                else { fx = f6x; fy = f6y; }
                laser_field(L, t + cs[j] * dt, rx, ry);
                const double e = fmax(fabs(fx - rx), fabs(fy - ry)) / L.c[0].F0;
                if (e > worst_simple) worst_simple = e;
            }
        }
    }
    // the Taylor series against the C library over their ranges, in ulp of the value
    double worst_taylor = 0;
    for (int i = 0; i < 5000000; ++i) {
        const int ks = i % 5;
        const double bound = ks == 0 ? 0.5 : ks == 1 ? 0.25 : ks == 2 ? 0.125 : ks == 3 ? 0.05 : 0.01;
        const double x = (2 * U(g) - 1) * bound;
        double s, c;
        if (ks == 0) sincos_taylor<7>(x, s, c);
        else if (ks == 1) sincos_taylor<5>(x, s, c);
        else if (ks == 2) sincos_taylor<4>(x, s, c);
        else if (ks == 3) sincos_taylor<3>(x, s, c);
        else sincos_taylor<2>(x, s, c);
        const double es = fabs(s - sin(x)) / (fabs(sin(x)) * 0x1.0p-52 + 1e-300), ec = fabs(c - cos(x)) / (cos(x) * 0x1.0p-52);
        if (es > worst_taylor) worst_taylor = es;
        if (ec > worst_taylor) worst_taylor = ec;
    }
    // the complex-time laser from its real-time parts (the SPA samplers' form, the carrier's exp
    // from the envelope's Taylor series when 2N is an integer) against the direct complex form
    double worst_tau = 0;
    for (int kind = 1; kind <= 2; ++kind)
    for (int cyc = 0; cyc < 4; ++cyc) {
        const double cycs[4] = {2.0, 1.5, 2.5, 3.7};
        etraj_laser_desc d{};
        d.n_colors = 1;
        d.color[0].kind = kind;
        d.color[0].peak_int = 2e14;
        d.color[0].wave_len = 800;
        d.color[0].cyc_num = cycs[cyc];
        d.color[0].ellip = 1.0;
        d.color[0].cep = 0.3;
        LaserP L;
        if (make_laser(d, L)) return 4;
        for (int i = 0; i < 20000; ++i) {
            const double tr = (2 * U(g) - 1) * L.c[0].halfw, tau = 45 * U(g);
            const LaserTr r = laser_tr(L, tr);
            Cx fx, fy, ax, ay, gx, gy, bx, by;
            laser_field_pot_tau(L, r, tr, tau, fx, fy, ax, ay);
            laser_field_pot_c(L, Cx{tr, tau}, gx, gy, bx, by);
            const double sF = fmax(hypot(gx.re, gx.im), hypot(gy.re, gy.im)) + 1e-300;
            const double sA = fmax(hypot(bx.re, bx.im), hypot(by.re, by.im)) + 1e-300;
            const double e = fmax(fmax(hypot(fx.re - gx.re, fx.im - gx.im), hypot(fy.re - gy.re, fy.im - gy.im)) / sF,
                                  fmax(hypot(ax.re - bx.re, ax.im - bx.im), hypot(ay.re - by.re, ay.im - by.im)) / sA);
            if (e > worst_tau) worst_tau = e;
        }
    }
    double worst_ch = 0, worst_sh = 0;
    for (int i = 0; i < 2000000; ++i) {
        const double s = (i & 3) == 0 ? 0.6 : (i & 3) == 1 ? 3.0 : (i & 3) == 2 ? 30.0 : 1e-3;
        const double x = (2 * U(g) - 1) * s;
        double c, sh;
        if (i & 4 && fabs(x) <= 0.5) cosh_sinh_small(x, c, sh);
        else cosh_sinh(x, c, sh);
        const double rc = cosh(x), rs = sinh(x);
        const double ec = fabs(c - rc) / rc, es = rs != 0 ? fabs(sh - rs) / fabs(rs) : fabs(sh);
        if (ec > worst_ch) worst_ch = ec;
        if (es > worst_sh) worst_sh = es;
    }
    printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", worst_field, worst_ch, worst_sh, worst_simple, worst_taylor,
           worst_tau);
    return 0;
}
"""


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("phys")
    (d / "phys.cpp").write_text(SRC)
    exe = d / "phys"
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    subprocess.run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(rocm, "include"),
                    "-I", os.path.join(ROOT, "etraj.jl_amd", "csrc"), str(d / "phys.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return [float(x) for x in out]


def test_stage_field_equals_the_direct_field(results):
    assert results[0] <= 1e-14


def test_cosh_sinh_from_one_exp(results):
    ulp = 2.0 ** -52
    assert results[1] <= 3 * ulp and results[2] <= 3 * ulp


def test_simple_laser_stage_fields_equal_the_direct_field(results):
    # the one-colour cos^n stage fields with the shorter series of Tsit5's stages 2 and 3 and
    # of the envelope (N >= 2), and the full sin/cos beyond their ranges
    assert results[3] <= 1e-14


def test_complex_time_laser_from_real_time_parts(results):
    # laser_field_pot_tau (envelope cosh/sinh by Taylor, the carrier's exp as a power of the
    # envelope's for 2N = 4, 3, 5) against laser_field_pot_c, relative to the largest component
    assert results[5] <= 1e-13


def test_taylor_sin_cos_orders_within_their_ranges(results):
    # orders 7, 5, 4, 3, 2 over |x| <= 1/2, 1/4, 1/8, 1/20, 1/100 against the C library, in ulp of
    # the value
    assert results[4] <= 2.0


# ---- the device's RK4 step, host-compiled, against the oracle's RK4 ------------------------

STEP_SRC = r"""
#include "etraj_physics.hpp"
#include <cstdio>
#include <cstdlib>
using namespace etraj;
// Integrates one electron (x y px py phase) from t0 to tf with the device's RK4 step sums in
// FP64: form 0 = hl_rk4_sums (hydrogen-like, exact soft-core phase rate), form 1 = gen_rk4_sums
// (the full EOM), the field evaluated at the stage times. Prints the final state.
int main(int argc, char** argv) {
    const int form = atoi(argv[1]);
    const double t0 = atof(argv[2]), tf = atof(argv[3]), dt = atof(argv[4]);
    double r[2] = {atof(argv[5]), atof(argv[6])}, p[2] = {atof(argv[7]), atof(argv[8])}, ph = atof(argv[9]);
    etraj_laser_desc d{};
    d.n_colors = 1;
    d.color[0].kind = ETRAJ_LASER_COS2; d.color[0].peak_int = 1e14; d.color[0].wave_len = 800; d.color[0].cyc_num = 2;
    etraj_target_desc td{};
    td.kind = ETRAJ_TARGET_HYDROGEN_LIKE; td.nucl_charge = 1; td.Ip = 0.5; td.asymp_coeff = 1; td.soft_core = 1e-10;
    LaserP L; TargetP T;
    if (make_laser(d, L) || make_target(td, T)) return 2;
    int64_t n = (int64_t)ceil((tf - t0) / dt); if (n < 1) n = 1;
    const double h = (tf - t0) / n, h6 = h / 6;
    const HlPhase<double> pc = hl_phase<double, ETRAJ_PHASE_SCTS, true>(h, T.Ip);
    for (int64_t s = 0; s < n; ++s) {
        const double t = t0 + s * h;
        double F0x, F0y, Fmx, Fmy, F1x, F1y, Sr[2], Sp[2], r2;
        laser_field(L, t, F0x, F0y);
        laser_field(L, t0 + (2 * s + 1) * (h / 2), Fmx, Fmy);
        laser_field(L, t0 + (2 * s + 2) * (h / 2), F1x, F1y);
        if (form == 0) {
            double SA, SQ;
            hl_rk4_sums<double, 2, ETRAJ_PHASE_SCTS, true, true>(T.Zd, T.sc, r, p, h, 0.5 * h, F0x, F0y, Fmx, Fmy, F1x, F1y,
                                                                   Sr, Sp, SA, SQ, r2);
            ph = fma(pc.cA, SA, fma(pc.cQ, SQ, ph));
        } else {
            double SPh;
            gen_rk4_sums<double, 2, ETRAJ_PHASE_SCTS, ETRAJ_TARGET_HYDROGEN_LIKE>(T, false, r, p, h, F0x, F0y, Fmx, Fmy,
                                                                              F1x, F1y, Sr, Sp, SPh, r2);
            ph = fma(h6, SPh, ph);
        }
        for (int i = 0; i < 2; ++i) { r[i] = fma(h6, Sr[i], r[i]); p[i] = fma(h6, Sp[i], p[i]); }
    }
    printf("%.17g %.17g %.17g %.17g %.17g\n", r[0], r[1], p[0], p[1], ph);
    return 0;
}
"""


@pytest.fixture(scope="module")
def step_exe(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("step")
    (d / "step.cpp").write_text(STEP_SRC)
    exe = d / "step"
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    # FMA contraction on, as the device compiler does
    subprocess.run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-mfma", "-ffp-contract=fast", "-D__HIP_PLATFORM_AMD__",
                    "-I", os.path.join(rocm, "include"), "-I", os.path.join(ROOT, "etraj.jl_amd", "csrc"),
                    str(d / "step.cpp"), "-o", str(exe)], check=True, capture_output=True)
    return str(exe)


def test_device_rk4_step_forms_match_the_oracle(step_exe):
    """The RK4 step in the arrangement the kernels use (stage sums Sr, Sp formed by FMAs; the
    hydrogen-like phase rate split into Z/sqrt(r2) (2 - a/r2) and |p|^2 parts), host-compiled in
    FP64 with FMA contraction, against the oracle's RK4 on the full (r, p, phase) system
    (TrajectorySimulation.jl:332-342 with the fixed step): tunnel exits of the H-atom LP pulse,
    integrated to t_final. Both forms agree with the oracle to 1e-11 relative in momentum and
    1e-9 rad in the SCTS phase (rounding differences over ~2000 steps)."""
    import math
    import sys
    sys.path[:0] = [os.path.join(ROOT, "etraj.jl_amd"), os.path.join(ROOT, "oracle")]
    import numpy as np
    import oracle as O
    from etraj_hip import Cos2Laser, get_atom
    from etraj_hip.simulation import run_desc
    laser = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    ld, td = laser._desc(), get_atom("H")._desc()
    tf = 2 * math.pi / (45.563352525 / 800.0) + 10.0
    rd = run_desc(dimension=2, traj_phase_method="SCTS", traj_t_final=tf, final_p_max=(1.5, 1.5),
                  final_p_num=(100, 100), traj_integrator="rk4", traj_dt=0.1, precision="fp64")
    rng = np.random.default_rng(7)
    worst_p = worst_ph = 0.0
    for t0 in (-80.0, -41.3, -3.0, 27.5, 55.0):
        Fx = O.laser_field(ld, [t0])[0][0]
        for kd in rng.uniform(-0.8, 0.8, 3):
            x0 = -(0.5 + kd * kd / 2) / Fx
            u0 = np.array([x0, 0.0, 0.0, kd, 0.0])
            ref, _, rc = O.integrate_state(ld, td, rd, t0, u0)
            assert rc == 0
            for form in (0, 1):
                out = subprocess.run([step_exe, str(form), "%.17g" % t0, "%.17g" % tf, "0.1"] + ["%.17g" % v for v in u0],
                                     capture_output=True, text=True, check=True).stdout.split()
                got = np.array([float(v) for v in out])
                worst_p = max(worst_p, np.linalg.norm(got[2:4] - ref[2:4]) / np.linalg.norm(ref[2:4]))
                worst_ph = max(worst_ph, abs(got[4] - ref[4]))
    assert worst_p < 1e-11, worst_p
    assert worst_ph < 1e-9, worst_ph
