"""Host-compiled checks of the device math that re-arranges a formula of the
reference for speed (etraj.jl_amd/csrc/etraj_physics.hpp), against the plain
form evaluated with the C library:

  * laser_field_at (the laser at t + d from sin/cos at t by angle addition, the
    cut-off window as logistic exps) equals laser_field(t + d)
    (Cos2Laser.jl:125-152 and siblings) to 1e-14 of the peak field, for every
    envelope, ellipticity, azimuth, CEP, time shift, two colours, increments up
    to 20 a.u. (past the Taylor range of the rotation) and times over the pulse
    and its edges;
  * cosh_sinh (both from one exp, sinh by its Taylor series below 1/2) equals
    cosh / sinh to 3 ulp over |x| up to 30.
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
    double worst_ch = 0, worst_sh = 0;
    for (int i = 0; i < 2000000; ++i) {
        const double s = (i & 3) == 0 ? 0.6 : (i & 3) == 1 ? 3.0 : (i & 3) == 2 ? 30.0 : 1e-3;
        const double x = (2 * U(g) - 1) * s;
        double c, sh;
        cosh_sinh(x, c, sh);
        const double rc = cosh(x), rs = sinh(x);
        const double ec = fabs(c - rc) / rc, es = rs != 0 ? fabs(sh - rs) / fabs(rs) : fabs(sh);
        if (ec > worst_ch) worst_ch = ec;
        if (es > worst_sh) worst_sh = es;
    }
    printf("%.17g %.17g %.17g\n", worst_field, worst_ch, worst_sh);
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
