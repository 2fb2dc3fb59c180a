"""The phase of a complex tunnelling prefactor against its closed form.

The reference's prefactor of an SAE atom with l > 0 (ADKSampler.jl:232-262,
shared_methods.jl:15-24) evaluates Y_lm of gen_sph_harm_funcs, which divides
the solid harmonic by (-i kappa)^l, at the field-frame components of the
saddle-point momentum k_ts = (kx - i ti Fx, ky - i ti Fy, kz) taken with
`kts ⋅ axis`, and LinearAlgebra.dot conjugates its first argument. For l = 1
the sum over k of the solid harmonic has a single term, so the ratio of the
l = 1 amplitude to the l = 0 amplitude of the same Ip, Z, laser and sample
(every other factor of the amplitude is shared, asymp_coeff = 1 for both) is,
with n = conj(k_ts) in the atom frame,
    m = 0:   sqrt(3) n_z / (-i kappa),
    m = +-1: -+sqrt(3/2) (n_x +- i n_y) / (-i kappa)
(Condon-Shortley Y_1m; Y_00 = 1/sqrt(4 pi)). ADK's sample momentum is
perpendicular to the field and ti |F| = sqrt(kappa^2 + k^2) (the saddle-point
condition k_ts . k_ts = -kappa^2), so both ratios are real here:
  * axis along the linear polarization x (axis parallel or antiparallel to the
    field, so the result does not depend on how the Euler angles of the
    field-frame -> atom-frame rotation are read): R = -sqrt(3) ti Fx / kappa
    (up to the kx ~ 1e-16 of the sample), negative where Fx > 0;
  * m = +-1 with the axis along z, for the samples with Fx > 0 (there the
    field-frame axes (z, -y, x) make a rotation that is its own inverse):
    R = +-sqrt(3/2) (ti Fx +- ky) / kappa, so m = +1 and m = -1 differ in
    magnitude, mirrored in ky.
Dropping the conjugation flips the sign of Im n_z and swaps the two m = +-1
magnitudes; (+i kappa)^l instead of (-i kappa)^l flips every sign. Copies of
the oracle with either change fail the CPU test below; the GPU test holds the
device sampler to the same closed forms.
"""
import math
import os
import shutil
import subprocess
import warnings

import numpy as np
import pytest

import oracle as O
from etraj_hip import _abi, Cos2Laser
from etraj_hip.samplers import init_sampler
from etraj_hip.targets import HydrogenLikeAtom

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IP = 0.579
KAPPA = math.sqrt(2 * IP)
N_T = 24


def _setup(dimension):
    laser = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    targets = {
        "s": HydrogenLikeAtom(Ip=IP, Z=1, l=0, asymp_coeff=1.0),
        "p0x": HydrogenLikeAtom(Ip=IP, Z=1, l=1, m=0, asymp_coeff=1.0, quan_ax_θ=math.pi / 2, quan_ax_ϕ=0.0),
        "p+1": HydrogenLikeAtom(Ip=IP, Z=1, l=1, m=1, asymp_coeff=1.0),
        "p-1": HydrogenLikeAtom(Ip=IP, Z=1, l=1, m=-1, asymp_coeff=1.0),
    }
    samplers = {}
    for k, t in targets.items():
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            samplers[k] = init_sampler(
                init_cond_method="ADK", laser=laser, target=t, dimension=dimension, sample_t_intv=(-90, 90),
                sample_t_num=N_T, sample_cutoff_limit=0.0, sample_monte_carlo=False, traj_phase_method="QTMC",
                rate_prefix="Full", ss_kd_max=1.5, ss_kd_num=40, ss_kz_max=1.0, ss_kz_num=7 if dimension == 3 else 0,
                mc_kt_num=1, mc_kd_max=0.0, mc_kz_max=0.0)
    return laser, targets, samplers


def _amplitudes(cols, dimension):
    """(sample columns, sign of Fx, complex amplitude) of a batch"""
    d = dimension
    amp = np.sqrt(cols[:, 2 * d + 1]) * np.exp(1j * cols[:, 2 * d + 2])
    # the tunnel exit lies opposite the field (ADKSampler.jl:173): sign(Fx) = -sign(x0)
    return cols, -np.sign(cols[:, 0]), amp


def _expected(cols, sfx, dimension, case):
    d = dimension
    kx, ky = cols[:, d], cols[:, d + 1]
    k2 = kx**2 + ky**2 + (cols[:, d + 2] ** 2 if d == 3 else 0.0)
    tiF = sfx * np.sqrt(KAPPA**2 + k2)                 # ti Fx, with Fy = 0
    nx, ny = kx + 1j * tiF, ky + 0j                     # conj(k_ts), lab (= atom for the z axis)
    if case == "p0x":
        return math.sqrt(3) * nx / (-1j * KAPPA)
    if case == "p+1":
        return -math.sqrt(1.5) * (nx + 1j * ny) / (-1j * KAPPA)
    return math.sqrt(1.5) * (nx - 1j * ny) / (-1j * KAPPA)


def _check(batch, dimension):
    """max relative deviation of the l = 1 / l = 0 amplitude ratios from the closed forms, and
    the numbers of samples compared (batch(key, b) -> columns or None)"""
    worst, counted = 0.0, {"p0x": 0, "p+1": 0, "p-1": 0}
    for b in range(1, N_T + 1):
        s = batch("s", b)
        if s is None or len(s) == 0:
            continue
        cols, sfx, a0 = _amplitudes(s, dimension)
        keep = np.abs(cols[:, 1]) < 1e-9 * np.abs(cols[:, 0])   # linear polarization: the exit is on the x axis
        assert keep.all()
        for case in counted:
            c1 = batch(case, b)
            assert c1.shape == s.shape
            np.testing.assert_array_equal(c1[:, :2 * dimension + 1], s[:, :2 * dimension + 1])
            _, _, a1 = _amplitudes(c1, dimension)
            # rates far from the field peaks are subnormal or 0 in FP64: keep normal ones
            sel = (s[:, 2 * dimension + 1] > 1e-290) & (c1[:, 2 * dimension + 1] > 1e-290)
            if case != "p0x":
                sel &= sfx > 0
            if not sel.any():
                continue
            exp = _expected(cols[sel], sfx[sel], dimension, case)
            got = a1[sel] / a0[sel]
            dev = np.abs(got - exp) / np.abs(exp)
            assert np.isfinite(dev).all(), case
            worst = max(worst, float(dev.max()))
            counted[case] += int(sel.sum())
    return worst, counted


def _oracle_batch(library=None, dimension=2):
    laser, targets, samplers = _setup(dimension)

    def batch(key, b):
        return O.gen_electron_batch(laser._desc(), targets[key]._desc(), samplers[key].desc, b, library=library)
    return batch


@pytest.mark.parametrize("dimension", [2, 3])
def test_oracle_prefactor_matches_closed_form(dimension):
    worst, counted = _check(_oracle_batch(dimension=dimension), dimension)
    assert min(counted.values()) > 100, counted
    assert worst < 1e-12, worst


MUTANTS = {
    "unconjugated k_ts": ("double complex c0 = conj(kts[0]), c1 = conj(kts[1]), c2 = conj(kts[2]);",
                          "double complex c0 = kts[0], c1 = kts[1], c2 = kts[2];"),
    "(+i kappa)^l": ("return nrm / cipow(-I * kappa, l) * tgamma(l + m + 1.0) * s;",
                     "return nrm / cipow(I * kappa, l) * tgamma(l + m + 1.0) * s;"),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_prefactor_mutants_fail_the_closed_form(name, tmp_path):
    """A copy of the oracle with the conjugation of k_ts dropped, or with (+i kappa)^l, is built
    on the CPU and must fail the closed-form check that the oracle passes."""
    old, new = MUTANTS[name]
    src = open(os.path.join(ROOT, "oracle", "etraj_oracle.c")).read()
    assert src.count(old) == 1, name
    path = tmp_path / "mutant.c"
    path.write_text(src.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-shared", "-I",
                    os.path.join(ROOT, "oracle"), "-o", str(so), str(path), "-lm", "-lpthread"], check=True)
    worst, counted = _check(_oracle_batch(O.load_variant(str(so))), 2)
    assert min(counted.values()) > 100, counted
    assert worst > 0.1, (name, worst)


@pytest.mark.gpu
@pytest.mark.parametrize("dimension", [2, 3])
def test_device_prefactor_matches_closed_form(ctx, dimension):
    laser, targets, samplers = _setup(dimension)

    def batch(key, b):
        sp = samplers[key]
        cap = max(int(ctx.lib.etraj_batch_max_size(sp.desc)), 1)
        out = np.empty((cap, sp.rows))
        n = _abi.C.c_int64()
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, laser._desc(), targets[key]._desc(), sp.desc, b,
                                                   _abi.dptr(out), cap, _abi.C.byref(n)))
        return out[:n.value]
    worst, counted = _check(batch, dimension)
    assert min(counted.values()) > 100, counted
    assert worst < 1e-12, worst
