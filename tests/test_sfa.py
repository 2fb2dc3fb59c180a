"""SFA-SPA (SPASampler.jl) and SPANE (SPANESampler.jl) samplers.

CPU pins of the oracle:
  * the reference's survivor-count KATs (ElectronSamplers_test.jl:133-291,
    atomic H: 14 cases) reproduced exactly;
  * dF/dt (the SPANE input) against 8th-order central differences of F;
  * the complex-time laser: equal to the real one on the real axis, and
    dA/d(tau) = -i F along t = tr + i tau (A is analytic in t there);
  * the SPA saddle point solves the full complex equation
    (p + A(ts))^2 + kz^2 = -2 Ip, and its integrals agree with scipy's
    adaptive quadrature of the same integrands;
  * for a long linearly polarized pulse at a field crest the SPA saddle point
    and tunnelling exponent reduce to the Keldysh/PPT closed forms
    sinh(w ti) = sqrt(2Ip + k^2)/A and G = (Ip + k^2/2) ti - (A^2/2)(sinh(2 w ti)/(4w) - ti/2);
  * the finite-difference Jacobian of the reference (steps 0.01) agrees with a
    Richardson-extrapolated derivative of the saddle-point momenta to O(h^2);
  * SPANE equals ADK for linear polarization (F x dF/dt = 0 and k_perp . dF/dt = 0);
  * atoms with l > 0 (Ar) are sampled (their prefactor is pinned in test_mol_prefactor.py).
GPU tests (marked gpu): device sampling and a whole SPA job against the oracle.
"""
import math
import warnings

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from cases import l1
from etraj_hip import (_abi, BichromaticLaser, Cos2Laser, Cos4Laser, GaussianLaser, get_atom)
from etraj_hip.samplers import init_sampler
from etraj_hip.simulation import run_desc

KAT_BASE = dict(dimension=2, sample_t_intv=(-100, 100), sample_t_num=400, sample_cutoff_limit=1e-14,
                sample_monte_carlo=False, traj_phase_method="CTMC", rate_prefix="Full", ss_kd_max=2.0,
                ss_kd_num=800, ss_kz_max=2.0, ss_kz_num=0, mc_kt_num=800, mc_kd_max=2.0, mc_kz_max=2.0)
_3D = {"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30}
KATS = [  # ElectronSamplers_test.jl:133-291: size(gen_electron_batch(sp, 200)) for get_atom("H")
    ("SPA", {}, (6, 633)), ("SPA", {"traj_phase_method": "QTMC"}, (7, 633)),
    ("SPA", {"traj_phase_method": "SCTS"}, (7, 633)), ("SPA", _3D, (8, 446)),
    ("SPA", dict(_3D, traj_phase_method="QTMC"), (9, 446)), ("SPA", dict(_3D, traj_phase_method="SCTS"), (9, 446)),
    ("SPA", {"rate_prefix": "Exp"}, (6, 611)), ("SPA", {"rate_prefix": "Pre"}, (6, 617)),
    ("SPA", {"rate_prefix": "PreCC"}, (6, 637)), ("SPA", {"rate_prefix": "Jac"}, (6, 607)),
    ("SPA", {"rate_prefix": {"Pre", "Jac"}}, (6, 613)),
    ("SPANE", {}, (6, 617)), ("SPANE", {"traj_phase_method": "QTMC"}, (7, 617)),
    ("SPANE", {"traj_phase_method": "SCTS"}, (7, 617)), ("SPANE", _3D, (8, 424)),
    ("SPANE", dict(_3D, traj_phase_method="SCTS"), (9, 424)),
    ("SPANE", {"rate_prefix": "Exp"}, (6, 573)), ("SPANE", {"rate_prefix": "Pre"}, (6, 590)),
    ("SPANE", {"rate_prefix": "PreCC"}, (6, 627)), ("SPANE", {"rate_prefix": "Jac"}, (6, 561)),
    ("SPANE", {"rate_prefix": {"Pre", "Jac"}}, (6, 579)),
]


def _kat(method, over, laser=None, target=None):
    l = laser or Cos4Laser(peak_int=1e15, wave_len=800.0, cyc_num=2, ellip=1.0)
    t = target or get_atom("H")
    kw = dict(KAT_BASE)
    kw.update(over)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method=method, laser=l, target=t, **kw)
    return l, t, sp


@pytest.mark.parametrize("method,over,expect", KATS)
def test_sfa_sampler_kats(method, over, expect):
    l, t, sp = _kat(method, over)
    got = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 200)
    assert (got.shape[1], got.shape[0]) == expect


def _lasers():
    return [Cos2Laser(peak_int=4e14, wave_len=800.0, cyc_num=2.0, ellip=1.0, azi=math.pi / 4, cep=math.pi, t_shift=10.0),
            Cos4Laser(peak_int=4e14, wave_len=800.0, cyc_num=2.0, ellip=0.5, azi=0.3, cep=1.0, t_shift=-5.0),
            GaussianLaser(peak_int=4e14, wave_len=800.0, spread_cyc_num=2.0, ellip=1.0, azi=math.pi / 4, cep=math.pi,
                          t_shift=10.0),
            BichromaticLaser(Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=8, ellip=1.0),
                             Cos2Laser(peak_int=2e14, wave_len=400.0, cyc_num=16, ellip=-1.0), 12.0)]


@pytest.mark.parametrize("k", range(4))
def test_field_derivative_matches_central_differences(k):
    ld = _lasers()[k]._desc()
    h = 1e-3
    c = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])
    for t in np.linspace(-90, 90, 19):
        Fx, Fy, dFx, dFy = O.laser_field_deriv(ld, t)
        f = O.laser_field(ld, t + h * np.arange(-4, 5))
        assert dFx == pytest.approx(c @ f[0] / h, rel=1e-7, abs=1e-11)
        assert dFy == pytest.approx(c @ f[1] / h, rel=1e-7, abs=1e-11)
        F = O.laser_field(ld, np.array([t]))
        assert (Fx, Fy) == (pytest.approx(F[0][0], rel=1e-15, abs=1e-300), pytest.approx(F[1][0], rel=1e-15, abs=1e-300))


@pytest.mark.parametrize("k", range(4))
def test_complex_time_laser_is_analytic_continuation(k):
    ld = _lasers()[k]._desc()
    for tr in np.linspace(-60, 60, 7):
        Fx, Fy, Ax, Ay = O.laser_complex(ld, complex(tr, 0.0))
        F = O.laser_field(ld, np.array([tr]))
        assert abs(Fx - F[0][0]) <= 1e-14 * (abs(F[0][0]) + 1e-3) and abs(Fy - F[1][0]) <= 1e-14 * (abs(F[1][0]) + 1e-3)
        assert abs(Ax - F[2][0]) <= 1e-14 * (abs(F[2][0]) + 1e-2) and abs(Ay - F[3][0]) <= 1e-14 * (abs(F[3][0]) + 1e-2)
        for tau in (3.0, 11.0):
            h = 1e-4
            ap = O.laser_complex(ld, complex(tr, tau + h))
            am = O.laser_complex(ld, complex(tr, tau - h))
            f = O.laser_complex(ld, complex(tr, tau))
            for i in (2, 3):
                dA = (ap[i] - am[i]) / (2 * h)
                assert abs(dA - (-1j) * f[i - 2]) <= 1e-6 * (abs(f[i - 2]) + 1e-6)


SPA_LASERS = [Cos4Laser(peak_int=1e15, wave_len=800.0, cyc_num=2, ellip=1.0),
              Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0),
              Cos2Laser(peak_int=2e14, wave_len=1300.0, cyc_num=4, ellip=0.4, azi=0.5)]


@pytest.mark.parametrize("k", range(3))
def test_spa_saddle_point_and_integrals(k):
    l = SPA_LASERS[k]
    ld, td = l._desc(), get_atom("H")._desc()
    Ip = 0.5
    n = 0
    for tr in (-40.0, -13.0, 0.5, 21.0):
        for kd, kz in ((0.4, 0.0), (-0.9, 0.3), (1.3, -0.6)):
            P = O.spa_point(ld, td, tr, kd, kz)
            if P is None:
                continue
            n += 1
            ti = P["ti"]
            _, _, Ax, Ay = O.laser_complex(ld, complex(tr, ti))
            # the full complex saddle-point equation
            lhs = (P["px"] + Ax) ** 2 + (P["py"] + Ay) ** 2 + kz * kz
            assert abs(lhs + 2 * Ip) <= 1e-7
            # integrals by scipy's adaptive quadrature
            f = lambda tau, i: O.laser_complex(ld, complex(tr, tau))[i]
            x0 = integrate.quad(lambda tau: f(tau, 2).imag, 0, ti, epsabs=0, epsrel=1e-12)[0]
            y0 = integrate.quad(lambda tau: f(tau, 3).imag, 0, ti, epsabs=0, epsrel=1e-12)[0]

            def g(tau):
                _, _, ax, ay = O.laser_complex(ld, complex(tr, tau))
                return ((P["px"] + ax) ** 2 + (P["py"] + ay) ** 2 + kz * kz) / 2 + Ip
            Gr = integrate.quad(lambda tau: g(tau).real, 0, ti, epsabs=0, epsrel=1e-12)[0]
            Gi = integrate.quad(lambda tau: g(tau).imag, 0, ti, epsabs=0, epsrel=1e-12)[0]
            assert P["x0"] == pytest.approx(x0, rel=1e-10, abs=1e-10)
            assert P["y0"] == pytest.approx(y0, rel=1e-10, abs=1e-10)
            assert P["G"] == pytest.approx(complex(Gr, Gi), rel=1e-10, abs=1e-10)
            # momentum at the tunnel exit: p + A(tr)
            _, _, Axr, Ayr = O.laser_complex(ld, complex(tr, 0.0))
            assert (P["kx0"], P["ky0"]) == (pytest.approx(P["px"] + Axr.real), pytest.approx(P["py"] + Ayr.real))
    assert n >= 8


def test_spa_reduces_to_keldysh_ppt_for_a_long_linear_pulse():
    w = 45.563352525 / 800.0
    l = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=40, ellip=0.0)
    ld, td = l._desc(), get_atom("H")._desc()
    tr = math.pi / (2 * w)                    # a crest of F: A(tr) = 0
    Ip = 0.5
    env = math.cos(w * tr / 80) ** 2
    A = math.sqrt(1e14 / 3.50944521e16) / w * env
    for kd in (0.05, 0.4, 0.8):
        P = O.spa_point(ld, td, tr, kd, 0.0)
        k2 = kd * kd
        ti = math.asinh(math.sqrt(2 * Ip + k2) / A) / w
        G = (Ip + k2 / 2) * ti - A * A / 2 * (math.sinh(2 * w * ti) / (4 * w) - ti / 2)
        assert P["ti"] == pytest.approx(ti, rel=2e-3)
        assert P["G"].real == pytest.approx(G, rel=2e-3)
        assert abs(P["G"].imag) <= 2e-3 * abs(G)
        assert P["kx0"] == pytest.approx(0.0, abs=2e-3) and P["ky0"] == pytest.approx(kd, rel=2e-3)


def test_spa_jacobian_matches_a_converged_derivative():
    l = SPA_LASERS[0]
    ld, td = l._desc(), get_atom("H")._desc()
    for tr, kd in ((-13.0, 0.4), (2.0, -0.8), (20.0, 1.1)):
        P = O.spa_point(ld, td, tr, kd, 0.0)

        def pvec(t, k):
            q = O.spa_point(ld, td, t, k, 0.0)
            return np.array([q["px"], q["py"]])

        def d(fn, h):   # Richardson-extrapolated central difference
            return (4 * (fn(h) - fn(-h)) / (2 * h) - (fn(2 * h) - fn(-2 * h)) / (4 * h)) / 3
        dt = d(lambda h: pvec(tr + h, kd), 1e-3)
        dk = d(lambda h: pvec(tr, kd + h), 1e-3)
        jac = abs(dt[0] * dk[1] - dk[0] * dt[1])
        assert P["jac"] == pytest.approx(jac, rel=1e-3)


def test_spane_equals_adk_for_linear_polarization():
    l = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    over = {"sample_cutoff_limit": 0.0, "sample_t_num": 64, "ss_kd_num": 120, "ss_kd_max": 1.0}
    n = 0
    for b in (17, 30, 41):
        l_, t, spa = _kat("SPANE", over, laser=l)
        _, _, adk = _kat("ADK", over, laser=l)
        ra = O.gen_electron_batch(l._desc(), t._desc(), adk.desc, b)
        rs = O.gen_electron_batch(l._desc(), t._desc(), spa.desc, b)
        assert ra.shape == rs.shape
        # equal up to rounding of F vs sqrt(F^2), amplified by the exponent (up to ~1e3 at the pulse edges)
        np.testing.assert_allclose(rs, ra, rtol=1e-10, atol=1e-300)
        n += len(ra)
    assert n > 100


def test_sfa_samplers_accept_l_gt_0():
    # Ar (l = 1): the Wigner-D / Y_lm prefactor (tests/test_mol_prefactor.py pins it)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ar = get_atom("Ar")
    for method in ("SPA", "SPANE"):
        l, _, sp = _kat(method, {}, target=ar)
        got = O.gen_electron_batch(l._desc(), ar._desc(), sp.desc, 200)
        assert got is not None and np.isfinite(got).all()


# ---- device ----------------------------------------------------------------------

def _device_batch(ctx, ld, td, sd, rows, b):
    cap = max(int(ctx.lib.etraj_batch_max_size(sd)), 1)
    out = np.empty((cap, rows))
    n = _abi.C.c_int64()
    ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(out), cap, _abi.C.byref(n)))
    return out[:n.value]


@pytest.mark.gpu
@pytest.mark.parametrize("method,over,expect", KATS)
def test_sfa_kats_on_device(ctx, method, over, expect):
    l, t, sp = _kat(method, over)
    got = _device_batch(ctx, l._desc(), t._desc(), sp.desc, sp.rows, 200)
    assert (got.shape[1], got.shape[0]) == expect
    ref = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 200)
    rtol = 1e-9 if method == "SPA" else 1e-12
    np.testing.assert_allclose(got[:, :-1] if sp.rows in (7, 9) else got, ref[:, :-1] if sp.rows in (7, 9) else ref,
                               rtol=rtol, atol=1e-12)
    if sp.rows in (7, 9):   # the phase angle(amp), compared modulo 2 pi
        d = np.angle(np.exp(1j * (got[:, -1] - ref[:, -1])))
        assert np.max(np.abs(d)) <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["SPA", "SPANE"])
@pytest.mark.parametrize("over", [{"sample_monte_carlo": True, "seed": 5},
                                  {"sample_monte_carlo": True, "dimension": 3, "seed": 5},
                                  {"sample_stratified": True, "kd_sigma": 0.5, "mc_kt_num": 300, "seed": 5,
                                   "traj_phase_method": "SCTS"}])
def test_sfa_seeded_sampling_on_device(ctx, method, over):
    l, t, sp = _kat(method, over, laser=SPA_LASERS[2])
    n = 0
    for b in (120, 200, 260):
        got = _device_batch(ctx, l._desc(), t._desc(), sp.desc, sp.rows, b)
        ref = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, b)
        ref = np.zeros((0, sp.rows)) if ref is None else ref
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9)
        n += len(ref)
    assert n > 0


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["SPA", "SPANE"])
def test_sfa_run_job_matches_oracle(ctx, method):
    w = 45.563352525 / 800.0
    half = 2 * math.pi / w
    kw = dict(laser=Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.5), target=get_atom("H"),
              dimension=2, sample_t_intv=(-half, half), sample_t_num=60, traj_t_final=half + 10.0,
              final_p_max=(1.5, 1.5), final_p_num=(150, 150), ss_kd_max=1.2, ss_kd_num=40,
              traj_phase_method="QTMC", rate_prefix="Full", sample_cutoff_limit=1e-16, traj_dt=0.1,
              precision="fp64")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method=method, **kw)
    ld, td, sd, rd = kw["laser"]._desc(), kw["target"]._desc(), sp.desc, run_desc(**kw)
    out = np.empty(150 * 150)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st.num_effective_traj == rst["num_effective_traj"] > 0
    assert l1(out, ref) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("dim,kd_num,kz_num", [(2, 3, 1), (3, 7, 5), (3, 2, 1)], ids=["2d-nk3", "3d-nk35", "3d-nk2"])
def test_spa_small_batches_run_job_matches_oracle(ctx, dim, kd_num, kz_num):
    """SFA-SPA batches far smaller than one wavefront (2 to 35 samples), in 2D
    and 3D, with fixed-point PMD sums: the same electrons, ion probability and
    PMD as the oracle."""
    w = 45.563352525 / 800.0
    half = 2 * math.pi / w
    p_num = (40,) * dim
    kw = dict(laser=Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=1.0), target=get_atom("H"),
              dimension=dim, sample_t_intv=(-half, half), sample_t_num=50, traj_t_final=half + 10.0,
              final_p_max=(1.5,) * dim, final_p_num=p_num, ss_kd_max=1.2, ss_kd_num=kd_num,
              traj_phase_method="QTMC", rate_prefix="Full", sample_cutoff_limit=1e-16, traj_dt=0.1,
              precision="fp64", pmd_accumulator="fixed")
    if dim == 3:
        kw.update(ss_kz_max=1.0, ss_kz_num=kz_num)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method="SPA", **kw)
    ld, td, sd, rd = kw["laser"]._desc(), kw["target"]._desc(), sp.desc, run_desc(**kw)
    out = np.empty(int(np.prod(p_num)))
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st.num_effective_traj == rst["num_effective_traj"] > 0
    assert st.ion_prob == pytest.approx(rst["ion_prob"], rel=1e-9)
    assert l1(out, ref) <= 1e-6
