"""WFAT sampler (WFATSampler.jl:131-249) with a tabulated G00 structure factor.

CPU pins of the oracle:
  * the FF -> MF Euler angles (shared_methods.jl:31-38) rebuild the rotation
    MF^-1 FF they were taken from, and cos(beta) is the projection of the
    molecular z axis on the field direction;
  * G00(beta, gamma) = sum c_ab P_a(cos beta) e^{-i b gamma} agrees with
    numpy's Legendre series;
  * with G = 1 the WFAT and ADK(:Exp) samplers of the same H-like target give
    the same exit points and momenta, and the rate ratio is the WFAT field
    factor W_00(F) = (kappa/2)(4 kappa^2/F)^(2n-1) (Tolstikhin, Morishita,
    Madsen, PRA 84, 053423 (2011)) times kappa/F, the k_perp -> 0 limit of the
    modified field factor (docs theory1_init_cond.md, WFAT section); |G|^2
    scales the rates.
GPU tests (marked gpu): device sampler and whole jobs against the oracle.
"""
import math
import warnings

import numpy as np
import pytest
from numpy.polynomial import legendre

import oracle as O
from cases import descriptors, l1, small_wfat
from etraj_hip import _abi, Cos2Laser, HydrogenLikeAtom, TabulatedMolecule
from etraj_hip.samplers import ADKSampler, WFATSampler, init_sampler


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def _ry(b):
    c, s = math.cos(b), math.sin(b)
    return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])


def _zyz(a, b, c):
    return _rz(a) @ _ry(b) @ _rz(c)


@pytest.mark.parametrize("seed", range(6))
def test_ff_mf_euler_rebuilds_the_rotation(seed):
    rng = np.random.default_rng(seed)
    rot = rng.uniform(-math.pi, math.pi, 3)
    th = rng.uniform(-math.pi, math.pi)
    F = rng.uniform(0.01, 0.2)
    Fx, Fy = F * math.cos(th), F * math.sin(th)
    a, b, c = O.ff_mf_euler(rot, Fx, Fy)
    MF = _zyz(*rot)
    FF = np.array([[0, Fy / F, Fx / F], [0, -Fx / F, Fy / F], [1.0, 0, 0]])
    np.testing.assert_allclose(_zyz(a, b, c), np.linalg.solve(MF, FF), atol=1e-13)
    assert 0.0 <= b <= math.pi
    z_mf = MF[:, 2]
    assert math.cos(b) == pytest.approx(z_mf @ np.array([Fx, Fy, 0.0]) / F, abs=1e-13)


def test_wfat_g00_is_a_legendre_fourier_series():
    rng = np.random.default_rng(7)
    c = rng.standard_normal((8, 16)) + 1j * rng.standard_normal((8, 16))
    mol = TabulatedMolecule(Ip=0.45, g00_coeff=c)
    sd = _abi.SamplerDesc()
    coeff = mol.coeff_table()
    sd.wfat_na, sd.wfat_nb = c.shape
    sd.wfat_coeff = _abi.dptr(coeff)
    for beta, gamma in rng.uniform([0, -math.pi], [math.pi, math.pi], (20, 2)):
        pa = legendre.legval(math.cos(beta), np.eye(8))        # P_0..P_7(cos beta)
        ref = np.sum(c * pa[:, None] * np.exp(-1j * np.arange(16) * gamma)[None, :])
        assert O.wfat_g00(sd, beta, gamma) == pytest.approx(ref, rel=1e-13, abs=1e-13)


def _pair(g00, dim=2, cutoff=0.0, **over):
    laser = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    kw = dict(laser=laser, dimension=dim, sample_t_intv=(-100.0, 100.0), sample_t_num=64, sample_cutoff_limit=cutoff,
              traj_phase_method="CTMC", ss_kd_max=1.0, ss_kd_num=40, ss_kz_max=0.8, ss_kz_num=7 if dim == 3 else 0)
    kw.update(over)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        adk = ADKSampler(target=HydrogenLikeAtom(Ip=0.5, Z=1), rate_prefix="Exp", **kw)
        mol = TabulatedMolecule(Ip=0.5, charge=0, g00_coeff=np.array([[g00]]), rot_α=0.4, rot_β=1.2, rot_γ=-0.3)
        wfat = WFATSampler(target=mol, **kw)
    return laser, adk, wfat, mol


@pytest.mark.parametrize("dim", [2, 3])
def test_wfat_over_adk_exp_is_the_wfat_field_factor(dim):
    laser, adk, wfat, mol = _pair(1.0, dim)
    h = HydrogenLikeAtom(Ip=0.5, Z=1)
    kappa, n = 1.0, 1.0
    tested = 0
    for b in (20, 28, 33, 41):
        ra = O.gen_electron_batch(laser._desc(), h._desc(), adk.desc, b)
        rw = O.gen_electron_batch(laser._desc(), mol._desc(), wfat.desc, b)
        if dim == 3:    # ADK drops |kd| < 1e-4 (whatever kz), WFAT kd^2 + kz^2 < 1e-8: no such point here
            assert ra.shape == rw.shape
        np.testing.assert_array_equal(ra[:, :-1], rw[:, :-1])        # exit point, momentum, t0
        tr = ra[0, -2]
        Fx, Fy, _, _ = O.laser_field(laser._desc(), np.array([tr]))
        F = math.hypot(Fx[0], Fy[0])
        W00 = kappa / 2 * (4 * kappa ** 2 / F) ** (2 * n - 1)
        np.testing.assert_allclose(rw[:, -1] / ra[:, -1], W00 * kappa / F, rtol=1e-12)
        tested += len(ra)
    assert tested > 100


def test_wfat_rate_scales_with_abs_g_squared():
    laser, _, w1, m1 = _pair(1.0)
    _, _, w2, m2 = _pair(2.0 - 1.5j)
    r1 = O.gen_electron_batch(laser._desc(), m1._desc(), w1.desc, 30)
    r2 = O.gen_electron_batch(laser._desc(), m2._desc(), w2.desc, 30)
    np.testing.assert_allclose(r2[:, -1], r1[:, -1] * abs(2.0 - 1.5j) ** 2, rtol=1e-14)


def test_wfat_seeded_modes_draw_the_adk_points():
    for over in ({"sample_monte_carlo": True, "mc_kt_num": 300, "mc_kd_max": 1.0, "seed": 11},
                 {"sample_stratified": True, "mc_kt_num": 300, "kd_sigma": 0.4, "seed": 11}):
        laser, adk, wfat, mol = _pair(1.0, **over)
        h = HydrogenLikeAtom(Ip=0.5, Z=1)
        ra = O.gen_electron_batch(laser._desc(), h._desc(), adk.desc, 25)
        rw = O.gen_electron_batch(laser._desc(), mol._desc(), wfat.desc, 25)
        np.testing.assert_array_equal(ra[:, :-1], rw[:, :-1])


def test_wfat_sampler_interface():
    laser = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    with pytest.raises(TypeError):
        WFATSampler(laser=laser, target=HydrogenLikeAtom(Ip=0.5, Z=1), dimension=2, sample_t_intv=(-1, 1),
                    sample_t_num=10, ss_kd_max=1.0, ss_kd_num=10)
    mol = TabulatedMolecule(Ip=0.45, g00_coeff=np.ones((2, 3)))
    with pytest.raises(NotImplementedError):
        ADKSampler(laser=laser, target=mol, dimension=2, sample_t_intv=(-1, 1), sample_t_num=10, ss_kd_max=1.0,
                   ss_kd_num=10)
    with pytest.raises(AssertionError):
        TabulatedMolecule(Ip=-0.1, g00_coeff=np.ones((1, 1)))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method="WFAT", laser=laser, target=mol, dimension=2, sample_t_intv=(-1, 1),
                          sample_t_num=10, ss_kd_max=1.0, ss_kd_num=10)
    assert isinstance(sp, WFATSampler) and sp.rows == 6
    assert (sp.desc.wfat_na, sp.desc.wfat_nb) == (2, 3)
    assert np.ctypeslib.as_array(sp.desc.wfat_coeff, (12,)).tolist() == [1.0, 0.0] * 6


def test_wfat_oracle_job_is_deterministic_across_threads():
    kw = small_wfat(n_t=16, n_k=64)
    ld, td, sd, rd = descriptors(kw)
    a, sa = O.run_job(ld, td, sd, rd, nthreads=1)
    b, sb = O.run_job(ld, td, sd, rd, nthreads=4)
    assert sa["num_effective_traj"] == sb["num_effective_traj"] > 0
    assert l1(b, a) < 1e-12


# ---- device ----------------------------------------------------------------------

def _device_batch(ctx, ld, td, sd, rows, b):
    cap = max(int(ctx.lib.etraj_batch_max_size(sd)), 1)
    out = np.empty((cap, rows))
    n = _abi.C.c_int64()
    ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(out), cap, _abi.C.byref(n)))
    return out[:n.value]


D3 = {"dimension": 3, "final_p_max": (2.0, 2.0, 1.0), "final_p_num": (40, 40, 10)}


@pytest.mark.gpu
@pytest.mark.parametrize("over", [{}, dict(D3, ss_kz_num=9, ss_kz_max=0.6),
                                  {"sample_monte_carlo": True, "mc_kt_num": 500, "mc_kd_max": 1.0},
                                  dict(D3, sample_monte_carlo=True, mc_kt_num=500, mc_kd_max=1.0, mc_kz_max=0.5),
                                  {"sample_stratified": True, "mc_kt_num": 500, "kd_sigma": 0.4},
                                  dict(D3, sample_stratified=True, mc_kt_num=500, kd_sigma=0.4, kz_sigma=0.3,
                                       traj_phase_method="SCTS")])
def test_wfat_device_batches_match_oracle(ctx, over):
    kw = small_wfat(n_t=50, sample_stratified=False, ss_kd_max=1.2, ss_kd_num=64, seed=99, sample_cutoff_limit=1e-20)
    kw.update(over)
    ld, td, sd, rd = descriptors(kw)
    rows = O.rows_of(sd.dimension, sd.phase_method)
    tt = np.empty(sd.sample_t_num)
    ctx.check(ctx.lib.etraj_batch_times(ctx.handle, sd, _abi.dptr(tt)))
    np.testing.assert_array_equal(tt, O.batch_times(sd))
    n = 0
    for b in (1, 10, 22, 25, 31, 50):
        got = _device_batch(ctx, ld, td, sd, rows, b)
        ref = O.gen_electron_batch(ld, td, sd, b)
        ref = np.zeros((0, rows)) if ref is None else ref
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
        n += len(ref)
    assert n > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-3)])
def test_wfat_run_job_matches_oracle(ctx, precision, tol):
    kw = small_wfat(n_t=40, n_k=256, precision=precision)
    ld, td, sd, rd = descriptors(kw)
    npix = rd.final_p_num[0] * rd.final_p_num[1]
    got = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(got), _abi.C.byref(st)))
    rd64 = descriptors(dict(kw, precision="fp64"))[3]
    ref, rst = O.run_job(ld, td, sd, rd64, nthreads=4)
    assert st.num_effective_traj == rst["num_effective_traj"] > 0
    assert st.ion_prob == pytest.approx(rst["ion_prob"], rel=1e-6 if precision == "fp64" else 1e-3)
    err = l1(got, ref)
    print(precision, "WFAT PMD L1 vs oracle", err)
    assert err <= tol
