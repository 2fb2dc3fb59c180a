"""The ADK / SPA / SPANE prefactor beyond l = 0 (ADKSampler.jl:188-265,
SPASampler.jl:204-278, SPANESampler.jl:208-266): molecules with the
reference's asymptotic coefficients C_lm (MO-ADK / MO-SFA) and atoms with l > 0.

CPU pins of the oracle:
  * the reference's molecular KATs for H2 (ElectronSamplers_test.jl:78-100,
    :160-174, :240-254: MO-ADK 562 / 332, MO-SPANE 608 / 400, MO-SPA 619 / 422,
    2D and 3D, CTMC / QTMC / SCTS rows), reproduced with the C_lm of
    Molecule_Hydrogen.jld2;
  * Y_lm of gen_sph_harm_funcs at a real k, times (-i kappa)^l / |k|^l, is the
    Condon-Shortley spherical harmonic (scipy's sph_harm_y);
  * the Wigner-D rotation is unitary: for an atom with l = 2 the sum over m of
    |prefactor|^2 at a complex k does not depend on the quantization axis;
  * a molecule whose only coefficient is C_00 gives the scalar l = 0 form's
    rates and phases.
The H2 KATs do not distinguish the conjugation of k_ts nor the sign of
(-i kappa)^l (a copy of the oracle with either changed passes them); those
follow the reference's code as written (LinearAlgebra.dot conjugates its first
argument), and the GPU tests hold the device to the oracle on molecules and
atoms whose prefactors are complex (CO, H2O, Ar with m = 0 and m = 1).
"""
import math
import warnings

import numpy as np
import pytest
from scipy.special import sph_harm_y

import oracle as O
from cases import descriptors, fixture_molecule, l1
from etraj_hip import _abi, Cos2Laser, Cos4Laser, get_atom
from etraj_hip.samplers import init_sampler
from etraj_hip.targets import GenericMolecule, HydrogenLikeAtom

KAT_BASE = dict(dimension=2, sample_t_intv=(-100, 100), sample_t_num=400, sample_cutoff_limit=1e-14,
                sample_monte_carlo=False, traj_phase_method="CTMC", rate_prefix="Full", ss_kd_max=2.0,
                ss_kd_num=800, ss_kz_max=2.0, ss_kz_num=0, mc_kt_num=800, mc_kd_max=2.0, mc_kz_max=2.0,
                mol_orbit_ridx=0)
_3D = {"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30}
KATS = [  # size(gen_electron_batch(sp, 200)) for get_mol("Hydrogen")
    ("ADK", {}, (6, 562)), ("ADK", {"traj_phase_method": "QTMC"}, (7, 562)),
    ("ADK", {"traj_phase_method": "SCTS"}, (7, 562)), ("ADK", _3D, (8, 332)),
    ("ADK", dict(_3D, traj_phase_method="QTMC"), (9, 332)), ("ADK", dict(_3D, traj_phase_method="SCTS"), (9, 332)),
    ("SPANE", {}, (6, 608)), ("SPANE", {"traj_phase_method": "QTMC"}, (7, 608)),
    ("SPANE", {"traj_phase_method": "SCTS"}, (7, 608)), ("SPANE", _3D, (8, 400)),
    ("SPA", {}, (6, 619)), ("SPA", {"traj_phase_method": "QTMC"}, (7, 619)),
    ("SPA", {"traj_phase_method": "SCTS"}, (7, 619)), ("SPA", _3D, (8, 422)),
]


def _sampler(method, target, laser=None, **over):
    l = laser or Cos4Laser(peak_int=1e15, wave_len=800.0, cyc_num=2, ellip=1.0)
    kw = dict(KAT_BASE)
    kw.update(over)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method=method, laser=l, target=target, **kw)
    return l, sp


@pytest.mark.parametrize("method,over,expect", KATS)
def test_oracle_reproduces_the_h2_molecular_kats(method, over, expect):
    t = fixture_molecule("Hydrogen")
    l, sp = _sampler(method, t, **over)
    got = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 200)
    assert (got.shape[1], got.shape[0]) == expect


@pytest.mark.parametrize("l", range(5))
def test_sph_harm_at_a_real_k_is_the_condon_shortley_harmonic(l):
    rng = np.random.default_rng(l)
    kappa = 0.9
    for _ in range(4):
        k = rng.normal(size=3)
        r = np.linalg.norm(k)
        theta, phi = math.acos(k[2] / r), math.atan2(k[1], k[0])
        for m in range(-l, l + 1):
            got = O.sph_harm_c(l, m, kappa, *k) * (-1j * kappa) ** l / r ** l
            assert got == pytest.approx(complex(sph_harm_y(l, m, theta, phi)), abs=1e-12)


def _atom_l2(m, theta, phi):
    return HydrogenLikeAtom(Ip=0.6, Z=1, l=2, m=m, asymp_coeff=1.3, quan_ax_θ=theta, quan_ax_ϕ=phi)


def test_the_rotation_of_the_prefactor_is_unitary():
    # sum_m |N_m|^2 = |C|^2 sum_m' |Y_lm'(k')|^2 for every quantization axis (D^l unitary)
    l = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.4)
    sums = []
    for theta, phi in [(0.0, 0.0), (0.7, 1.9), (2.2, -0.8)]:
        tot = None
        for m in range(-2, 3):
            t = _atom_l2(m, theta, phi)
            _, sp = _sampler("ADK", t, laser=l, sample_t_num=30, ss_kd_num=40, sample_cutoff_limit=0.0,
                             rate_prefix="Pre")
            got = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 14)
            tot = got[:, 5] if tot is None else tot + got[:, 5]
        sums.append(tot)
    np.testing.assert_allclose(sums[1], sums[0], rtol=1e-11)
    np.testing.assert_allclose(sums[2], sums[0], rtol=1e-11)
    assert not np.allclose(O.gen_electron_batch(l._desc(), _atom_l2(1, 0.7, 1.9)._desc(), sp.desc, 14)[:, 5],
                           sums[0] / 5)


def _c00_molecule(C):
    data = dict(name="c00", charge=0, spin=0, energy_levels=np.array([-0.5, 0.2]), orbit_occ=np.array([2.0, 0.0]),
                wfat_intdata={}, wfat_mu={}, asymp_coeff={0: np.array([[C]], dtype=complex)})
    return GenericMolecule(data, 0.3, 1.1, -0.4, soft_core=1e-10)


@pytest.mark.parametrize("method", ["ADK", "SPANE", "SPA"])
def test_a_c00_molecule_gives_the_scalar_l0_prefactor(method):
    mol = _c00_molecule(0.83)
    atom = HydrogenLikeAtom(Ip=0.5, Z=1, asymp_coeff=0.83)
    l = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.5)
    kw = dict(sample_t_num=30, ss_kd_num=60, traj_phase_method="QTMC", sample_cutoff_limit=1e-20)
    _, sm = _sampler(method, mol, laser=l, **kw)
    _, sa = _sampler(method, atom, laser=l, **kw)
    for b in (8, 15, 22):
        a = O.gen_electron_batch(l._desc(), mol._desc(), sm.desc, b)
        r = O.gen_electron_batch(l._desc(), atom._desc(), sa.desc, b)
        assert a.shape == r.shape
        np.testing.assert_allclose(a[:, :6], r[:, :6], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(np.angle(np.exp(1j * (a[:, 6] - r[:, 6]))), 0.0, atol=1e-10)


def test_sampler_descriptor_of_a_molecule():
    t = fixture_molecule("CarbonMonoxide", (0.1, 0.2, 0.3))
    _, sp = _sampler("ADK", t, mol_orbit_ridx=-1)
    d = sp.desc
    C = t.MolAsympCoeff(-1)
    assert d.asymp_lmax == C.shape[0] - 1 and d.orbit_ip == pytest.approx(t.ion_potential(-1))
    assert tuple(d.mol_rot) == (0.1, 0.2, 0.3)
    flat = np.ctypeslib.as_array(d.asymp_coeff, shape=(2 * C.size,))
    np.testing.assert_array_equal(flat[0::2] + 1j * flat[1::2], C.ravel(order="F"))


# ---- device ----------------------------------------------------------------------

def _device_batch(ctx, ld, td, sd, rows, b):
    cap = max(int(ctx.lib.etraj_batch_max_size(sd)), 1)
    out = np.empty((cap, rows))
    n = _abi.C.c_int64()
    ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(out), cap, _abi.C.byref(n)))
    return out[:n.value]


@pytest.mark.gpu
@pytest.mark.parametrize("method,over,expect", KATS)
def test_device_reproduces_the_h2_molecular_kats(ctx, method, over, expect):
    t = fixture_molecule("Hydrogen")
    l, sp = _sampler(method, t, **over)
    got = _device_batch(ctx, l._desc(), t._desc(), sp.desc, expect[0], 200)
    assert (got.shape[1], got.shape[0]) == expect


def _targets():
    ar = get_atom("Ar")
    ar1 = get_atom("Ar", m=1, quan_ax_θ=0.6, quan_ax_ϕ=-1.1)
    return [("CO", fixture_molecule("CarbonMonoxide", (0.3, 1.1, -0.7)), 0),
            ("CO-1", fixture_molecule("CarbonMonoxide", (2.0, 0.4, 1.0)), -1),
            ("H2O", fixture_molecule("Water", (0.5, 2.0, 0.2)), 0),
            ("O2", fixture_molecule("Oxygen", (0.1, 0.9, 0.4)), (1, 0)),
            ("Ar", ar, 0), ("Ar m=1", ar1, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["ADK", "SPANE", "SPA"])
@pytest.mark.parametrize("dim", [2, 3])
def test_device_prefactor_batches_match_oracle(ctx, method, dim):
    l = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.6)
    over = dict(_3D, ss_kd_num=12, ss_kz_num=7) if dim == 3 else {"ss_kd_num": 64}
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        targets = _targets()
    n = 0
    for name, t, ridx in targets:
        _, sp = _sampler(method, t, laser=l, sample_t_num=40, traj_phase_method="SCTS", sample_cutoff_limit=1e-24,
                         mol_orbit_ridx=ridx, **over)
        ld, td, sd = l._desc(), t._desc(), sp.desc
        rows = O.rows_of(sd.dimension, sd.phase_method)
        for b in (9, 17, 24, 31):
            got = _device_batch(ctx, ld, td, sd, rows, b)
            ref = O.gen_electron_batch(ld, td, sd, b)
            ref = np.zeros((0, rows)) if ref is None else ref
            assert got.shape == ref.shape, (name, b)
            np.testing.assert_allclose(got[:, :-1], ref[:, :-1], rtol=1e-10, atol=1e-300, err_msg=name)
            np.testing.assert_allclose(np.angle(np.exp(1j * (got[:, -1] - ref[:, -1]))), 0.0, atol=1e-9,
                                       err_msg=name)
            n += len(ref)
    assert n > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-3)])
def test_carbon_monoxide_mo_adk_job_matches_oracle(ctx, precision, tol):
    from cases import small_scts
    t = fixture_molecule("CarbonMonoxide", (0.2, 1.2, 0.5))
    kw = small_scts(n_t=40, n_k=256, precision=precision)
    kw.update(target=t, mol_orbit_ridx=0)
    ld, td, sd, rd = descriptors(kw)
    npix = rd.final_p_num[0] * rd.final_p_num[1]
    got = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(got), _abi.C.byref(st)))
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st.num_effective_traj == rst["num_effective_traj"] > 0
    assert l1(got, ref) <= tol
