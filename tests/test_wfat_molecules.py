"""WFAT with the reference's own molecular data (every channel (nxi, m) of
the WFAT integrals, GenericMolecule.jl:236-260, WFATSampler.jl:132-249).

CPU pins of the oracle:
  * the reference's WFAT survivor-count KATs for H2 (ElectronSamplers_test.jl:
    296-319: 566 in 2D, 566 with rot_beta = rot_gamma = pi/2, 332 in 3D, and
    non-empty Monte-Carlo batches), reproduced with the H2 data of
    Molecule_Hydrogen.jld2;
  * Wigner's d^l_{m m'}(beta) equals <l m| exp(-i beta J_y) |l m'> computed by
    a matrix exponential of the Condon-Shortley J_y;
  * for data with m' = 0 only (H2) G_00 is the Legendre series sum_l I_l P_l(cos beta);
  * the dipole factor: G with mu equals G with mu = 0 times exp(-kappa mu_z(beta));
  * the channel sum: a molecule whose only non-zero channel is (0, 0) gives the
    table form's rates.
The file reader (etraj_hip.moldata) is checked against values the HDF5 tools
print for the same files, where the reference tree is present.
GPU tests: the KATs on the device, device batches of CO, H2O and O2 against
the oracle, and a WFAT job of O2 against the oracle.
"""
import math
import os
import warnings

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O
from cases import descriptors, fixture_molecule, l1
from etraj_hip import _abi, Cos2Laser, Cos4Laser
from etraj_hip.samplers import init_sampler
from etraj_hip.targets import GenericMolecule, TabulatedMolecule

KAT_BASE = dict(dimension=2, sample_t_intv=(-100, 100), sample_t_num=400, sample_cutoff_limit=1e-14,
                sample_monte_carlo=False, traj_phase_method="CTMC", ss_kd_max=2.0, ss_kd_num=800, ss_kz_max=2.0,
                ss_kz_num=0, mc_kt_num=800, mc_kd_max=2.0, mc_kz_max=2.0, mol_orbit_ridx=0)
KATS = [  # ElectronSamplers_test.jl:296-319, size(gen_electron_batch(sp, 200))
    ((0.0, 0.0, 0.0), {}, (6, 566)),
    ((0.0, math.pi / 2, math.pi / 2), {}, (6, 566)),
    ((0.0, 0.0, 0.0), {"dimension": 3, "ss_kd_num": 30, "ss_kz_num": 30}, (8, 332)),
]
MC = [{"sample_monte_carlo": True}, {"sample_monte_carlo": True, "dimension": 3}]


def _kat_sampler(rot, over, key="Hydrogen"):
    l = Cos4Laser(peak_int=1e15, wave_len=800.0, cyc_num=2, ellip=1.0)
    t = fixture_molecule(key, rot)
    kw = dict(KAT_BASE)
    kw.update(over)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sp = init_sampler(init_cond_method="WFAT", laser=l, target=t, **kw)
    return l, t, sp


@pytest.mark.parametrize("rot,over,expect", KATS)
def test_oracle_reproduces_the_h2_wfat_kats(rot, over, expect):
    l, t, sp = _kat_sampler(rot, over)
    got = O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 200)
    assert (got.shape[1], got.shape[0]) == expect


@pytest.mark.parametrize("over", MC)
def test_oracle_h2_monte_carlo_batches_are_not_empty(over):
    l, t, sp = _kat_sampler((0.0, 0.0, 0.0), dict(over, seed=5))
    assert O.gen_electron_batch(l._desc(), t._desc(), sp.desc, 200) is not None


def _jy(l):
    """J_y in the |l m> basis, m = l..-l (Condon-Shortley J_+)."""
    ms = np.arange(l, -l - 1, -1)
    jp = np.zeros((2 * l + 1, 2 * l + 1))
    for i in range(1, 2 * l + 1):          # <m+1| J+ |m>
        m = ms[i]
        jp[i - 1, i] = math.sqrt(l * (l + 1) - m * (m + 1))
    return (jp - jp.T) / 2j, ms


@pytest.mark.parametrize("l", range(7))
def test_wigner_d_is_the_exponential_of_jy(l):
    jy, ms = _jy(l)
    for beta in (0.0, 0.3, 1.2, 2.0, math.pi):
        D = expm(-1j * beta * jy).real
        for i, m in enumerate(ms):
            for j, mp in enumerate(ms):
                assert O.wigner_d(l, int(m), int(mp), beta) == pytest.approx(D[i, j], abs=1e-12)


def _wfat_desc(t, ridx=0, **over):
    l = Cos2Laser(peak_int=1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    kw = dict(KAT_BASE, sample_t_num=40, ss_kd_num=64, mol_orbit_ridx=ridx)
    kw.update(over)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return init_sampler(init_cond_method="WFAT", laser=l, target=t, **kw).desc


def test_h2_g00_is_the_legendre_series_of_the_m0_integrals():
    t = fixture_molecule("Hydrogen")
    sd = _wfat_desc(t)
    mu, I = t.wfat_data(0)
    mM, L = (I.shape[1] - 1) // 2, I.shape[2] - 1
    kappa = math.sqrt(2 * t.ion_potential(0))
    for beta in (0.0, 0.4, 1.3, 2.9):
        for gamma in (0.0, 1.0, -2.5):
            x = math.cos(beta)
            ref = sum(I[0, mM, l, l] * np.polynomial.legendre.legval(x, [0] * l + [1]) for l in range(L + 1))
            ref *= math.exp(-kappa * (-math.sin(beta) * mu[0] + x * mu[2]))
            assert O.wfat_G(sd, beta, gamma, 0, 0, kappa) == pytest.approx(ref, rel=1e-12, abs=1e-14)


def test_the_dipole_factor_of_carbon_monoxide():
    t = fixture_molecule("CarbonMonoxide")
    sd = _wfat_desc(t)
    mu, _ = t.wfat_data(0)
    assert abs(mu[2]) > 1.0                                # CO's HOMO dipole, 1.56 a.u.
    sd0 = _wfat_desc(t)
    for k in range(3):
        sd0.wfat_mu[k] = 0.0
    kappa = 1.1
    for beta, gamma, nxi, m in [(0.3, 0.2, 0, 0), (1.7, -1.0, 1, 1), (2.6, 2.0, 0, -2)]:
        muz = -math.sin(beta) * mu[0] + math.cos(beta) * mu[2]
        assert O.wfat_G(sd, beta, gamma, nxi, m, kappa) == pytest.approx(
            O.wfat_G(sd0, beta, gamma, nxi, m, kappa) * math.exp(-kappa * muz), rel=1e-12, abs=1e-300)


def test_a_single_channel_molecule_gives_the_table_rates():
    # integrals with only I[0, m=0, l, m'=0] = c_l: G_00 = sum_l c_l P_l(cos beta), no other channel
    c = np.array([1.3, -0.4, 0.25, 0.1])
    L = c.size - 1
    I = np.zeros((3, 5, L + 1, 2 * L + 1), complex)
    for l in range(L + 1):
        I[0, 2, l, l] = c[l]
    data = dict(name="test", charge=0, spin=0, energy_levels=np.array([-0.6, 0.1]), orbit_occ=np.array([2.0, 0.0]),
                wfat_intdata={0: I}, wfat_mu={0: np.zeros(3)})
    g = GenericMolecule(data, 0.4, 1.0, -0.3)
    tab = TabulatedMolecule(Ip=0.6, charge=0, g00_coeff=c[:, None], rot_α=0.4, rot_β=1.0, rot_γ=-0.3)
    l = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.5)
    for b in (5, 20, 33):
        a = O.gen_electron_batch(l._desc(), g._desc(), _wfat_desc(g, sample_cutoff_limit=1e-30), b)
        r = O.gen_electron_batch(l._desc(), tab._desc(), _wfat_desc(tab, sample_cutoff_limit=1e-30), b)
        assert a.shape == r.shape
        np.testing.assert_allclose(a, r, rtol=1e-12)


def test_open_shell_orbital_indices():
    t = fixture_molecule("Oxygen")
    assert t.open_shell and t.Ip == pytest.approx(0.5542638748810138)
    assert t.ion_potential((1, 0)) == pytest.approx(0.5542638748810138)
    with pytest.raises(ValueError):
        t.ion_potential(0)
    sd = _wfat_desc(t, ridx=(1, -1))
    assert sd.wfat_form == _abi.WFAT_INTDATA and sd.orbit_ip == pytest.approx(t.ion_potential((1, -1)))


REF_DB = "/root/reference/src/Targets/Database"


@pytest.mark.skipif(not os.path.isdir(REF_DB), reason="the reference tree is not present")
def test_reader_matches_the_hdf5_tools_view_of_the_h2_file():
    from etraj_hip.moldata import read_molecule_jld2
    m = read_molecule_jld2(os.path.join(REF_DB, "Molecule_Hydrogen.jld2"))
    # values h5dump prints for /energy_levels, /orbit_occ, /atom_coords, /charge of this file
    assert m["energy_levels"][:3] == pytest.approx([-0.594444, 0.147584, 0.231618], abs=1e-6)
    assert m["orbit_occ"][0] == 2 and not m["orbit_occ"][1:].any()
    assert m["charge"] == 0 and m["spin"] == 0 and m["atoms"] == ["H", "H"]
    np.testing.assert_allclose(m["atom_coords"], [[0, 0, -0.37072], [0, 0, 0.37072]], atol=1e-5)
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "mol_Hydrogen.npz"))
    np.testing.assert_array_equal(m["wfat_intdata"][0], f["wfat_intdata_0"])


# ---- device ----------------------------------------------------------------------

def _device_batch(ctx, ld, td, sd, rows, b):
    cap = max(int(ctx.lib.etraj_batch_max_size(sd)), 1)
    out = np.empty((cap, rows))
    n = _abi.C.c_int64()
    ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(out), cap, _abi.C.byref(n)))
    return out[:n.value]


@pytest.mark.gpu
@pytest.mark.parametrize("rot,over,expect", KATS)
def test_device_reproduces_the_h2_wfat_kats(ctx, rot, over, expect):
    l, t, sp = _kat_sampler(rot, over)
    got = _device_batch(ctx, l._desc(), t._desc(), sp.desc, expect[0], 200)
    assert (got.shape[1], got.shape[0]) == expect


@pytest.mark.gpu
@pytest.mark.parametrize("key,ridx,dim", [("CarbonMonoxide", 0, 2), ("CarbonMonoxide", -1, 3), ("Water", 0, 2),
                                          ("Water", -2, 3), ("Oxygen", (1, 0), 2), ("Oxygen", (1, -1), 3)])
def test_device_molecule_batches_match_oracle(ctx, key, ridx, dim):
    t = fixture_molecule(key, (0.3, 1.1, -0.7))
    l = Cos2Laser(peak_int=2e14, wave_len=800.0, cyc_num=2, ellip=0.6)
    over = {"dimension": dim, "ss_kz_num": 9, "ss_kz_max": 0.6} if dim == 3 else {}
    sd = _wfat_desc(t, ridx=ridx, sample_cutoff_limit=1e-24, **over)
    ld, td = l._desc(), t._desc()
    rows = O.rows_of(sd.dimension, sd.phase_method)
    n = 0
    for b in (3, 12, 20, 27, 36):
        got = _device_batch(ctx, ld, td, sd, rows, b)
        ref = O.gen_electron_batch(ld, td, sd, b)
        ref = np.zeros((0, rows)) if ref is None else ref
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-300)
        n += len(ref)
    assert n > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-3)])
def test_oxygen_wfat_job_matches_oracle(ctx, precision, tol):
    from etraj_hip import workloads
    t = fixture_molecule("Oxygen", (0.2, 0.9, 0.4))
    kw = workloads.mol_lp_scts_wfat(n_t=40, n_k=256, precision=precision, p_num=160)
    kw.update(target=t, mol_orbit_ridx=(1, 0))
    ld, td, sd, rd = descriptors(kw)
    npix = rd.final_p_num[0] * rd.final_p_num[1]
    got = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(got), _abi.C.byref(st)))
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st.num_effective_traj == rst["num_effective_traj"] > 0
    assert l1(got, ref) <= tol
