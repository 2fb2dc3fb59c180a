"""GPU parity of the remaining workload families against the CPU FP64 oracle:

  * SAE (Tong-Lin) targets inside the integrators (SAEAtom.jl:98-186): He and Ar,
    2D SCTS with RK4 (FP64 and mixed, whose FP32 pass runs the FP32 SAE force) and
    with Tsit5, with the reference's 2D SCTS quirk (SAEAtom.jl:136) on and off;
  * the Ar SFA-SPA 3D grid with adaptive Tsit5 (rtol 1e-8) and QTMC phase, on a
    small grid (k_tsit5 in 3D, the SAE force, the SPA sampler, the 3D PMD);
  * bichromatic counter-rotating CP jobs (BichromaticLaser.jl:59-85) of the
    intensity scan, H ADK + SCTS, each job of one scan call against the oracle;
  * the H-atom SCTS job through the reference's own integrator, Tsit5 at
    traj_rtol 1e-6 (TrajectorySimulation.jl:342), against the oracle's Tsit5;
  * mixed Tsit5 with a soft-core (molecular, a = 1) target and SCTS.

Tolerances (relative L1 of the PMD):
  FP64 RK4                      1e-6  (same algorithm; libm ulps)
  FP64 Tsit5                    1e-6, or 3x the L1 between two oracle runs whose
                                initial conditions differ by one ulp where that is
                                larger (accept / reject decisions flip on ulps)
  mixed                         1e-3  (the engine's stated bound)
"""
import warnings

import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_ctmc, small_scts, small_wfat
from etraj_hip import _abi, Cos2Laser, get_atom, workloads

pytestmark = pytest.mark.gpu


def _job(ctx, ld, td, sd, rd):
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    return out, st.as_dict()


def _scan(ctx, jobs):
    ds = [descriptors(kw) for kw in jobs]
    n = len(ds)
    ld = (_abi.LaserDesc * n)(*[d[0] for d in ds])
    sd = (_abi.SamplerDesc * n)(*[d[2] for d in ds])
    st = (_abi.Stats * n)()
    rd = ds[0][3]
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(n * npix)
    ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, ld, ds[0][1], sd, rd, _abi.dptr(out), st))
    return out.reshape(n, npix), [st[j].as_dict() for j in range(n)], ds


def _columns(kw, batches):
    ld, td, sd, rd = descriptors(kw)
    cols = [O.gen_electron_batch(ld, td, sd, b) for b in batches]
    return ld, td, sd, rd, np.ascontiguousarray(np.concatenate([c for c in cols if c is not None]))


def _device_records(ctx, ld, td, rd, cols):
    n, rows = cols.shape
    rec = np.empty((n, 12))
    st = np.empty(n, np.int32)
    ctx.check(ctx.lib.etraj_integrate_records(ctx.handle, ld, td, rd, _abi.dptr(cols), n, rows, _abi.dptr(rec),
                                              st.ctypes.data_as(_abi.C.POINTER(_abi.C.c_int32))))
    return rec, st


def _device_bin(ctx, ld, td, rd, cols, npix, coherent):
    pmd = np.zeros(2 * npix if coherent else npix)
    ip, iu, nf = _abi.C.c_double(0.0), _abi.C.c_double(0.0), _abi.C.c_int64(0)
    ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, ld, td, rd, _abi.dptr(cols), cols.shape[0], cols.shape[1],
                                              pmd.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                              _abi.C.byref(iu), _abi.C.byref(nf)))
    return pmd, ip.value, nf.value


def _amp_abs2(pmd, coherent):
    return pmd[0::2] ** 2 + pmd[1::2] ** 2 if coherent else pmd


def _ulp_conditioning(ld, td, rd, cols, npix, coherent):
    """L1 between the oracle's PMD of `cols` and of the same columns with every position and
    momentum moved by one ulp: how far an FP64 integration of these electrons moves when its
    arithmetic differs in the last bit."""
    ref = np.zeros(2 * npix if coherent else npix)
    O.integrate_and_bin(ld, td, rd, cols, ref)
    pert = cols.copy()
    dim = rd.dimension
    pert[:, :2 * dim] = np.nextafter(pert[:, :2 * dim], np.inf)
    alt = np.zeros_like(ref)
    O.integrate_and_bin(ld, td, rd, np.ascontiguousarray(pert), alt)
    return ref, l1(_amp_abs2(alt, coherent), _amp_abs2(ref, coherent))


# ---- SAE targets in the integrators ----------------------------------------------

def _sae_atom(name):
    """He (l = 0), or Ar with its l = 1, m = 0 orbital quantized along the field (x): a p_z
    orbital gives no 2D (xy-plane) ionization through its prefactor."""
    return get_atom(name, quan_ax_θ=np.pi / 2) if name == "Ar" else get_atom(name)


def _sae_kw(name, integrator, precision, quirk):
    kw = small_scts(n_t=24, n_k=128, precision=precision, p_num=240, target=_sae_atom(name),
                    reproduce_quirks=quirk, traj_integrator=integrator, restep_radius=2.0)
    if integrator == "tsit5":
        kw["traj_rtol"] = 1e-8
    return kw


@pytest.mark.parametrize("quirk", [True, False], ids=["quirk", "fixed"])
@pytest.mark.parametrize("name", ["He", "Ar"])
def test_sae_scts_rk4_fp64_matches_oracle(ctx, name, quirk):
    kw = _sae_kw(name, "rk4", "fp64", quirk)
    ld, td, sd, rd = descriptors(kw)
    got, st = _job(ctx, ld, td, sd, rd)
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st["num_effective_traj"] == rst["num_effective_traj"] > 0
    assert st["n_failed"] == rst["n_failed"]
    assert st["ion_prob"] == pytest.approx(rst["ion_prob"], rel=1e-9)
    err = l1(got, ref)
    print(f"SAE {name} quirk={quirk} RK4 FP64: {st['num_effective_traj']} electrons, PMD L1 {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("quirk", [True, False], ids=["quirk", "fixed"])
@pytest.mark.parametrize("name", ["He", "Ar"])
def test_sae_scts_rk4_mixed_matches_oracle(ctx, name, quirk):
    """The FP32 pass runs the FP32 Tong-Lin force and potential (three exp per evaluation)."""
    kw = _sae_kw(name, "rk4", "mixed", quirk)
    ld, td, sd, rd = descriptors(kw)
    got, st = _job(ctx, ld, td, sd, rd)
    ref, rst = O.run_job(ld, td, sd, descriptors(dict(kw, precision="fp64"))[3], nthreads=4)
    assert st["num_effective_traj"] == rst["num_effective_traj"]
    assert 0 < st["n_restep"] < st["num_effective_traj"]
    err = l1(got, ref)
    print(f"SAE {name} quirk={quirk} RK4 mixed: PMD L1 {err:.2e}, re-integrated {st['n_restep']}")
    assert err <= 1e-3


def test_sae_quirk_changes_the_result(ctx):
    """The quirk switch reaches the device: the two phase potentials give different PMDs, each
    equal to the oracle's with the same switch (the tests above)."""
    a = _job(ctx, *descriptors(_sae_kw("Ar", "rk4", "fp64", True)))[0]
    b = _job(ctx, *descriptors(_sae_kw("Ar", "rk4", "fp64", False)))[0]
    assert l1(a, b) > 1e-3


@pytest.mark.parametrize("name", ["He", "Ar"])
def test_sae_scts_records_fp64_match_oracle(ctx, name):
    # He (Ip 0.904) at 4e14 W/cm^2, Ar at 1e14, for comparable yields on the grid
    laser = Cos2Laser(peak_int=4e14 if name == "He" else 1e14, wave_len=800.0, cyc_num=2, ellip=0.0)
    kw = small_ctmc(n_t=60, n_kd=40, traj_phase_method="SCTS", target=_sae_atom(name), laser=laser)
    ld, td, sd, rd, cols = _columns(kw, range(8, 60, 4))
    rec, st = _device_records(ctx, ld, td, rd, cols)
    rrec, rst = O.integrate_records(ld, td, rd, cols)
    both = (st == 0) & (rst == 0)
    assert (st != rst).mean() <= 0.01 and both.sum() > 100
    dp = np.linalg.norm(rec[both, 8:11] - rrec[both, 8:11], axis=1)
    dphi = np.abs(np.angle(np.exp(1j * (rec[both, 11] - rrec[both, 11]))))
    print(f"SAE {name} records: median dp {np.median(dp):.1e}, median dphi {np.median(dphi):.1e}")
    assert np.median(dp) < 1e-12
    assert np.median(dphi) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sae_scts_tsit5_matches_oracle(ctx, precision):
    kw = _sae_kw("Ar", "tsit5", precision, True)
    ld, td, sd, rd, cols = _columns(dict(kw, precision="fp64"), range(1, 24, 2))
    rd_dev = descriptors(kw)[3]
    npix = 240 * 240
    ref, cond = _ulp_conditioning(ld, td, rd, cols, npix, True)
    got, ip, nf = _device_bin(ctx, ld, td, rd_dev, cols, npix, True)
    err = l1(_amp_abs2(got, True), _amp_abs2(ref, True))
    print(f"SAE Ar Tsit5 {precision}: {len(cols)} electrons, PMD L1 {err:.2e} (one-ulp conditioning {cond:.2e})")
    assert err <= (max(1e-6, 3 * cond) if precision == "fp64" else 1e-3)


# ---- C3: Ar SFA-SPA 3D grid, Tsit5 rtol 1e-8, QTMC ---------------------------------

def _ar_spa_kw(precision):
    return workloads.ar_cp_spa_qtmc_3d(n_t=40, n_k=12, p_num=40, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ar_spa_3d_tsit5_qtmc_run_job_matches_oracle(ctx, precision):
    kw = _ar_spa_kw(precision)
    ld, td, sd, rd = descriptors(kw)
    got, st = _job(ctx, ld, td, sd, rd)
    ref, rst = O.run_job(ld, td, sd, descriptors(dict(kw, precision="fp64"))[3], nthreads=4)
    assert st["num_effective_traj"] == rst["num_effective_traj"] > 100
    err = l1(got, ref)
    print(f"Ar SPA 3D Tsit5 {precision}: {st['num_effective_traj']} electrons, steps {st['n_steps']} "
          f"(oracle {rst['n_steps']}), PMD L1 {err:.2e}")
    if precision == "fp64":
        assert st["n_failed"] == rst["n_failed"]
        assert st["ion_prob"] == pytest.approx(rst["ion_prob"], rel=1e-9)
        # the conditioning of these electrons at this tolerance, from the same columns
        ld2, td2, sd2, rd2, cols = _columns(kw, range(1, 41))
        _, cond = _ulp_conditioning(ld2, td2, rd2, cols, 40 ** 3, True)
        print(f"   one-ulp conditioning {cond:.2e}")
        assert err <= max(1e-6, 3 * cond)
    else:
        assert err <= 1e-3


def test_ar_spa_3d_records_match_oracle(ctx):
    """Per electron, 3D Tsit5 FP64 against the oracle on the SPA sampler's columns (sampled on the
    device): the bulk agrees to the integrator's tolerance."""
    kw = _ar_spa_kw("fp64")
    ld, td, sd, rd = descriptors(kw)
    cap = int(ctx.lib.etraj_batch_max_size(sd))
    buf = np.empty((cap, 9))
    cnt = _abi.C.c_int64()
    cols = []
    for b in range(1, 41, 3):
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(buf), cap, _abi.C.byref(cnt)))
        ref = O.gen_electron_batch(ld, td, sd, b)
        assert cnt.value == (0 if ref is None else len(ref))
        if cnt.value:
            # the rate is exp(-2 Re G) with a tunnelling exponent Re G of up to ~35 here, so the
            # quadratures' 1e-14 relative agreement on G becomes ~2 Re G x 1e-14 on the rate
            np.testing.assert_allclose(buf[:cnt.value], ref, rtol=5e-12, atol=1e-300)
            cols.append(buf[:cnt.value].copy())
    cols = np.ascontiguousarray(np.concatenate(cols))
    rec, st = _device_records(ctx, ld, td, rd, cols)
    rrec, rst = O.integrate_records(ld, td, rd, cols)
    both = (st == 0) & (rst == 0)
    assert (st != rst).mean() <= 0.01 and both.sum() > 50
    rel = np.linalg.norm(rec[both, 8:11] - rrec[both, 8:11], axis=1) / np.linalg.norm(rrec[both, 8:11], axis=1)
    print(f"Ar SPA 3D records: {both.sum()} electrons, median rel dp {np.median(rel):.1e}, p95 "
          f"{np.quantile(rel, 0.95):.1e}")
    assert np.median(rel) < 1e-10
    assert np.quantile(rel, 0.95) < 1e-5


# ---- C4: bichromatic counter-rotating CP scan jobs ----------------------------------

@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-3)])
def test_bichromatic_scan_jobs_match_oracle(ctx, precision, tol):
    ints = [5e13, 1.6e14, 5e14]
    jobs = workloads.bichromatic_scan_jobs(ints, n_t=48, n_k=48, precision=precision)
    got, sts, ds = _scan(ctx, jobs)
    for j, kw in enumerate(jobs):
        ld, td, sd, rd = descriptors(dict(kw, precision="fp64"))
        ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
        assert sts[j]["num_effective_traj"] == rst["num_effective_traj"] > 0
        err = l1(got[j], ref)
        print(f"bichromatic {ints[j]:.1e} W/cm2 {precision}: {rst['num_effective_traj']} electrons, PMD L1 {err:.2e}")
        assert err <= tol


# ---- the H-atom SCTS job through the reference's integrator ---------------------------

def test_scts_stratified_tsit5_rtol_1e6_matches_oracle(ctx):
    kw = small_scts(n_t=40, n_k=256, precision="fp64", traj_integrator="tsit5", traj_rtol=1e-6)
    ld, td, sd, rd = descriptors(kw)
    got, st = _job(ctx, ld, td, sd, rd)
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st["num_effective_traj"] == rst["num_effective_traj"]
    err = l1(got, ref)
    print(f"H SCTS Tsit5 rtol 1e-6: {st['num_effective_traj']} electrons, {st['n_steps']} steps (oracle "
          f"{rst['n_steps']}), PMD L1 {err:.2e}")
    assert abs(st["n_steps"] - rst["n_steps"]) <= 1e-3 * rst["n_steps"]
    assert err <= 1e-3


# ---- mixed Tsit5 with a soft-core target --------------------------------------------

def _mixed_vs_fp64_tsit5(ctx, kw, cols_kw):
    ld, td, sd, rd64, cols = _columns(dict(cols_kw, precision="fp64"), range(1, 24, 2))
    td = kw["target"]._desc()
    rd64 = descriptors(dict(kw, precision="fp64"))[3]
    rdm = descriptors(dict(kw, precision="mixed"))[3]
    rec, st = _device_records(ctx, ld, td, rdm, cols)
    r64, s64 = _device_records(ctx, ld, td, rd64, cols)
    both = (st == 0) & (s64 == 0)
    dphi = np.abs(np.angle(np.exp(1j * (rec[both, 11] - r64[both, 11]))))
    dp = np.linalg.norm(rec[both, 8:11] - r64[both, 8:11], axis=1)
    return dphi, dp, len(cols)


def test_tsit5_mixed_soft_core_scts_close_to_fp64(ctx):
    """A soft core of 1 (the molecular potential of GenericMolecule.jl:453-454) keeps the
    -Z a / r2^(3/2) term of the SCTS phase rate in the FP32 stage derivatives: mixed Tsit5 is as
    close to FP64 Tsit5 with the soft core of 1 as with a negligible one (same sampler, same
    columns). Dropping the term moves the phase by ~1e-2 rad per a.u. of time spent near r ~ 3."""
    base = small_wfat(n_t=24, n_k=96, traj_integrator="tsit5", traj_rtol=1e-8)
    soft, dp_soft, n = _mixed_vs_fp64_tsit5(ctx, base, base)
    mol = base["target"]
    hard_target = type(mol)(Ip=mol.Ip, charge=mol.charge, g00_coeff=mol.g00_coeff, rot_α=mol.rot[0],
                            rot_β=mol.rot[1], rot_γ=mol.rot[2], soft_core=1e-10, name="negligible soft core")
    hard, dp_hard, _ = _mixed_vs_fp64_tsit5(ctx, dict(base, target=hard_target), base)
    print(f"mixed vs FP64 Tsit5 over {n} electrons: soft core 1: median dphi {np.median(soft):.1e}, p95 "
          f"{np.quantile(soft, 0.95):.1e}, median dp {np.median(dp_soft):.1e}; soft core 1e-10: median dphi "
          f"{np.median(hard):.1e}, p95 {np.quantile(hard, 0.95):.1e}, median dp {np.median(dp_hard):.1e}")
    assert np.median(soft) <= 3 * np.median(hard) + 1e-6
    assert np.quantile(soft, 0.95) <= 3 * np.quantile(hard, 0.95) + 1e-5
    assert np.median(dp_soft) <= 3 * np.median(dp_hard) + 1e-9
