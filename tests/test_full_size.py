"""The workloads at their full sizes, checked through properties that do not
need the oracle to run at that size:

  * the 1e8-electron H-atom ADK+SCTS job (10000 t0 strata x 34400 Gaussian
    k_perp, coherent 1000^2 PMD): the mixed and FP64 runs sample the same
    >= 1e8 electrons; with fixed-point sums two mixed runs are the same bits;
    the mixed PMD is within relative L1 1e-3 of the FP64 PMD (the stated bound),
    and both runs give the same ion probability;
  * that job with 16000 strata (1.6e8 electrons) through Tsit5 (CTMC, fixed-point
    sums, hit counts), whose columns fill three slabs: every electron is integrated once (the CTMC
    identity sum(PMD) = ion_prob - ion_prob_uncollected, hit counts within the
    collected electrons), with no failures, and its ionization and PMD are those
    of the FP64 RK4 run of the same electrons up to the two integrators'
    truncation difference (1e8 electrons: ion_prob 1.2 % apart, PMD L1 7.8e-2);
  * the reference's CPU-runnable case (2000 t0 x 50 kd CTMC grid, 500^2 PMD) in
    FP64 against the oracle at full size (relative L1 <= 1e-6), with the
    CTMC identity sum(PMD) = ion_prob - ion_prob_uncollected and the hit counts
    summing to the collected electrons.
"""
import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1
from etraj_hip import _abi, workloads

pytestmark = pytest.mark.gpu


def _job(ctx, kw, hits=False):
    ld, td, sd, rd = descriptors(kw)
    npix = int(np.prod(kw["final_p_num"]))
    out = np.empty(npix)
    h = np.zeros(npix, dtype=np.uint64) if hits else None
    if hits:
        rd.hit_counts = h.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64))
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    return out, st.as_dict(), h, (ld, td, sd, rd)


def test_headline_job_at_full_size(ctx):
    kw = dict(workloads.h_lp_scts_stratified(), pmd_accumulator="fixed")
    mixed, sm, _, _ = _job(ctx, kw)
    again, sa, _, _ = _job(ctx, kw)
    fp64, s64, _, _ = _job(ctx, dict(kw, precision="fp64"))
    assert sm["num_effective_traj"] == sa["num_effective_traj"] == s64["num_effective_traj"] >= 1e8
    assert mixed.tobytes() == again.tobytes() and sm["ion_prob"] == sa["ion_prob"]
    err = l1(mixed, fp64)
    print("full-size PMD L1 mixed vs FP64", err, "re-integrated", sm["n_restep"])
    assert err <= 1e-3
    # a few of the 1e8 electrons end with an energy within rounding of zero, where FP32 and FP64
    # may disagree on ionized / bound (one electron's weight is ~1e-8 of the total)
    assert sm["ion_prob"] == pytest.approx(s64["ion_prob"], rel=1e-6)
    assert 0 < sm["n_restep"] < 0.1 * sm["num_effective_traj"]


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
    return out.reshape(n, npix), [st[j].as_dict() for j in range(n)]


def test_bichromatic_scan_at_full_size(ctx):
    """The 32-intensity counter-rotating bichromatic scan (~1e8 electrons, per-intensity counts from
    the ADK yield): every job of the scan equals the same job run alone (same counters, PMD L1
    <= 1e-12: only the order of the atomic sums differs), and each job's mixed PMD is within the
    stated bound of its FP64 PMD."""
    ints = list(np.logspace(np.log10(5e13), np.log10(5e14), 32))
    _, pst = _scan(ctx, workloads.bichromatic_scan_jobs(ints, n_t=200, n_k=200))
    n_k = workloads.bichromatic_scan_counts([s["ion_prob"] for s in pst], [s["num_effective_traj"] for s in pst],
                                            200 * 200)
    jobs = workloads.bichromatic_scan_jobs(ints, n_t=2000, n_k=n_k)
    mixed, sm = _scan(ctx, jobs)
    total = sum(s["num_effective_traj"] for s in sm)
    print(f"bichromatic scan: {total} electrons")
    assert 0.8e8 < total < 1.3e8
    for j, kw in enumerate(jobs):
        alone, sa, _, _ = _job(ctx, kw)
        for k in ("num_effective_traj", "n_failed", "n_bound", "n_restep", "n_steps", "n_steps_fp64"):
            assert sm[j][k] == sa[k], (j, k)
        assert l1(mixed[j], alone) <= 1e-12, j
    fp64, s64 = _scan(ctx, [dict(kw, precision="fp64") for kw in jobs])
    errs = [l1(mixed[j], fp64[j]) for j in range(len(jobs))]
    print("per-intensity PMD L1 mixed vs FP64: max", max(errs), "median", float(np.median(errs)))
    assert max(errs) <= 1e-3


def test_ar_spa_3d_at_full_size(ctx):
    """The Ar SFA-SPA 3D grid (2000 x 158 x 158, Tsit5 rtol 1e-8, QTMC, 256^3 PMD): with fixed-point
    sums two FP64 runs are the same bits; mixed and FP64 sample the same electrons with the same
    ion probability; and the mixed PMD is as close to an FP64 run at rtol 1e-10 as the FP64 run
    at rtol 1e-8 is (at 9.6M electrons over 16.7M cells the L1 of either is set by the
    integrator's tolerance moving electrons between neighbouring cells)."""
    kw = dict(workloads.ar_cp_spa_qtmc_3d(), pmd_accumulator="fixed")
    a, sa, _, _ = _job(ctx, kw)
    b, sb, _, _ = _job(ctx, kw)
    assert a.tobytes() == b.tobytes() and sa["num_effective_traj"] == sb["num_effective_traj"] > 9e6
    m, smx, _, _ = _job(ctx, dict(kw, precision="mixed"))
    ref, sr, _, _ = _job(ctx, dict(kw, traj_rtol=1e-10))
    assert smx["num_effective_traj"] == sa["num_effective_traj"]
    assert smx["ion_prob"] == pytest.approx(sa["ion_prob"], rel=1e-6)
    trunc, mix = l1(a, ref), l1(m, ref)
    print(f"Ar SPA 3D full: L1 vs FP64 rtol 1e-10: FP64 rtol 1e-8 {trunc:.2e}, mixed {mix:.2e}; "
          f"mixed vs FP64 {l1(m, a):.2e}, re-integrated {smx['n_restep']}")
    assert mix <= 1.5 * trunc


def test_tsit5_job_in_three_slabs_at_full_size(ctx):
    kw = dict(workloads.h_lp_scts_stratified(n_t=16000, precision="fp64", p_num=500), traj_phase_method="CTMC",
              pmd_accumulator="fixed")
    ts5, st, hits, _ = _job(ctx, dict(kw, traj_integrator="tsit5", traj_rtol=1e-6), hits=True)
    rk4, sr, _, _ = _job(ctx, kw)
    print("Tsit5 1.6e8 CTMC:", st["n_slabs"], "slabs,", st["num_effective_traj"], "electrons, ion_prob",
          st["ion_prob"], "(RK4", sr["ion_prob"], "), PMD L1 vs RK4", l1(ts5, rk4), flush=True)
    assert st["n_slabs"] == 3
    assert st["num_effective_traj"] == sr["num_effective_traj"] >= 1.5e8
    assert st["n_failed"] == 0 and st["n_attempts"] >= st["n_steps"] > st["num_effective_traj"]
    assert ts5.sum() == pytest.approx(st["ion_prob"] - st["ion_prob_uncollected"], rel=1e-12)
    assert 0 < hits.sum() <= st["num_effective_traj"] - st["n_failed"] - st["n_bound"]
    # RK4 at dt 0.1 against Tsit5 at rtol 1e-6 on the 1e8 electrons of 10000 strata: ion_prob 1.2 %
    # apart, PMD L1 7.8e-2 (measured); a slab lost or counted twice would move ion_prob by tens of
    # per cent
    assert st["ion_prob"] == pytest.approx(sr["ion_prob"], rel=3e-2)
    assert l1(ts5, rk4) <= 0.15


def test_ctmc_grid_at_full_size_matches_oracle(ctx):
    kw = workloads.h_lp_ctmc_grid(n_t=2000, n_kd=50, precision="fp64")
    got, st, hits, (ld, td, sd, rd) = _job(ctx, dict(kw, pmd_accumulator="fixed"), hits=True)
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=16)
    assert st["num_effective_traj"] == rst["num_effective_traj"]
    assert st["ion_prob"] == pytest.approx(rst["ion_prob"], rel=1e-9)
    assert l1(got, ref) <= 1e-6
    assert got.sum() == pytest.approx(st["ion_prob"] - st["ion_prob_uncollected"], rel=1e-12)
    collected = st["num_effective_traj"] - st["n_failed"] - st["n_bound"]
    assert 0 < hits.sum() <= collected
