"""The workloads at their full sizes, checked through properties that do not
need the oracle to run at that size:

  * the 1e8-electron H-atom ADK+SCTS job (10000 t0 strata x 32000 Gaussian
    k_perp, coherent 1000^2 PMD): the mixed and FP64 runs sample the same
    93 463 129 electrons; with fixed-point sums two mixed runs are the same bits;
    the mixed PMD is within relative L1 1e-3 of the FP64 PMD (the stated bound),
    and both runs give the same ion probability;
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
    assert sm["num_effective_traj"] == sa["num_effective_traj"] == s64["num_effective_traj"] == 93463129
    assert mixed.tobytes() == again.tobytes() and sm["ion_prob"] == sa["ion_prob"]
    err = l1(mixed, fp64)
    print("full-size PMD L1 mixed vs FP64", err, "re-integrated", sm["n_restep"])
    assert err <= 1e-3
    assert sm["ion_prob"] == pytest.approx(s64["ion_prob"], rel=1e-9)
    assert 0 < sm["n_restep"] < 0.1 * sm["num_effective_traj"]


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
