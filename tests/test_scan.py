"""Parameter scans (etraj_run_scan): several jobs in shared launches.

Every job of a scan must equal the same job run alone (etraj_run_job): the
per-electron arithmetic is the same, only the order of the PMD's atomic sums
differs, so PMDs agree to 1e-12 relative L1 and every counter exactly. WFAT
orientation scans are also checked against the oracle per orientation.
"""
import math
import warnings

import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_scts, small_wfat
from etraj_hip import (_abi, BichromaticLaser, Cos2Laser, perform_traj_scan, perform_traj_simulation,
                       workloads)


def _run_scan(ctx, descs):
    n = len(descs)
    ld = (_abi.LaserDesc * n)(*[d[0] for d in descs])
    sd = (_abi.SamplerDesc * n)(*[d[2] for d in descs])
    td, rd = descs[0][1], descs[0][3]
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(n * npix)
    st = (_abi.Stats * n)()
    ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, ld, td, sd, rd, _abi.dptr(out), st))
    return out.reshape(n, npix), [st[j].as_dict() for j in range(n)]


def _run_job(ctx, d):
    ld, td, sd, rd = d
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    return out, st.as_dict()


COUNTERS = ("num_effective_traj", "n_failed", "n_bound", "n_restep", "n_steps", "n_steps_fp64")


def _assert_scan_equals_singles(ctx, kws):
    keep = [descriptors(kw) for kw in kws]     # the descriptors point into sampler tables kept here
    got, gst = _run_scan(ctx, keep)
    for j, d in enumerate(keep):
        ref, rst = _run_job(ctx, d)
        for k in COUNTERS:
            assert gst[j][k] == rst[k], (j, k)
        assert gst[j]["ion_prob"] == pytest.approx(rst["ion_prob"], rel=1e-12)
        assert rst["num_effective_traj"] > 0
        assert l1(got[j], ref) <= 1e-12, j
    return gst


def _intensity_jobs(precision, n_t=24, n_k=96, ints=(5e13, 1.2e14, 3e14)):
    return [small_scts(n_t=n_t, n_k=n_k, precision=precision,
                       laser=Cos2Laser(peak_int=I, wave_len=800.0, cyc_num=2, ellip=0.0), seed=1000 + i)
            for i, I in enumerate(ints)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_intensity_scan_equals_single_jobs(ctx, precision):
    _assert_scan_equals_singles(ctx, _intensity_jobs(precision))


@pytest.mark.gpu
def test_bichromatic_counter_rotating_scan(ctx):
    def bichro(I):
        return BichromaticLaser(Cos2Laser(peak_int=I, wave_len=800.0, cyc_num=8, ellip=1.0),
                                Cos2Laser(peak_int=I, wave_len=400.0, cyc_num=16, ellip=-1.0), 0.0)
    half = 8 * math.pi / (45.563352525 / 800.0)
    kws = [small_scts(n_t=20, n_k=64, laser=bichro(I), sample_t_intv=(-half, half), traj_t_final=half + 10.0,
                      rate_prefix={"Pre", "Jac"}, seed=7 + i, precision="mixed")
           for i, I in enumerate((5e13, 2e14))]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        _assert_scan_equals_singles(ctx, kws)


@pytest.mark.gpu
def test_scan_jobs_of_different_sizes(ctx):
    kws = _intensity_jobs("mixed", ints=(1e14, 2e14))
    kws[1] = dict(kws[1], mc_kt_num=300, sample_t_num=13)
    _assert_scan_equals_singles(ctx, kws)


@pytest.mark.gpu
def test_scan_slabs_cross_job_boundaries(ctx):
    # 2 jobs x 1100 batches x 32000 columns = 7.0e7 columns > 2^26: two slabs, the first ending inside job 1
    kws = _intensity_jobs("mixed", n_t=1100, n_k=32000, ints=(1e14, 1.5e14))
    gst = _assert_scan_equals_singles(ctx, kws)
    assert gst[0]["n_slabs"] == 2


@pytest.mark.gpu
def test_tsit5_scan_equals_single_jobs(ctx):
    kws = [dict(kw, traj_integrator="tsit5", precision="fp64") for kw in _intensity_jobs("fp64", n_t=16, n_k=64)]
    _assert_scan_equals_singles(ctx, kws)


@pytest.mark.gpu
def test_wfat_orientation_scan_matches_oracle(ctx):
    rots = [(0.0, 0.0, 0.0), (0.0, math.pi / 2, 0.0), (1.0, 2.0, -0.5), (-2.0, 0.4, 2.5)]
    kws = [small_wfat(n_t=24, n_k=128, target=workloads.wfat_molecule(r)) for r in rots]
    keep = [descriptors(kw) for kw in kws]
    got, gst = _run_scan(ctx, keep)
    yields = []
    for j, (ld, td, sd, rd) in enumerate(keep):
        ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
        assert gst[j]["num_effective_traj"] == rst["num_effective_traj"] > 0
        assert l1(got[j], ref) <= 1e-6
        yields.append(gst[j]["ion_prob"])
    assert len(set(np.round(yields, 12))) == len(rots)      # the orientation changes the yield


@pytest.mark.gpu
def test_perform_traj_scan_host_api(ctx):
    kws = _intensity_jobs("mixed", n_t=16, n_k=64, ints=(1e14, 2e14))
    res = perform_traj_scan(kws)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        one = perform_traj_simulation(**kws[1])
    assert res[1]["num_effective_traj"] == one["num_effective_traj"]
    assert l1(res[1]["momentum_spec"], one["momentum_spec"]) <= 1e-12
    assert res[0]["momentum_spec"].shape == one["momentum_spec"].shape


def test_perform_traj_scan_rejects_different_runs():
    kws = _intensity_jobs("mixed", ints=(1e14, 2e14))
    kws[1] = dict(kws[1], final_p_num=(100, 100))
    with pytest.raises(ValueError, match="job 1"):
        perform_traj_scan(kws)
    kws = _intensity_jobs("mixed", ints=(1e14, 2e14))
    kws[1] = dict(kws[1], precision="fp64")
    with pytest.raises(ValueError, match="job 1"):
        perform_traj_scan(kws)
    with pytest.raises(ValueError):
        perform_traj_scan([])


@pytest.mark.gpu
def test_wfat_orientation_scan_of_a_database_molecule(ctx):
    # every channel of H2's WFAT integrals, four orientations in one launch; each equals its job alone
    from cases import fixture_molecule
    rots = [(0.0, 0.0, 0.0), (0.0, math.pi / 2, math.pi / 2), (1.0, 2.0, -0.5), (-2.0, 0.4, 2.5)]
    kws = [dict(small_wfat(n_t=24, n_k=128), target=fixture_molecule("Hydrogen", r)) for r in rots]
    _assert_scan_equals_singles(ctx, kws)


@pytest.mark.gpu
def test_mo_adk_intensity_scan(ctx):
    # CO's asymptotic coefficients (MO-ADK), three intensities in one launch
    from cases import fixture_molecule
    t = fixture_molecule("CarbonMonoxide", (0.3, 1.1, -0.7))
    kws = [dict(kw, target=t, mol_orbit_ridx=0) for kw in _intensity_jobs("mixed")]
    _assert_scan_equals_singles(ctx, kws)
