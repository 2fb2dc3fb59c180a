"""ETRAJ_PMD_FIXED (order-independent 128-bit fixed-point PMD sums) and the
per-cell hit counts (the integer-count PMD of a CTMC run), on the device.

Checks:
  * bit-identical PMD and ion_prob from repeated runs, and for a job run alone
    against the same job inside a scan (different slabs, different launch and
    atomic order);
  * fixed against FP64-atomic sums: relative L1 <= 1e-14 (both round each
    contribution once; the fixed sum is exact to 2^-100);
  * against the FP64 oracle: the usual FP64 bound (relative L1 <= 1e-6);
  * hit counts equal, cell by cell, the histogram of the same electrons'
    asymptotic momenta from the records entry point (same device arithmetic),
    binned with the reference's rounding (TrajectorySimulation.jl:375-384).
"""
import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_ctmc, small_scts
from etraj_hip import _abi
from etraj_hip.samplers import batch_num, init_sampler

pytestmark = pytest.mark.gpu


def _job(ctx, ld, td, sd, rd, hits=False):
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    h = np.zeros(npix, dtype=np.uint64) if hits else None
    rd.hit_counts = h.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64)) if hits else None
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    rd.hit_counts = None
    return out, st.as_dict(), h


@pytest.mark.parametrize("make", [lambda: small_ctmc(n_t=100, n_kd=32),
                                  lambda: small_scts(n_t=40, n_k=256),
                                  lambda: small_ctmc(n_t=60, n_kd=24, traj_integrator="tsit5")],
                         ids=["ctmc-rk4", "scts-rk4", "ctmc-tsit5"])
def test_fixed_pmd_is_reproducible_and_matches_float(ctx, make):
    kw = make()
    ld, td, sd, rd = descriptors(dict(kw, pmd_accumulator="fixed"))
    a, sa, _ = _job(ctx, ld, td, sd, rd)
    b, sb, _ = _job(ctx, ld, td, sd, rd)
    assert a.tobytes() == b.tobytes()
    assert sa["ion_prob"] == sb["ion_prob"] and sa["ion_prob_uncollected"] == sb["ion_prob_uncollected"]
    rdf = descriptors(dict(kw, pmd_accumulator="float"))[3]
    f, sf, _ = _job(ctx, ld, td, sd, rdf)
    assert l1(a, f) <= 1e-14
    assert sa["ion_prob"] == pytest.approx(sf["ion_prob"], rel=1e-14)
    ref, _ = O.run_job(ld, td, sd, descriptors(dict(kw, pmd_accumulator="float"))[3], nthreads=4)
    assert l1(a, ref) <= 1e-6


def test_fixed_pmd_of_a_scan_job_equals_the_job_alone(ctx):
    base = small_ctmc(n_t=80, n_kd=32, pmd_accumulator="fixed")
    other = dict(base, laser=type(base["laser"])(peak_int=2e14, wave_len=800.0, cyc_num=2.0, ellip=0.0))
    ld, td, sd, rd = descriptors(base)
    ld2, _, sd2, _ = descriptors(other)
    alone, st, _ = _job(ctx, ld, td, sd, rd)
    npix = alone.size
    out = np.empty(2 * npix)
    stats = (_abi.Stats * 2)()
    lasers = (_abi.LaserDesc * 2)(ld2, ld)
    samplers = (_abi.SamplerDesc * 2)(sd2, sd)
    ctx.check(ctx.lib.etraj_run_scan(ctx.handle, 2, lasers, td, samplers, rd, _abi.dptr(out), stats))
    assert out[npix:].tobytes() == alone.tobytes()
    assert stats[1].ion_prob == st["ion_prob"]


@pytest.mark.parametrize("phase", ["CTMC", "SCTS"])
def test_hit_counts_equal_the_histogram_of_the_records(ctx, phase):
    kw = small_ctmc(n_t=60, n_kd=24, traj_phase_method=phase)
    ld, td, sd, rd = descriptors(kw)
    _, st, hits = _job(ctx, ld, td, sd, rd, hits=True)
    sp = init_sampler(**kw)
    cols = []
    for b in range(1, batch_num(sp) + 1):
        cap = max(int(ctx.lib.etraj_batch_max_size(sp.desc)), 1)
        buf = np.empty((cap, sp.rows))
        n = _abi.C.c_int64()
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sp.desc, b, _abi.dptr(buf), cap,
                                                   _abi.C.byref(n)))
        cols.append(buf[:n.value])
    cols = np.concatenate(cols)
    assert len(cols) == st["num_effective_traj"]
    rec = np.empty((len(cols), 12))
    status = np.empty(len(cols), np.int32)
    ctx.check(ctx.lib.etraj_integrate_records(ctx.handle, ld, td, rd, _abi.dptr(cols), len(cols), sp.rows,
                                              _abi.dptr(rec), status.ctypes.data_as(_abi.C.POINTER(_abi.C.c_int32))))
    pmax = [rd.final_p_max[k] for k in range(2)]
    pnum = [rd.final_p_num[k] for k in range(2)]
    ok = status == 0
    idx = []
    for d in range(2):
        q = (rec[:, 8 + d] + pmax[d]) / (pmax[d] / pnum[d] * 2)
        i = np.where(np.abs(q) < 9.0e15, np.rint(q), 0).astype(np.int64)   # round(Int, ...) ties to even
        ok &= (i >= 1) & (i <= pnum[d])
        idx.append(i)
    lin = (idx[0] - 1) + pnum[0] * (idx[1] - 1)
    expect = np.bincount(lin[ok], minlength=pnum[0] * pnum[1]).astype(np.uint64)
    assert hits.sum() > 0
    np.testing.assert_array_equal(hits, expect)


def test_hit_counts_of_the_batch_path(ctx):
    """etraj_integrate_and_bin counts on the host, in column order."""
    kw = small_ctmc(n_t=40, n_kd=24)
    ld, td, sd, rd = descriptors(kw)
    _, st, hits = _job(ctx, ld, td, sd, rd, hits=True)
    sp = init_sampler(**kw)
    npix = hits.size
    h2 = np.zeros(npix, dtype=np.uint64)
    rd.hit_counts = h2.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64))
    pmd = np.zeros(npix)
    ip, iu, nf = _abi.C.c_double(0), _abi.C.c_double(0), _abi.C.c_int64(0)
    for b in range(1, batch_num(sp) + 1):
        cap = max(int(ctx.lib.etraj_batch_max_size(sp.desc)), 1)
        buf = np.empty((cap, sp.rows))
        n = _abi.C.c_int64()
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sp.desc, b, _abi.dptr(buf), cap,
                                                   _abi.C.byref(n)))
        if n.value:
            cols = np.ascontiguousarray(buf[:n.value])
            ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, ld, td, rd, _abi.dptr(cols), n.value, sp.rows,
                                                      pmd.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                                      _abi.C.byref(iu), _abi.C.byref(nf)))
    rd.hit_counts = None
    np.testing.assert_array_equal(h2, hits)


def test_perform_traj_simulation_returns_hit_counts(ctx):
    from etraj_hip import perform_traj_simulation
    kw = small_ctmc(n_t=40, n_kd=24)
    r = perform_traj_simulation(**kw, pmd_accumulator="fixed", hit_counts=True)
    h = r["hit_counts"]
    assert h.shape == r["momentum_spec"].shape == tuple(kw["final_p_num"])
    assert h.dtype == np.uint64 and h.sum() > 0
    assert ((h > 0) == (r["momentum_spec"] > 0)).all()     # CTMC: a cell has weight iff it has electrons
    r2 = perform_traj_simulation(**kw, pmd_accumulator="fixed", hit_counts=True)
    assert r2["momentum_spec"].tobytes() == r["momentum_spec"].tobytes()
    np.testing.assert_array_equal(r2["hit_counts"], h)
