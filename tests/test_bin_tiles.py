"""PMD binning through LDS-privatised tiles (ETRAJ_BIN_TILES) and the launch tuning, on the device.

The tile binning adds the same contributions as the integrator's direct atomics
(ETRAJ_BIN_ATOMIC), in another order. Checks, per workload (RK4 and Tsit5, FP64
and mixed -- whose FP64 re-integration pass keeps its atomics --, CTMC weights
and coherent amplitudes, 2D and 3D, one job and a scan):
  * ETRAJ_PMD_FIXED: the PMD, ion_prob and the hit counts are bit-identical to
    the atomic path's (integer sums);
  * ETRAJ_PMD_FLOAT: relative L1 <= 1e-13 against the atomic path (FP64
    reassociation only), equal counters;
  * one case against the FP64 oracle (relative L1 <= 1e-6, the FP64 bound);
  * small tiles and chunks (many tiles, many workgroups per tile) give the same
    fixed-point PMD;
  * every launch choice of etraj_set_tuning (workgroup size, Tsit5 refill and
    claim, FP64-pass overlap) leaves a fixed-point PMD bit-identical;
  * invalid tuning values are rejected; the autotuner returns a valid set.
"""
import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_ctmc, small_scts
from etraj_hip import _abi, workloads

pytestmark = pytest.mark.gpu

CASES = {
    "ctmc-rk4-fp64": lambda: small_ctmc(n_t=100, n_kd=32),
    "scts-rk4-mixed": lambda: small_scts(n_t=40, n_k=256, precision="mixed"),
    "ctmc-tsit5": lambda: small_ctmc(n_t=60, n_kd=24, traj_integrator="tsit5"),
    "scts-tsit5-mixed": lambda: small_scts(n_t=24, n_k=128, traj_integrator="tsit5", precision="mixed"),
    "qtmc-3d-tsit5": lambda: workloads.ar_cp_spa_qtmc_3d(n_t=40, n_k=12, p_num=40),
}


def _job(ctx, kw, binning, hits=False):
    ld, td, sd, rd = descriptors(dict(kw, pmd_binning=binning))
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    h = np.zeros(npix, dtype=np.uint64) if hits else None
    rd.hit_counts = h.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64)) if hits else None
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    rd.hit_counts = None
    return out, st.as_dict(), h


COUNTERS = ("num_effective_traj", "n_failed", "n_bound", "n_restep", "n_steps", "n_steps_fp64", "n_attempts")


@pytest.mark.parametrize("name", list(CASES))
def test_tiles_equal_atomics_fixed_point(ctx, name):
    kw = dict(CASES[name](), pmd_accumulator="fixed")
    a, sa, ha = _job(ctx, kw, "atomic", hits=True)
    t, stt, ht = _job(ctx, kw, "tiles", hits=True)
    print(f"{name}: {sa['num_effective_traj']} electrons, ion_prob {sa['ion_prob']:.6e}")
    assert a.tobytes() == t.tobytes()
    assert sa["ion_prob"] == stt["ion_prob"] and sa["ion_prob_uncollected"] == stt["ion_prob_uncollected"]
    assert np.array_equal(ha, ht)
    assert int(ht.sum()) > 0
    for k in COUNTERS:
        assert sa[k] == stt[k], k


@pytest.mark.parametrize("name", list(CASES))
def test_tiles_equal_atomics_float(ctx, name):
    kw = CASES[name]()
    a, sa, _ = _job(ctx, kw, "atomic")
    t, stt, _ = _job(ctx, kw, "tiles")
    assert l1(t, a) <= 1e-13
    assert stt["ion_prob"] == pytest.approx(sa["ion_prob"], rel=1e-13)
    for k in COUNTERS:
        assert sa[k] == stt[k], k


def test_tiles_against_the_oracle(ctx):
    kw = small_ctmc(n_t=100, n_kd=32)
    t, st, _ = _job(ctx, kw, "tiles")
    ld, td, sd, rd = descriptors(kw)
    ref, rst = O.run_job(ld, td, sd, rd, nthreads=4)
    assert st["num_effective_traj"] == rst["num_effective_traj"]
    assert l1(t, ref) <= 1e-6


def test_small_tiles_and_chunks(ctx):
    """tiles of 64 cells (many tiles) and chunks of 256 records (many workgroups per tile)"""
    kw = dict(small_scts(n_t=40, n_k=256), pmd_accumulator="fixed")
    a, _, ha = _job(ctx, kw, "atomic", hits=True)
    old = ctx.get_tuning()
    try:
        for shift, chunk in ((6, 256), (9, 1000), (0, 65536)):
            ctx.set_tuning(bin_tile_shift=shift, bin_chunk=chunk)
            t, _, ht = _job(ctx, kw, "tiles", hits=True)
            assert a.tobytes() == t.tobytes(), (shift, chunk)
            assert np.array_equal(ha, ht)
    finally:
        ctx.set_tuning(**old)


def test_scan_tiles_equal_atomics(ctx):
    base = dict(small_scts(n_t=30, n_k=200, precision="mixed"), pmd_accumulator="fixed")
    jobs = [dict(base, laser=type(base["laser"])(peak_int=I, wave_len=800.0, cyc_num=2.0, ellip=0.0), seed=s)
            for I, s in ((1e14, 1), (1.6e14, 2), (0.7e14, 3))]
    outs = {}
    for binning in ("atomic", "tiles"):
        ds = [descriptors(dict(kw, pmd_binning=binning)) for kw in jobs]
        n = len(ds)
        ld = (_abi.LaserDesc * n)(*[d[0] for d in ds])
        sd = (_abi.SamplerDesc * n)(*[d[2] for d in ds])
        st = (_abi.Stats * n)()
        rd = ds[0][3]
        npix = rd.final_p_num[0] * rd.final_p_num[1]
        out = np.empty(n * npix)
        ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, ld, ds[0][1], sd, rd, _abi.dptr(out), st))
        outs[binning] = (out, [st[j].ion_prob for j in range(n)])
    assert outs["atomic"][0].tobytes() == outs["tiles"][0].tobytes()
    assert outs["atomic"][1] == outs["tiles"][1]


@pytest.mark.parametrize("choice", [dict(rk4_block=64), dict(rk4_block=128), dict(overlap_fp64=0),
                                    dict(bin_tiles=1)],
                         ids=["block64", "block128", "serial-fp64", "auto-tiles"])
def test_rk4_launch_choices_leave_the_pmd_unchanged(ctx, choice):
    kw = dict(small_scts(n_t=40, n_k=256, precision="mixed"), pmd_accumulator="fixed")
    ref, sref, _ = _job(ctx, kw, "auto")
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(**choice)
        out, st, _ = _job(ctx, kw, "auto")
    finally:
        ctx.set_tuning(**old)
    assert ref.tobytes() == out.tobytes()
    assert sref["ion_prob"] == st["ion_prob"]
    for k in COUNTERS:
        assert sref[k] == st[k], k


@pytest.mark.parametrize("choice", [dict(tsit5_refill=1), dict(tsit5_refill=64), dict(tsit5_claim=64),
                                    dict(tsit5_claim=4096)], ids=["refill1", "refill64", "claim64", "claim4096"])
def test_tsit5_launch_choices_leave_the_pmd_unchanged(ctx, choice):
    kw = dict(small_scts(n_t=24, n_k=128, traj_integrator="tsit5"), pmd_accumulator="fixed")
    ref, sref, _ = _job(ctx, kw, "auto")
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(**choice)
        out, st, _ = _job(ctx, kw, "auto")
    finally:
        ctx.set_tuning(**old)
    assert ref.tobytes() == out.tobytes()
    for k in COUNTERS:
        assert sref[k] == st[k], k


def test_invalid_tuning_is_rejected(ctx):
    for bad in (dict(rk4_block=96), dict(rk4_block=512), dict(tsit5_refill=0), dict(tsit5_claim=32),
                dict(bin_tile_shift=3), dict(bin_chunk=100), dict(overlap_fp64=2)):
        with pytest.raises(AssertionError):
            ctx.set_tuning(**bad)


def test_autotune_returns_a_valid_set(ctx):
    from etraj_hip import autotune
    old = ctx.get_tuning()
    try:
        res = autotune(**small_scts(n_t=24, n_k=128, precision="mixed"))
        assert res["best_ms"] > 0 and res["start_ms"] > 0
        assert res["best_ms"] <= res["start_ms"] * 1.05
        assert ctx.get_tuning() == res["best"]
        ctx.set_tuning(**res["best"])   # a valid set
    finally:
        ctx.set_tuning(**old)
