"""Graph replay of a job (etraj_tuning.graph = 1), on the device.

A job whose batches fit one slab is captured into a HIP graph at its first call
and replayed while its launch arguments repeat. Checks, per workload (RK4 and
Tsit5, FP64 and mixed -- whose FP64 re-integration runs inside the graph --,
CTMC weights with hit counts, coherent amplitudes, 2D and 3D, one job and a
scan):
  * ETRAJ_PMD_FIXED: every one of several calls in a row (the capture, then
    replays, with the process's own eager device work and synchronisations
    between them) gives the PMD, ion_prob, hit counts and counters of the
    eager launches bit for bit; the graph path is the one taken (its stats
    carry ms_sample == 0);
  * a change of the job (seed, intensity) between calls is captured anew, and
    a return to an earlier job is right again;
  * ETRAJ_PMD_FLOAT: relative L1 <= 1e-13 against the eager launches (FP64
    atomics in another order only);
  * a job of several slabs ignores the choice and runs eagerly.
"""
import numpy as np
import pytest
import torch

from cases import descriptors, l1, small_ctmc, small_scts
from etraj_hip import _abi, workloads

pytestmark = pytest.mark.gpu

CASES = {
    "ctmc-rk4-fp64-hits": lambda: small_ctmc(n_t=100, n_kd=32),
    "scts-rk4-mixed": lambda: small_scts(n_t=40, n_k=256, precision="mixed"),
    "scts-rk4-mixed-tiles": lambda: small_scts(n_t=40, n_k=256, precision="mixed", pmd_binning="tiles"),
    "scts-tsit5": lambda: small_scts(n_t=24, n_k=128, traj_integrator="tsit5"),
    "scts-tsit5-mixed": lambda: small_scts(n_t=24, n_k=128, traj_integrator="tsit5", precision="mixed"),
    "qtmc-3d-tsit5": lambda: workloads.ar_cp_spa_qtmc_3d(n_t=40, n_k=12, p_num=40),
}

COUNTERS = ("num_effective_traj", "n_failed", "n_bound", "n_restep", "n_steps", "n_steps_fp64", "n_attempts")


def _job(ctx, kw, hits=False):
    ld, td, sd, rd = descriptors(kw)
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    h = np.zeros(npix, dtype=np.uint64) if hits else None
    rd.hit_counts = h.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint64)) if hits else None
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    rd.hit_counts = None
    return out, st.as_dict(), h


def _eager_work():
    """Device work of the process's own between two calls: torch launches on its stream, then a sync."""
    x = torch.arange(1 << 20, device="cuda:0", dtype=torch.float64)
    assert float((x * 2.0).sum()) == float((1 << 20) * ((1 << 20) - 1))
    torch.cuda.synchronize()


def _with_graph(ctx, fn):
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(graph=1)
        return fn()
    finally:
        ctx.set_tuning(**old)


def _same(ref, got):
    (a, sa, ha), (b, sb, hb) = ref, got
    assert a.tobytes() == b.tobytes()
    assert sa["ion_prob"] == sb["ion_prob"] and sa["ion_prob_uncollected"] == sb["ion_prob_uncollected"]
    for k in COUNTERS:
        assert sa[k] == sb[k], k
    if ha is not None:
        assert np.array_equal(ha, hb)


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replays_equal_eager_fixed_point(ctx, name):
    kw = dict(CASES[name](), pmd_accumulator="fixed")
    hits = name.startswith("ctmc")
    ref = _job(ctx, kw, hits)
    assert ref[1]["ms_sample"] > 0 and ref[1]["n_slabs"] == 1
    print(f"{name}: {ref[1]['num_effective_traj']} electrons, ion_prob {ref[1]['ion_prob']:.6e}", flush=True)

    def calls():
        got = []
        for _ in range(4):
            got.append(_job(ctx, kw, hits))
            _eager_work()
        return got

    for r, got in enumerate(_with_graph(ctx, calls)):
        assert got[1]["ms_sample"] == 0.0 and got[1]["ms_integrate"] > 0, f"call {r}: graph path not taken"
        _same(ref, got)
    if hits:
        assert int(ref[2].sum()) > 0
    # eager again after the replays: unchanged
    _same(ref, _job(ctx, kw, hits))


def test_graph_follows_a_change_of_job(ctx):
    base = dict(small_scts(n_t=30, n_k=200, precision="mixed"), pmd_accumulator="fixed")
    a = dict(base, seed=11)
    b = dict(base, seed=12, laser=type(base["laser"])(peak_int=1.3e14, wave_len=800.0, cyc_num=2.0, ellip=0.0))
    ref_a, ref_b = _job(ctx, a), _job(ctx, b)
    assert ref_a[0].tobytes() != ref_b[0].tobytes()
    seq = _with_graph(ctx, lambda: [_job(ctx, kw) for kw in (a, a, b, b, a, b)])
    for kw_ref, got in zip((ref_a, ref_a, ref_b, ref_b, ref_a, ref_b), seq):
        _same(kw_ref, got)


def test_graph_scan_equals_eager(ctx):
    base = dict(small_scts(n_t=30, n_k=200, precision="mixed"), pmd_accumulator="fixed")
    jobs = [dict(base, laser=type(base["laser"])(peak_int=I, wave_len=800.0, cyc_num=2.0, ellip=0.0), seed=s)
            for I, s in ((1e14, 1), (1.6e14, 2), (0.7e14, 3))]
    ds = [descriptors(kw) for kw in jobs]
    n = len(ds)
    ld = (_abi.LaserDesc * n)(*[d[0] for d in ds])
    sd = (_abi.SamplerDesc * n)(*[d[2] for d in ds])
    rd = ds[0][3]
    npix = rd.final_p_num[0] * rd.final_p_num[1]

    def scan():
        st = (_abi.Stats * n)()
        out = np.empty(n * npix)
        ctx.check(ctx.lib.etraj_run_scan(ctx.handle, n, ld, ds[0][1], sd, rd, _abi.dptr(out), st))
        return out, [st[j].as_dict() for j in range(n)]

    ref = scan()
    got = _with_graph(ctx, lambda: [scan(), scan(), scan()])
    for out, sts in got:
        assert out.tobytes() == ref[0].tobytes()
        for j in range(n):
            assert sts[j]["ms_sample"] == 0.0
            assert sts[j]["ion_prob"] == ref[1][j]["ion_prob"]
            for k in COUNTERS:
                assert sts[j][k] == ref[1][j][k], k


def test_graph_float_pmd_matches_eager(ctx):
    kw = small_scts(n_t=40, n_k=256, precision="mixed")
    ref = _job(ctx, kw)
    got = _with_graph(ctx, lambda: [_job(ctx, kw) for _ in range(3)])
    for out, st, _ in got:
        assert st["ms_sample"] == 0.0
        assert l1(out, ref[0]) <= 1e-13
        for k in COUNTERS:
            assert st[k] == ref[1][k], k


def test_graph_choice_ignored_for_several_slabs(ctx):
    # 2 x (2^25 + 1) init columns: two slabs, run eagerly whatever the choice
    kw = dict(workloads.h_lp_scts_stratified(n_t=2, n_k=(1 << 25) + 1, p_num=64), pmd_accumulator="fixed")
    ref = _job(ctx, kw)
    assert ref[1]["n_slabs"] == 2
    got = _with_graph(ctx, lambda: _job(ctx, kw))
    assert got[1]["ms_sample"] > 0 and got[1]["n_slabs"] == 2
    _same(ref, got)
