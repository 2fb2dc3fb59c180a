"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must keep reproducing them (CPU) and the engine must match them (GPU)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from cases import descriptors, l1, small_ctmc, small_scts
from etraj_hip import _abi

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dense(f, n):
    a = np.zeros(int(n))
    a[f["idx"]] = f["val"]
    return a


def test_kat_fixture_matches_reference_table():
    k = json.load(open(os.path.join(G, "kat_counts.json")))
    assert [c["size"][1] for c in k["cases"]] == [590, 590, 590, 376, 376, 376, 536, 562, 610, 514, 540]


def test_oracle_reproduces_golden_ctmc():
    f = np.load(os.path.join(G, "ctmc_grid_small.npz"))
    ld, td, sd, rd = descriptors(small_ctmc(n_t=100, n_kd=32))
    pmd, st = O.run_job(ld, td, sd, rd, nthreads=1)
    assert l1(pmd, dense(f, f["npix"])) < 1e-12
    assert st["num_effective_traj"] == int(f["num_effective_traj"])


def test_oracle_reproduces_golden_scts():
    f = np.load(os.path.join(G, "scts_strat_small.npz"))
    ld, td, sd, rd = descriptors(small_scts(n_t=40, n_k=256))
    amp, st = O.run_job_amplitude(ld, td, sd, rd)
    assert l1(amp, dense(f, f["n"])) < 1e-12


def test_oracle_reproduces_golden_records():
    f = np.load(os.path.join(G, "records_small.npz"))
    ld, td, sd, rd = descriptors(small_ctmc(n_t=60, n_kd=40))
    rec, st = O.integrate_records(ld, td, rd, f["cols"])
    np.testing.assert_array_equal(st, f["status"])
    ok = st == 0
    np.testing.assert_allclose(rec[ok, 8:11], f["rec"][ok, 8:11], rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_engine_matches_golden_ctmc(ctx):
    f = np.load(os.path.join(G, "ctmc_grid_small.npz"))
    ld, td, sd, rd = descriptors(small_ctmc(n_t=100, n_kd=32))
    out = np.empty(int(f["npix"]))
    st = _abi.Stats()
    ctx.check(ctx.lib.etraj_run_job(ctx.handle, ld, td, sd, rd, _abi.dptr(out), _abi.C.byref(st)))
    assert st.num_effective_traj == int(f["num_effective_traj"])
    assert st.ion_prob == pytest.approx(float(f["ion_prob"]), rel=1e-9)
    assert l1(out, dense(f, f["npix"])) <= 1e-6


@pytest.mark.gpu
def test_engine_matches_golden_scts_amplitudes(ctx):
    """Amplitude-level (before abs2) comparison through the per-batch path."""
    f = np.load(os.path.join(G, "scts_strat_small.npz"))
    ld, td, sd, rd = descriptors(small_scts(n_t=40, n_k=256))
    amp = np.zeros(int(f["n"]))
    ip, iu, nf = _abi.C.c_double(0), _abi.C.c_double(0), _abi.C.c_int64(0)
    cap = int(ctx.lib.etraj_batch_max_size(sd))
    buf = np.empty((cap, 7))
    cnt = _abi.C.c_int64()
    for b in range(1, sd.sample_t_num + 1):
        ctx.check(ctx.lib.etraj_gen_electron_batch(ctx.handle, ld, td, sd, b, _abi.dptr(buf), cap, _abi.C.byref(cnt)))
        if cnt.value:
            ctx.check(ctx.lib.etraj_integrate_and_bin(ctx.handle, ld, td, rd, _abi.dptr(buf), cnt.value, 7,
                                                      amp.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                                      _abi.C.byref(iu), _abi.C.byref(nf)))
    assert ip.value == pytest.approx(float(f["ion_prob"]), rel=1e-9)
    assert l1(amp, dense(f, f["n"])) <= 1e-6


@pytest.mark.gpu
def test_engine_matches_golden_records(ctx):
    f = np.load(os.path.join(G, "records_small.npz"))
    ld, td, sd, rd = descriptors(small_ctmc(n_t=60, n_kd=40))
    cols = np.ascontiguousarray(f["cols"])
    n, rows = cols.shape
    rec = np.empty((n, 12))
    st = np.empty(n, np.int32)
    ctx.check(ctx.lib.etraj_integrate_records(ctx.handle, ld, td, rd, _abi.dptr(cols), n, rows, _abi.dptr(rec),
                                              st.ctypes.data_as(_abi.C.POINTER(_abi.C.c_int32))))
    assert (st != f["status"]).mean() <= 0.01
    both = (st == 0) & (f["status"] == 0)
    err = np.linalg.norm(rec[both, 8:11] - f["rec"][both, 8:11], axis=1) / np.linalg.norm(f["rec"][both, 8:11], axis=1)
    assert np.median(err) < 1e-12 and (err > 1e-8).mean() <= 0.01
