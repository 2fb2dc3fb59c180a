"""etraj_integrate_and_bin_noalloc (device buffers in the library's static
scratch array) against etraj_integrate_and_bin (device buffers from hipMalloc):
the same kernels on the same columns, so the host-binned PMDs and counters
are identical. Two calls in a row, the second with fewer columns, for each
integrator form; and the column limit is an error."""
import numpy as np
import pytest

import oracle as O
from cases import descriptors, small_ctmc, small_scts
from etraj_hip import _abi

pytestmark = pytest.mark.gpu


def _columns(kw, batches):
    ld, td, sd, rd = descriptors(kw)
    cols = [O.gen_electron_batch(ld, td, sd, b) for b in batches]
    return ld, td, rd, np.ascontiguousarray(np.concatenate([c for c in cols if c is not None]))


def _bin(ctx, fn, ld, td, rd, cols, npix, coherent):
    pmd = np.zeros(2 * npix if coherent else npix)
    ip, iu, nf = _abi.C.c_double(0.0), _abi.C.c_double(0.0), _abi.C.c_int64(0)
    ctx.check(fn(ctx.handle, ld, td, rd, _abi.dptr(cols), cols.shape[0], cols.shape[1],
                 pmd.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip), _abi.C.byref(iu), _abi.C.byref(nf)))
    return pmd, ip.value, iu.value, nf.value


CASES = {
    "rk4-mixed-scts": lambda: small_scts(n_t=40, n_k=256, precision="mixed", p_num=300, restep_radius=2.0),
    "rk4-fp64-ctmc": lambda: small_ctmc(n_t=60, n_kd=40),
    "tsit5-mixed-scts": lambda: small_ctmc(n_t=60, n_kd=40, traj_phase_method="SCTS", traj_integrator="tsit5",
                                           precision="mixed", traj_rtol=1e-8),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_equals_integrate_and_bin(ctx, case):
    kw = CASES[case]()
    ld, td, rd, cols = _columns(kw, range(2, kw["sample_t_num"] - 2, 2))
    npix = int(np.prod(kw["final_p_num"]))
    coh = kw["traj_phase_method"] != "CTMC"
    print(f"no-alloc {case}: {cols.shape[0]} columns, then {cols.shape[0] // 3}", flush=True)
    for sub in (cols, np.ascontiguousarray(cols[: cols.shape[0] // 3])):   # second call: a smaller input
        ref = _bin(ctx, ctx.lib.etraj_integrate_and_bin, ld, td, rd, sub, npix, coh)
        got = _bin(ctx, ctx.lib.etraj_integrate_and_bin_noalloc, ld, td, rd, sub, npix, coh)
        assert got[1:] == ref[1:]
        assert got[0].tobytes() == ref[0].tobytes()
        assert ref[1] > 0


def test_noalloc_rejects_more_columns_than_its_scratch(ctx):
    kw = small_ctmc(n_t=60, n_kd=40)
    ld, td, sd, rd = descriptors(kw)
    cols = np.zeros((65537, 6))
    cols[:, 4] = -50.0
    pmd = np.zeros(500 * 500)
    ip, iu, nf = _abi.C.c_double(0.0), _abi.C.c_double(0.0), _abi.C.c_int64(0)
    rc = ctx.lib.etraj_integrate_and_bin_noalloc(ctx.handle, ld, td, rd, _abi.dptr(cols), cols.shape[0], 6,
                                                 pmd.ctypes.data_as(_abi.C.c_void_p), _abi.C.byref(ip),
                                                 _abi.C.byref(iu), _abi.C.byref(nf))
    assert rc == -5   # ETRAJ_ECAPACITY
    assert b"65536" in ctx.lib.etraj_last_error(ctx.handle)
    assert not pmd.any()
