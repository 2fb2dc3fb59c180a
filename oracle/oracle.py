"""ctypes access to the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg as the checker / CPU baseline, never by
the product. Descriptors are the product's C-ABI structs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "etraj.jl_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from etraj_hip import _abi  # noqa: E402  (descriptor structs only)
from etraj_hip._abi import LaserDesc, TargetDesc, SamplerDesc, RunDesc, Stats  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")

P = C.POINTER
DP = P(C.c_double)


def build(force: bool = False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_philox4x64_10.argtypes = [P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]
        L.oracle_philox4x64_10.restype = None
        L.oracle_grid_point.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int]
        L.oracle_grid_point.restype = C.c_double
        L.oracle_bin_index.argtypes = [C.c_double, C.c_double, C.c_int32]
        L.oracle_bin_index.restype = C.c_int64
        L.oracle_tsit5_tableau.argtypes = [DP, DP, DP, DP]
        L.oracle_tsit5_tableau.restype = None
        L.oracle_laser_field.argtypes = [P(LaserDesc), DP, C.c_int64, DP, DP, DP, DP]
        L.oracle_unit_envelope.argtypes = [P(LaserDesc), C.c_double]
        L.oracle_unit_envelope.restype = C.c_double
        L.oracle_target_eval.argtypes = [P(TargetDesc), DP, C.c_int64, DP, DP]
        L.oracle_eom_rhs.argtypes = [P(LaserDesc), P(TargetDesc), C.c_int32, C.c_int32, C.c_int32, C.c_double, DP, DP]
        L.oracle_ff_mf_euler.argtypes = [DP, C.c_double, C.c_double, DP]
        L.oracle_ff_mf_euler.restype = None
        L.oracle_wfat_G.argtypes = [P(SamplerDesc), C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, DP]
        L.oracle_wfat_G.restype = None
        L.oracle_wigner_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        L.oracle_wigner_d.restype = C.c_double
        L.oracle_sph_harm_c.argtypes = [C.c_int, C.c_int, C.c_double, DP, DP]
        L.oracle_sph_harm_c.restype = None
        L.oracle_laser_complex.argtypes = [P(LaserDesc), C.c_double, C.c_double, DP]
        L.oracle_laser_field_deriv.argtypes = [P(LaserDesc), C.c_double, DP]
        L.oracle_spa_point.argtypes = [P(LaserDesc), P(TargetDesc), C.c_double, C.c_double, C.c_double, DP]
        L.oracle_batch_times.argtypes = [P(SamplerDesc), DP]
        L.oracle_batch_max_size.argtypes = [P(SamplerDesc)]
        L.oracle_batch_max_size.restype = C.c_int64
        L.oracle_gen_electron_batch.argtypes = [P(LaserDesc), P(TargetDesc), P(SamplerDesc), C.c_int32, DP,
                                                C.c_int64, P(C.c_int64)]
        L.oracle_integrate_state.argtypes = [P(LaserDesc), P(TargetDesc), P(RunDesc), C.c_double, DP, P(C.c_int64)]
        L.oracle_post_process.argtypes = [P(TargetDesc), C.c_int32, C.c_int32, C.c_int32, DP, C.c_double, DP,
                                          P(C.c_int32)]
        L.oracle_integrate_records.argtypes = [P(LaserDesc), P(TargetDesc), P(RunDesc), DP, C.c_int64, C.c_int32,
                                               DP, P(C.c_int32)]
        L.oracle_integrate_and_bin.argtypes = [P(LaserDesc), P(TargetDesc), P(RunDesc), DP, C.c_int64, C.c_int32,
                                               DP, DP, DP, P(C.c_int64)]
        L.oracle_run_job.argtypes = [P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc), C.c_int, C.c_int32,
                                     C.c_int32, C.c_int32, DP, P(Stats)]
        L.oracle_run_job_dynamic.argtypes = L.oracle_run_job.argtypes
        L.oracle_run_job_amplitude.argtypes = [P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc), C.c_int, DP,
                                               P(Stats)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(DP)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


def philox(ctr, key):
    c = (C.c_uint64 * 4)(*ctr)
    k = (C.c_uint64 * 2)(*key)
    o = (C.c_uint64 * 4)()
    lib().oracle_philox4x64_10(c, k, o)
    return [int(x) for x in o]


def laser_field(ld, t):
    t = np.ascontiguousarray(np.atleast_1d(np.asarray(t, float)))
    out = [np.empty_like(t) for _ in range(4)]
    _chk(lib().oracle_laser_field(ld, _d(t), t.size, *[_d(o) for o in out]))
    return out


def target_eval(td, xyz):
    xyz = np.ascontiguousarray(np.asarray(xyz, float).reshape(-1, 3))
    V = np.empty(len(xyz))
    F = np.empty((len(xyz), 3))
    _chk(lib().oracle_target_eval(td, _d(xyz), len(xyz), _d(V), _d(F)))
    return V, F


def eom_rhs(ld, td, dim, phase, quirks, t, u):
    u = np.ascontiguousarray(u, float)
    du = np.zeros(7)
    _chk(lib().oracle_eom_rhs(ld, td, dim, phase, quirks, t, _d(u), _d(du)))
    return du[:len(u)]


def ff_mf_euler(mol_rot, Fx, Fy):
    r = np.ascontiguousarray(mol_rot, dtype=np.float64)
    out = np.empty(3)
    lib().oracle_ff_mf_euler(_d(r), Fx, Fy, _d(out))
    return out


def wfat_G(sd, beta, gamma, nxi=0, m=0, kappa=1.0):
    """G_{nxi m}(beta, gamma) of the sampler descriptor's structure factor;
    kappa = sqrt(2 Ip) enters the dipole factor of the WFAT-integral form."""
    out = np.empty(2)
    lib().oracle_wfat_G(sd, kappa, beta, gamma, nxi, m, _d(out))
    return complex(out[0], out[1])


def wfat_g00(sd, beta, gamma, kappa=1.0):
    return wfat_G(sd, beta, gamma, 0, 0, kappa)


def wigner_d(l, m, mp, beta):
    return lib().oracle_wigner_d(l, m, mp, beta)


def sph_harm_c(l, m, kappa, x, y, z):
    """Y_lm of gen_sph_harm_funcs (shared_methods.jl:15-24) at complex (x, y, z)."""
    xyz = np.array([complex(x).real, complex(x).imag, complex(y).real, complex(y).imag, complex(z).real,
                    complex(z).imag])
    out = np.empty(2)
    lib().oracle_sph_harm_c(l, m, kappa, _d(xyz), _d(out))
    return complex(out[0], out[1])


def laser_complex(ld, t):
    """(Fx, Fy, Ax, Ay) at complex time t."""
    out = np.empty(8)
    _chk(lib().oracle_laser_complex(ld, t.real, t.imag, _d(out)))
    return tuple(complex(out[2 * i], out[2 * i + 1]) for i in range(4))


def laser_field_deriv(ld, t):
    """(Fx, Fy, dFx/dt, dFy/dt) at real time t."""
    out = np.empty(4)
    _chk(lib().oracle_laser_field_deriv(ld, t, _d(out)))
    return tuple(out)


def spa_point(ld, td, tr, kd, kz=0.0):
    """SFA-SPA saddle point of one sample: dict or None if the solve fails."""
    out = np.empty(12)
    rc = lib().oracle_spa_point(ld, td, tr, kd, kz, _d(out))
    if rc == 1:
        return None
    _chk(rc)
    k = ("ti", "px", "py", "x0", "y0", "kx0", "ky0")
    d = dict(zip(k, out[:7]))
    d["G"] = complex(out[7], out[8])
    d["kdotF"] = complex(out[9], out[10])
    d["jac"] = out[11]
    return d


def batch_times(sd):
    t = np.empty(sd.sample_t_num)
    _chk(lib().oracle_batch_times(sd, _d(t)))
    return t


def rows_of(dim, phase):
    return (8 if phase == _abi.PHASE_CTMC else 9) if dim == 3 else (6 if phase == _abi.PHASE_CTMC else 7)


def gen_electron_batch(ld, td, sd, batch_id, library=None):
    """returns (N, rows) C-order == Julia rows x N, or None. library: another build of the
    oracle (load_variant), default liboracle.so"""
    L = library or lib()
    cap = max(int(L.oracle_batch_max_size(sd)), 1)
    rows = rows_of(sd.dimension, sd.phase_method)
    out = np.empty((cap, rows))
    n = C.c_int64()
    _chk(L.oracle_gen_electron_batch(ld, td, sd, batch_id, _d(out), cap, C.byref(n)))
    return out[:n.value].copy() if n.value else None


def integrate_state(ld, td, rd, t0, u):
    u = np.ascontiguousarray(u, float).copy()
    ns = C.c_int64()
    rc = lib().oracle_integrate_state(ld, td, rd, t0, _d(u), C.byref(ns))
    return u, ns.value, rc


def post_process(td, dim, phase, quirks, u, t0):
    u = np.ascontiguousarray(u, float)
    rec = np.empty(12)
    st = C.c_int32()
    _chk(lib().oracle_post_process(td, dim, phase, quirks, _d(u), t0, _d(rec), C.byref(st)))
    return rec, st.value


def integrate_records(ld, td, rd, cols):
    cols = np.ascontiguousarray(cols, float)
    n, rows = cols.shape
    rec = np.empty((n, 12))
    st = np.empty(n, np.int32)
    _chk(lib().oracle_integrate_records(ld, td, rd, _d(cols), n, rows, _d(rec), st.ctypes.data_as(P(C.c_int32))))
    return rec, st


def integrate_and_bin(ld, td, rd, cols, pmd):
    cols = np.ascontiguousarray(cols, float)
    ip, iu, nf = C.c_double(0), C.c_double(0), C.c_int64(0)
    _chk(lib().oracle_integrate_and_bin(ld, td, rd, _d(cols), cols.shape[0], cols.shape[1], _d(pmd), C.byref(ip),
                                        C.byref(iu), C.byref(nf)))
    return ip.value, iu.value, nf.value


def run_job(ld, td, sd, rd, nthreads=1, batch_begin=0, batch_end=0, batch_stride=1, dynamic=False, library=None):
    """The oracle's whole job over the batches batch_begin, +batch_stride, ... < batch_end
    (0: all). dynamic: batches taken one at a time by the threads (timing runs). library: an
    oracle build loaded with load_variant (default: this directory's liboracle.so)."""
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    out = np.empty(npix)
    st = Stats()
    L = library or lib()
    fn = L.oracle_run_job_dynamic if dynamic else L.oracle_run_job
    _chk(fn(ld, td, sd, rd, nthreads, batch_begin, batch_end, batch_stride, _d(out), C.byref(st)))
    return out, st.as_dict()


def load_variant(path):
    """Another build of etraj_oracle.c (e.g. the -O3 -march=native timing build), with the
    run_job and sampler entry points bound."""
    L = C.CDLL(path)
    for name in ("oracle_run_job", "oracle_run_job_dynamic"):
        getattr(L, name).argtypes = lib().oracle_run_job.argtypes
    for name in ("oracle_batch_max_size", "oracle_gen_electron_batch"):
        getattr(L, name).argtypes = getattr(lib(), name).argtypes
        getattr(L, name).restype = getattr(lib(), name).restype
    return L


def build_native(outdir):
    """Compiles etraj_oracle.c with -O3 -march=native (FMA contraction allowed) for this host's
    CPU into outdir/liboracle_native.so; returns the path, or None if the build fails."""
    out = os.path.join(outdir, "liboracle_native.so")
    r = subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out}"], capture_output=True, text=True)
    return out if r.returncode == 0 and os.path.exists(out) else None


def run_job_amplitude(ld, td, sd, rd):
    npix = int(np.prod([rd.final_p_num[k] for k in range(rd.dimension)]))
    n = npix * (2 if rd.phase_method != _abi.PHASE_CTMC else 1)
    out = np.empty(n)
    st = Stats()
    _chk(lib().oracle_run_job_amplitude(ld, td, sd, rd, 1, _d(out), C.byref(st)))
    return out, st.as_dict()


def tsit5_tableau():
    c, a, b, bt = np.empty(7), np.empty(49), np.empty(7), np.empty(7)
    lib().oracle_tsit5_tableau(_d(c), _d(a), _d(b), _d(bt))
    return c, a.reshape(7, 7), b, bt
