"""ctypes mirror of include/etraj.h and the loader of libetraj_hip.so.

The library is loaded from this package directory (or from $ETRAJ_HIP_LIB). There is no fallback:
if the shared object is missing or cannot be loaded, every entry point that
needs the device raises :class:`EngineUnavailable`.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ETRAJ_HIP_LIB names another build of the same library (the Julia binding reads the same variable)
LIB_PATH = os.environ.get("ETRAJ_HIP_LIB", os.path.join(HERE, "libetraj_hip.so"))

ABI_VERSION = 4

OK, EINVAL, EHIP, ENOMEM, EUNSUPPORTED, ECAPACITY, ERANGE = 0, -1, -2, -3, -4, -5, -6

LASER_COS2, LASER_COS4, LASER_GAUSSIAN = 1, 2, 3
TARGET_HYDROGEN_LIKE, TARGET_SAE = 1, 2
INIT_ADK, INIT_SPA, INIT_SPANE, INIT_WFAT = 1, 2, 3, 4
WFAT_TABLE, WFAT_INTDATA = 0, 1
PHASE_CTMC, PHASE_QTMC, PHASE_SCTS = 0, 1, 2
PREFIX_PRE, PREFIX_PRECC, PREFIX_JAC = 1, 2, 4
SAMPLING_STEP, SAMPLING_MONTECARLO, SAMPLING_STRATIFIED = 0, 1, 2
INTEG_RK4, INTEG_TSIT5 = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
QUIRK_SCTS3D_XPY, QUIRK_SAE2D_SCTS = 1, 2
QUIRKS_REFERENCE = QUIRK_SCTS3D_XPY | QUIRK_SAE2D_SCTS
PMD_FLOAT, PMD_FIXED = 0, 1
BIN_AUTO, BIN_ATOMIC, BIN_TILES = 0, 1, 2

PHASE_METHODS = {"CTMC": PHASE_CTMC, "QTMC": PHASE_QTMC, "SCTS": PHASE_SCTS}


class MonoLaser(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("peak_int", C.c_double),
                ("wave_len", C.c_double), ("cyc_num", C.c_double), ("ellip", C.c_double),
                ("azi", C.c_double), ("cep", C.c_double), ("t_shift", C.c_double)]


class LaserDesc(C.Structure):
    _fields_ = [("n_colors", C.c_int32), ("reserved", C.c_int32), ("delay21", C.c_double),
                ("color", MonoLaser * 2)]


class TargetDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nucl_charge", C.c_int32), ("l", C.c_int32), ("m", C.c_int32),
                ("Ip", C.c_double), ("asymp_coeff", C.c_double), ("quan_ax_theta", C.c_double),
                ("quan_ax_phi", C.c_double), ("soft_core", C.c_double),
                ("a1", C.c_double), ("b1", C.c_double), ("a2", C.c_double), ("b2", C.c_double),
                ("a3", C.c_double), ("b3", C.c_double)]


class SamplerDesc(C.Structure):
    _fields_ = [("init_method", C.c_int32), ("dimension", C.c_int32), ("phase_method", C.c_int32),
                ("rate_prefix", C.c_int32), ("sampling", C.c_int32), ("sample_t_num", C.c_int32),
                ("ss_kd_num", C.c_int32), ("ss_kz_num", C.c_int32), ("mc_kt_num", C.c_int64),
                ("sample_t_lo", C.c_double), ("sample_t_hi", C.c_double),
                ("ss_kd_max", C.c_double), ("ss_kz_max", C.c_double),
                ("mc_kd_max", C.c_double), ("mc_kz_max", C.c_double),
                ("kd_sigma", C.c_double), ("kz_sigma", C.c_double),
                ("cutoff_limit", C.c_double), ("seed", C.c_uint64), ("mol_rot", C.c_double * 3),
                ("wfat_na", C.c_int32), ("wfat_nb", C.c_int32), ("wfat_coeff", C.POINTER(C.c_double)),
                ("wfat_form", C.c_int32), ("wfat_dims", C.c_int32 * 4), ("wfat_reserved", C.c_int32),
                ("wfat_mu", C.c_double * 3), ("orbit_ip", C.c_double),
                ("asymp_lmax", C.c_int32), ("asymp_reserved", C.c_int32), ("asymp_coeff", C.POINTER(C.c_double))]


class RunDesc(C.Structure):
    _fields_ = [("dimension", C.c_int32), ("phase_method", C.c_int32), ("integrator", C.c_int32),
                ("precision", C.c_int32), ("traj_t_final", C.c_double), ("dt", C.c_double),
                ("rtol", C.c_double), ("atol", C.c_double), ("restep_radius", C.c_double),
                ("final_p_max", C.c_double * 3), ("final_p_num", C.c_int32 * 3),
                ("quirks", C.c_int32), ("maxiters", C.c_int64), ("pmd_mode", C.c_int32),
                ("pmd_binning", C.c_int32), ("hit_counts", C.POINTER(C.c_uint64))]


class Stats(C.Structure):
    _fields_ = [("ion_prob", C.c_double), ("ion_prob_uncollected", C.c_double),
                ("num_effective_traj", C.c_int64), ("n_failed", C.c_int64), ("n_bound", C.c_int64),
                ("n_restep", C.c_int64), ("n_steps", C.c_int64), ("n_steps_fp64", C.c_int64),
                ("ms_sample", C.c_double),
                ("ms_integrate", C.c_double), ("ms_integrate_main", C.c_double), ("ms_total", C.c_double),
                ("n_slabs", C.c_int64), ("n_attempts", C.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("bin_tiles", C.c_int32), ("bin_tile_shift", C.c_int32), ("bin_chunk", C.c_int32),
                ("overlap_fp64", C.c_int32), ("rk4_block", C.c_int32), ("tsit5_refill", C.c_int32),
                ("tsit5_claim", C.c_int32), ("graph", C.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


P = C.POINTER
DP = P(C.c_double)

# name -> (restype, argtypes); every symbol include/etraj.h declares
SIGNATURES = {
    "etraj_abi_version": (C.c_int, []),
    "etraj_ctx_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
    "etraj_ctx_destroy": (None, [C.c_void_p]),
    "etraj_last_error": (C.c_char_p, [C.c_void_p]),
    "etraj_laser_field": (C.c_int, [C.c_void_p, P(LaserDesc), DP, C.c_int64, DP, DP, DP, DP]),
    "etraj_target_eval": (C.c_int, [C.c_void_p, P(TargetDesc), DP, C.c_int64, DP, DP]),
    "etraj_batch_times": (C.c_int, [C.c_void_p, P(SamplerDesc), DP]),
    "etraj_batch_max_size": (C.c_int64, [P(SamplerDesc)]),
    "etraj_gen_electron_batch": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(SamplerDesc), C.c_int32,
                                           DP, C.c_int64, P(C.c_int64)]),
    "etraj_integrate_and_bin": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(RunDesc), DP, C.c_int64,
                                          C.c_int32, C.c_void_p, DP, DP, P(C.c_int64)]),
    "etraj_integrate_and_bin_noalloc": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(RunDesc), DP,
                                                  C.c_int64, C.c_int32, C.c_void_p, DP, DP, P(C.c_int64)]),
    "etraj_integrate_records": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(RunDesc), DP, C.c_int64,
                                          C.c_int32, DP, P(C.c_int32)]),
    "etraj_run_job": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc), DP,
                                P(Stats)]),
    "etraj_run_job_device": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc),
                                       P(DP), P(Stats)]),
    "etraj_run_scan": (C.c_int, [C.c_void_p, C.c_int32, P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc), DP,
                                 P(Stats)]),
    "etraj_run_scan_device": (C.c_int, [C.c_void_p, C.c_int32, P(LaserDesc), P(TargetDesc), P(SamplerDesc),
                                        P(RunDesc), P(DP), P(Stats)]),
    "etraj_get_tuning": (C.c_int, [C.c_void_p, P(Tuning)]),
    "etraj_set_tuning": (C.c_int, [C.c_void_p, P(Tuning)]),
    "etraj_autotune": (C.c_int, [C.c_void_p, P(LaserDesc), P(TargetDesc), P(SamplerDesc), P(RunDesc), P(Tuning),
                                 DP, DP]),
}


class EngineUnavailable(RuntimeError):
    """libetraj_hip.so is missing or no HIP device is usable."""


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[etraj] error {code}: {msg}")
        self.code = code


class AssertionFailure(AssertionError):
    """Argument validation failure (the reference raises AssertionError)."""


_lib = None
_lib_lock = threading.Lock()


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libetraj_hip.so (no fallback) and binds every declared symbol."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise EngineUnavailable(f"{path} is not built; run __graft_entry__.build() or make -C etraj.jl_amd/csrc")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.etraj_abi_version() != ABI_VERSION:
            raise EngineUnavailable("ABI version mismatch")
        _lib = lib
        return lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(DP) if a is not None else None


class Context:
    """One etraj_ctx (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.etraj_ctx_create(device, C.byref(h))
        if rc != OK:
            raise EngineUnavailable(f"etraj_ctx_create(device={device}) failed with code {rc} (no usable HIP device?)")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.etraj_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_tuning(self) -> dict:
        t = Tuning()
        self.check(self.lib.etraj_get_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **choices) -> dict:
        """Changes the named launch choices (etraj_tuning fields) and returns the full set."""
        t = Tuning()
        self.check(self.lib.etraj_get_tuning(self.handle, C.byref(t)))
        for k, v in choices.items():
            if k not in t.as_dict():
                raise ValueError(f"unknown tuning field {k}")
            setattr(t, k, int(v))
        self.check(self.lib.etraj_set_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def check(self, rc: int):
        if rc != OK:
            msg = self.lib.etraj_last_error(self.handle).decode(errors="replace")
            if rc == EINVAL:
                raise AssertionFailure(msg)
            raise EngineError(rc, msg)


_ctx = None


def default_context() -> Context:
    global _ctx
    if _ctx is None:
        _ctx = Context(int(os.environ.get("ETRAJ_DEVICE", "0")))
    return _ctx
