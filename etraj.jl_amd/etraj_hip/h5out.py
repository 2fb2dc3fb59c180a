"""The `output_fmt = :h5` file of write_output! (TrajectorySimulation.jl:502-546),
written through the HDF5 C library (the loader of etraj_hip.moldata).

Datasets, as HDF5.jl writes them for the reference:
  info                  "Trajectory simulation output generated by ..." (UTF-8 string,
                        scalar; written when the job opens the file, :239)
  params_text           the YAML text of the run parameters (UTF-8 string, scalar)
  px, py[, pz]          range(-pmax, pmax, length=num) as Float64 vectors
  momentum_spec         the PMD (Float64; |amplitude|^2 for QTMC/SCTS); Julia's
                        size (nx, ny[, nz]) is HDF5's reversed dimensions
                        (ny, nx), px fastest; shuffle + deflate level 6 when
                        output_compress (:526-540)
  ion_prob, ion_prob_uncollected   Float64 scalars
  num_effective_traj    Int64 scalar
The file is written as `path!` and renamed when complete (:199-206, :546).
"""
import ctypes
import os

import numpy as np

from . import __version__
from .moldata import _g, _hdf5

__all__ = ["write_h5"]

_H5F_ACC_TRUNC = 2
_H5S_SCALAR = 0
_H5T_CSET_UTF8 = 1
_H5T_STR_NULLPAD = 1


def _lib():
    lib = _hdf5()
    hid = ctypes.c_int64
    lib.H5Fcreate.restype = hid
    lib.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid, hid]
    lib.H5Screate_simple.restype = hid
    lib.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
    lib.H5Screate.restype = hid
    lib.H5Screate.argtypes = [ctypes.c_int]
    lib.H5Dcreate2.restype = hid
    lib.H5Dcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]
    lib.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    lib.H5Pcreate.restype = hid
    lib.H5Pcreate.argtypes = [hid]
    lib.H5Pset_chunk.argtypes = [hid, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    lib.H5Pset_shuffle.argtypes = [hid]
    lib.H5Pset_deflate.argtypes = [hid, ctypes.c_uint]
    lib.H5Pclose.argtypes = [hid]
    lib.H5Tset_strpad.argtypes = [hid, ctypes.c_int]
    return lib


class _Writer:
    def __init__(self, path):
        self.lib = _lib()
        self.id = self.lib.H5Fcreate(os.fsencode(path), _H5F_ACC_TRUNC, 0, 0)
        if self.id < 0:
            raise OSError(f"cannot create {path}")

    def close(self):
        self.lib.H5Fclose(self.id)

    def _dataset(self, name, ftype, space, dcpl=0):
        d = self.lib.H5Dcreate2(self.id, name.encode(), ftype, space, 0, dcpl, 0)
        if d < 0:
            raise OSError(f"H5Dcreate2 {name} failed")
        return d

    def _put(self, name, ftype, mtype, buf, dims, compress=False):
        lib = self.lib
        if dims:
            cd = (ctypes.c_uint64 * len(dims))(*dims)
            space = lib.H5Screate_simple(len(dims), cd, None)
        else:
            space = lib.H5Screate(_H5S_SCALAR)
        dcpl = 0
        if compress and dims:
            dcpl = lib.H5Pcreate(_g("H5P_CLS_DATASET_CREATE_ID_g"))
            chunk = list(dims)
            while int(np.prod(chunk)) * 8 > (1 << 20) and chunk[0] > 1:   # chunks of <= 1 MiB
                chunk[0] = (chunk[0] + 1) // 2
            lib.H5Pset_chunk(dcpl, len(chunk), (ctypes.c_uint64 * len(chunk))(*chunk))
            lib.H5Pset_shuffle(dcpl)
            lib.H5Pset_deflate(dcpl, 6)
        d = self._dataset(name, ftype, space, dcpl)
        try:
            if lib.H5Dwrite(d, mtype, 0, 0, 0, buf) < 0:
                raise OSError(f"H5Dwrite {name} failed")
        finally:
            lib.H5Dclose(d)
            lib.H5Sclose(space)
            if dcpl:
                lib.H5Pclose(dcpl)

    def array(self, name, a, compress=False):
        """A Julia array given in Julia's index order (numpy, any layout)."""
        a = np.asarray(a, dtype=np.float64)
        buf = np.ascontiguousarray(a.ravel(order="F"))          # Julia's column-major memory
        self._put(name, _g("H5T_IEEE_F64LE_g"), _g("H5T_NATIVE_DOUBLE_g"), buf.ctypes.data, a.shape[::-1], compress)

    def f64(self, name, v):
        buf = np.array([float(v)])
        self._put(name, _g("H5T_IEEE_F64LE_g"), _g("H5T_NATIVE_DOUBLE_g"), buf.ctypes.data, ())

    def i64(self, name, v):
        buf = np.array([int(v)], dtype=np.int64)
        self._put(name, _g("H5T_STD_I64LE_g"), _g("H5T_NATIVE_LLONG_g"), buf.ctypes.data, ())

    def string(self, name, text):
        lib = self.lib
        raw = text.encode("utf-8")
        t = lib.H5Tcopy(_g("H5T_C_S1_g"))
        lib.H5Tset_size(t, max(len(raw), 1))
        lib.H5Tset_cset(t, _H5T_CSET_UTF8)
        lib.H5Tset_strpad(t, _H5T_STR_NULLPAD)
        buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
        try:
            self._put(name, t, t, ctypes.addressof(buf), ())
        finally:
            lib.H5Tclose(t)


def write_h5(path, result, params_text, compress=True):
    """Writes `result` (the dict perform_traj_simulation returns) to `path` as
    write_output! does for :h5, via `path!` renamed when complete."""
    tmp = path + "!"
    w = _Writer(tmp)
    try:
        w.string("info", f"Trajectory simulation output generated by etraj_hip @ {__version__}")
        w.string("params_text", params_text)
        w.array("px", result["px"])
        w.array("py", result["py"])
        if "pz" in result:
            w.array("pz", result["pz"])
        w.array("momentum_spec", result["momentum_spec"], compress)
        w.f64("ion_prob", result["ion_prob"])
        w.f64("ion_prob_uncollected", result["ion_prob_uncollected"])
        w.i64("num_effective_traj", result["num_effective_traj"])
    finally:
        w.close()
    os.replace(tmp, path)
    return path
