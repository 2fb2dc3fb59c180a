"""The `output_fmt = :h5` and `:jld2` files of write_output!
(TrajectorySimulation.jl:502-546), written through the HDF5 C library (the
loader of etraj_hip.moldata).

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

A JLD2 file (the reference's default `output_fmt`, written by JLD2.jl's
jldopen) is an HDF5 file behind a 512-byte user block that opens with the
signature "HDF5-based Julia Data Format, version 0.1.1\0" (the format version
of the reference's own .jld2 files), with a version-2 superblock at byte 512
and HDF5 1.8 object headers. write_jld2 writes that layout: the user block
and signature, the 1.8 file format, the root group's links kept in its object
header, strings as fixed-size null-padded UTF-8 (as JLD2 stores a String), and
momentum_spec deflated without the shuffle filter (JLD2's `compress = true`).
Its datasets are those of the :h5 file. The reference's JLD2 file also holds
`params`, the job's parameters as Julia objects (laser and target structs);
no Julia type can be written from here, and `params_text` holds the same
parameters as YAML.
"""
import ctypes
import os

import numpy as np

from . import __version__
from .moldata import _g, _hdf5

__all__ = ["write_h5", "write_jld2", "file_layout"]

_H5F_ACC_TRUNC = 2
_H5F_LIBVER_V18 = 1
_H5S_SCALAR = 0
JLD2_USERBLOCK = 512
JLD2_SIGNATURE = b"HDF5-based Julia Data Format, version 0.1.1\x00"
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
    lib.H5Pset_userblock.argtypes = [hid, ctypes.c_uint64]
    lib.H5Pset_libver_bounds.argtypes = [hid, ctypes.c_int, ctypes.c_int]
    lib.H5Pset_link_phase_change.argtypes = [hid, ctypes.c_uint, ctypes.c_uint]
    return lib


class _Writer:
    def __init__(self, path, jld2=False):
        lib = self.lib = _lib()
        self.shuffle = not jld2
        fcpl = fapl = 0
        if jld2:
            fcpl = lib.H5Pcreate(_g("H5P_CLS_FILE_CREATE_ID_g"))
            fapl = lib.H5Pcreate(_g("H5P_CLS_FILE_ACCESS_ID_g"))
            ok = (lib.H5Pset_userblock(fcpl, JLD2_USERBLOCK) >= 0 and
                  lib.H5Pset_link_phase_change(fcpl, 64, 48) >= 0 and
                  lib.H5Pset_libver_bounds(fapl, _H5F_LIBVER_V18, _H5F_LIBVER_V18) >= 0)
            if not ok:
                lib.H5Pclose(fcpl)
                lib.H5Pclose(fapl)
                raise OSError("cannot set the JLD2 file layout")
        try:
            self.id = lib.H5Fcreate(os.fsencode(path), _H5F_ACC_TRUNC, fcpl, fapl)
        finally:
            if fcpl:
                lib.H5Pclose(fcpl)
                lib.H5Pclose(fapl)
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
            if self.shuffle:
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


def write_h5(path, result, params_text, compress=True, jld2=False):
    """Writes `result` (the dict perform_traj_simulation returns) to `path` as
    write_output! does for :h5 (jld2: for :jld2), via `path!` renamed when complete."""
    tmp = path + "!"
    w = _Writer(tmp, jld2)
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
    if jld2:
        with open(tmp, "r+b") as fh:
            head = fh.read(JLD2_USERBLOCK)
            if len(head) != JLD2_USERBLOCK or any(head):
                raise OSError(f"{tmp}: the JLD2 user block is missing or not empty")
            info = JLD2_SIGNATURE + f" (etraj_hip {__version__} 64-bit LE)".encode() + b"\x00"
            fh.seek(0)
            fh.write(info.ljust(JLD2_USERBLOCK, b"\x00"))
    os.replace(tmp, path)
    return path


def write_jld2(path, result, params_text, compress=True):
    """write_output! for :jld2 (see the module docstring for the layout)."""
    return write_h5(path, result, params_text, compress, jld2=True)


class _GroupInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", ctypes.c_uint64), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


def file_layout(path, string_dataset):
    """The container facts that tell a JLD2 file from a plain HDF5 file: the user
    block's size and signature (its bytes up to the first NUL), the superblock
    version, how the root group stores its links (0 symbol table, 1 compact,
    2 dense) and the encoding of the scalar string dataset `string_dataset`
    (fixed or variable size, character set, padding)."""
    lib = _lib()
    hid = ctypes.c_int64
    lib.H5Fget_create_plist.restype = hid
    lib.H5Fget_create_plist.argtypes = [hid]
    lib.H5Pget_userblock.argtypes = [hid, ctypes.POINTER(ctypes.c_uint64)]
    lib.H5Gget_info.argtypes = [hid, ctypes.POINTER(_GroupInfo)]
    lib.H5Tget_strpad.argtypes = [hid]
    f = lib.H5Fopen(os.fsencode(path), 0, 0)
    if f < 0:
        raise OSError(f"cannot open {path} as HDF5")
    try:
        pl = lib.H5Fget_create_plist(f)
        ub = ctypes.c_uint64()
        lib.H5Pget_userblock(pl, ctypes.byref(ub))
        lib.H5Pclose(pl)
        gi = _GroupInfo()
        if lib.H5Gget_info(f, ctypes.byref(gi)) < 0:
            raise OSError("H5Gget_info failed")
        d = lib.H5Dopen2(f, string_dataset.encode(), 0)
        if d < 0:
            raise OSError(f"no dataset {string_dataset}")
        t = lib.H5Dget_type(d)
        string = {"class": int(lib.H5Tget_class(t)), "variable": int(lib.H5Tis_variable_str(t)),
                  "cset": int(lib.H5Tget_cset(t)), "strpad": int(lib.H5Tget_strpad(t))}
        lib.H5Tclose(t)
        lib.H5Dclose(d)
    finally:
        lib.H5Fclose(f)
    with open(path, "rb") as fh:
        head = fh.read(int(ub.value) + 16)
    sig = head[:int(ub.value)].split(b"\x00", 1)[0] if ub.value else b""
    return {"userblock": int(ub.value), "signature": sig.decode("latin-1"),
            "superblock_version": head[int(ub.value) + 8] if len(head) > ub.value + 8 else -1,
            "root_link_storage": int(gi.storage_type), "string": string}
