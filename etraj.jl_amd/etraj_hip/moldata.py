"""Reader of eTraj.jl's molecule database files (`Molecule_*.jld2`).

The reference loads them with JLD2 (`GenericMolecule.jl:99-119`, `LoadMolecule`):
`atoms, atom_coords, charge, spin, name, energy_levels, orbit_occ, wfat_indices,
wfat_intdata, wfat_mu, asymp_coeff_indices, asymp_coeff`. A JLD2 file is an HDF5
file; this module reads it through the HDF5 C library (libhdf5, located by
`ETRAJ_HDF5_LIB`, then the system search path, then /opt/conda/lib), which
executes nothing from the file. JLD2's encodings that occur in these files:

* plain arrays and scalars: HDF5 datasets, dimensions reversed (Julia is
  column-major), read here back into Julia's index order;
* `ComplexF64`: a compound {re, im} of two doubles;
* `Dict{K,V}`: a compound {kvvec: object reference} pointing at a dataset of
  references to compound {first: reference, second: reference} pairs, each
  pointing at the key and the value datasets;
* `Tuple{Int,Int}` (open-shell orbital keys): a compound {"1", "2"} of integers.

Only the fields the trajectory path uses are decoded: energies, occupations,
charge, spin, atoms, the WFAT data (`wfat_intdata[ridx]`, a 4-index complex
array intdata[nxi, m, l, m'], and `wfat_mu[ridx]`, a 3-vector) and the
asymptotic coefficients (`asymp_coeff[ridx]`, C_lm at [l, m + l]).
"""
import ctypes
import ctypes.util
import os

import numpy as np

__all__ = ["HDF5Unavailable", "read_molecule_jld2"]


class HDF5Unavailable(RuntimeError):
    pass


_lib = None


def _hdf5():
    global _lib
    if _lib is not None:
        return _lib
    cands = [os.environ.get("ETRAJ_HDF5_LIB"), ctypes.util.find_library("hdf5"), "/opt/conda/lib/libhdf5.so"]
    for c in cands:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError:
            continue
    else:
        raise HDF5Unavailable("libhdf5 not found (set ETRAJ_HDF5_LIB to its path)")
    hid = ctypes.c_int64
    lib.H5open.restype = ctypes.c_int
    lib.H5Fopen.restype = hid
    lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
    lib.H5Fclose.argtypes = [hid]
    lib.H5Dopen2.restype = hid
    lib.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
    lib.H5Dclose.argtypes = [hid]
    lib.H5Dget_space.restype = hid
    lib.H5Dget_space.argtypes = [hid]
    lib.H5Dget_type.restype = hid
    lib.H5Dget_type.argtypes = [hid]
    lib.H5Sget_simple_extent_ndims.argtypes = [hid]
    lib.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
    lib.H5Sclose.argtypes = [hid]
    lib.H5Tclose.argtypes = [hid]
    lib.H5Tget_class.argtypes = [hid]
    lib.H5Tget_size.restype = ctypes.c_size_t
    lib.H5Tget_size.argtypes = [hid]
    lib.H5Tget_nmembers.argtypes = [hid]
    lib.H5Tget_member_name.restype = ctypes.c_void_p
    lib.H5Tget_member_name.argtypes = [hid, ctypes.c_uint]
    lib.H5Tcreate.restype = hid
    lib.H5Tcreate.argtypes = [ctypes.c_int, ctypes.c_size_t]
    lib.H5Tinsert.argtypes = [hid, ctypes.c_char_p, ctypes.c_size_t, hid]
    lib.H5Tcopy.restype = hid
    lib.H5Tcopy.argtypes = [hid]
    lib.H5Tset_size.argtypes = [hid, ctypes.c_size_t]
    lib.H5Tis_variable_str.argtypes = [hid]
    lib.H5Tget_cset.argtypes = [hid]
    lib.H5Tset_cset.argtypes = [hid, ctypes.c_int]
    lib.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    lib.H5Rdereference2.restype = hid
    lib.H5Rdereference2.argtypes = [hid, hid, ctypes.c_int, ctypes.c_void_p]
    lib.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
    lib.H5free_memory.argtypes = [ctypes.c_void_p]
    lib.H5Dvlen_reclaim.argtypes = [hid, hid, hid, ctypes.c_void_p]
    if lib.H5open() < 0:
        raise HDF5Unavailable("H5open failed")
    _lib = lib
    return lib


def _g(name):
    return ctypes.c_int64.in_dll(_hdf5(), name).value


_H5T_COMPOUND, _H5T_STRING, _H5T_REFERENCE = 6, 3, 7


class _File:
    def __init__(self, path):
        self.lib = _hdf5()
        self.id = self.lib.H5Fopen(os.fsencode(path), 0, 0)   # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.id < 0:
            raise OSError(f"cannot open {path} as HDF5")

    def close(self):
        self.lib.H5Fclose(self.id)

    def exists(self, name):
        return self.lib.H5Lexists(self.id, name.encode(), 0) > 0

    def _shape(self, d):
        sp = self.lib.H5Dget_space(d)
        nd = self.lib.H5Sget_simple_extent_ndims(sp)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        if nd > 0:
            self.lib.H5Sget_simple_extent_dims(sp, dims, None)
        self.lib.H5Sclose(sp)
        return tuple(int(dims[i]) for i in range(nd))

    def _members(self, t):
        out = []
        for i in range(self.lib.H5Tget_nmembers(t)):
            p = self.lib.H5Tget_member_name(t, i)
            out.append(ctypes.string_at(p).decode())
            self.lib.H5free_memory(p)
        return out

    def read(self, d):
        """Value of an open dataset, Julia index order."""
        lib = self.lib
        shape = self._shape(d)
        t = lib.H5Dget_type(d)
        try:
            cls = lib.H5Tget_class(t)
            n = int(np.prod(shape)) if shape else 1
            if cls == _H5T_COMPOUND:
                names = self._members(t)
                if names == ["re", "im"]:
                    mt = lib.H5Tcreate(_H5T_COMPOUND, 16)
                    lib.H5Tinsert(mt, b"re", 0, _g("H5T_NATIVE_DOUBLE_g"))
                    lib.H5Tinsert(mt, b"im", 8, _g("H5T_NATIVE_DOUBLE_g"))
                    buf = np.empty(n, np.complex128)
                    self._read(d, mt, buf.ctypes.data)
                    lib.H5Tclose(mt)
                    return self._julia_order(buf, shape)
                classes = [lib.H5Tget_class(self._member_type(t, i)) for i in range(len(names))]
                if all(c == _H5T_REFERENCE for c in classes):
                    return self._read_refs(d, names, n, shape)
                if all(c == 0 for c in classes):   # a tuple of integers, e.g. Tuple{Int,Int} (members "1", "2")
                    k = len(names)
                    mt = lib.H5Tcreate(_H5T_COMPOUND, 8 * k)
                    for i, nm in enumerate(names):
                        lib.H5Tinsert(mt, nm.encode(), 8 * i, _g("H5T_NATIVE_LLONG_g"))
                    buf = np.empty(n * k, np.int64)
                    self._read(d, mt, buf.ctypes.data)
                    lib.H5Tclose(mt)
                    recs = [tuple(int(v) for v in buf[j * k:(j + 1) * k]) for j in range(n)]
                    return recs[0] if not shape else recs
                raise NotImplementedError(f"compound {names}")
            if cls == _H5T_STRING:
                if lib.H5Tis_variable_str(t) > 0:
                    mt = lib.H5Tcopy(_g("H5T_C_S1_g"))
                    lib.H5Tset_size(mt, ctypes.c_size_t(-1).value)   # H5T_VARIABLE
                    lib.H5Tset_cset(mt, lib.H5Tget_cset(t))
                    ptrs = (ctypes.c_char_p * n)()
                    self._read(d, mt, ctypes.addressof(ptrs))
                    vals = [p.decode() if p else "" for p in ptrs]
                    sp = lib.H5Dget_space(d)
                    lib.H5Dvlen_reclaim(mt, sp, 0, ctypes.addressof(ptrs))
                    lib.H5Sclose(sp)
                    lib.H5Tclose(mt)
                else:
                    size = lib.H5Tget_size(t)
                    buf = ctypes.create_string_buffer(size * n)
                    self._read(d, t, ctypes.addressof(buf))
                    vals = [buf.raw[i * size:(i + 1) * size].rstrip(b"\0").decode() for i in range(n)]
                return vals[0] if not shape else vals
            size = lib.H5Tget_size(t)
            if cls == _H5T_REFERENCE:   # an array of references to separately stored objects
                buf = np.empty(n, np.uint64)
                self._read(d, _g("H5T_STD_REF_OBJ_g"), buf.ctypes.data)
                vals = [self._deref(d, int(r)) for r in buf]
                return vals[0] if not shape else vals
            if cls == 0:      # H5T_INTEGER
                mt, dt = _g("H5T_NATIVE_LLONG_g"), np.int64
            elif cls == 1:    # H5T_FLOAT
                mt, dt = _g("H5T_NATIVE_DOUBLE_g"), np.float64
            else:
                raise NotImplementedError(f"HDF5 class {cls} (size {size})")
            buf = np.empty(n, dt)
            self._read(d, mt, buf.ctypes.data)
            return buf[0] if not shape else self._julia_order(buf, shape)
        finally:
            lib.H5Tclose(t)

    def _member_type(self, t, i):
        self.lib.H5Tget_member_type.restype = ctypes.c_int64
        self.lib.H5Tget_member_type.argtypes = [ctypes.c_int64, ctypes.c_uint]
        return self.lib.H5Tget_member_type(t, i)

    def _read(self, d, mt, ptr):
        if self.lib.H5Dread(d, mt, 0, 0, 0, ptr) < 0:   # H5S_ALL, H5S_ALL, H5P_DEFAULT
            raise OSError("H5Dread failed")

    @staticmethod
    def _julia_order(buf, shape):
        # HDF5 dims are Julia's reversed: the row-major array of shape `shape` read
        # in reversed axis order is Julia's column-major array
        return buf.reshape(shape).transpose(tuple(range(len(shape) - 1, -1, -1)))

    def _deref(self, d, ref):
        r = ctypes.c_uint64(ref)
        o = self.lib.H5Rdereference2(d, 0, 0, ctypes.byref(r))   # H5P_DEFAULT, H5R_OBJECT
        if o < 0:
            raise OSError("H5Rdereference2 failed")
        try:
            return self.read(o)
        finally:
            self.lib.H5Dclose(o)

    def _read_refs(self, d, names, n, shape):
        """A compound whose members are all object references (JLD2's Pair,
        Dict and Set wrappers): each member is dereferenced and read."""
        lib = self.lib
        k = len(names)
        mt = lib.H5Tcreate(_H5T_COMPOUND, 8 * k)
        for i, nm in enumerate(names):
            lib.H5Tinsert(mt, nm.encode(), 8 * i, _g("H5T_STD_REF_OBJ_g"))
        buf = np.empty(n * k, np.uint64)
        self._read(d, mt, buf.ctypes.data)
        lib.H5Tclose(mt)
        recs = [_unwrap({nm: self._deref(d, int(buf[j * k + i])) for i, nm in enumerate(names)}) for j in range(n)]
        return recs[0] if not shape else recs

    def get(self, name):
        d = self.lib.H5Dopen2(self.id, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        try:
            return self.read(d)
        finally:
            self.lib.H5Dclose(d)


def _unwrap(rec):
    """JLD2 wrappers: Dict = {kvvec: [Pair...]}, Pair = {first, second},
    Set = {dict: Dict}."""
    if set(rec) == {"kvvec"}:
        kv = rec["kvvec"]
        kv = kv if isinstance(kv, list) else [kv]
        return {p["first"]: p["second"] for p in kv}
    if set(rec) == {"dict"}:
        return rec["dict"]
    return rec


def read_molecule_jld2(path):
    """The fields of `LoadMolecule` (`GenericMolecule.jl:99-119`) that the
    trajectory path uses, as plain Python / numpy values in Julia's index
    order (0-based): name, atoms, atom_coords (natoms x 3), charge, spin,
    energy_levels, orbit_occ, wfat_intdata {ridx: complex array
    [nxi, m, l, m']}, wfat_mu {ridx: 3-vector}, asymp_coeff {ridx: complex
    array [l, m + l]} (MolAsympCoeff, GenericMolecule.jl:304-309)."""
    f = _File(path)
    try:
        out = {k: f.get(k) for k in ("name", "atoms", "atom_coords", "charge", "spin", "energy_levels",
                                     "orbit_occ")}
        out["wfat_intdata"] = f.get("wfat_intdata") if f.exists("wfat_intdata") else {}
        out["wfat_mu"] = f.get("wfat_μ") if f.exists("wfat_μ") else {}
        out["asymp_coeff"] = f.get("asymp_coeff") if f.exists("asymp_coeff") else {}
    finally:
        f.close()
    for k in ("wfat_intdata", "wfat_mu", "asymp_coeff"):
        if not isinstance(out[k], dict):
            out[k] = {}
    out["charge"] = int(out["charge"])
    out["spin"] = int(out["spin"])
    return out
