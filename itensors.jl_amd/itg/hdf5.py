"""HDF5 read / write of ITensors in the layout ITensors.jl and NDTensors use,
so that files move between the two.

Layout (each object a group carrying string attribute "type" and Int64
attribute "version" = 1; scalars are 0-d datasets, strings fixed-length
UTF-8 NULLTERM, as HDF5.jl writes them):

* ITensor (ext/ITensorsHDF5Ext/itensor.jl:6-16): "inds" (IndexSet),
  "storage" (read also accepts the older key "store", itensor.jl:27-38);
* IndexSet (indexset.jl:4-15): "length", "index_1" ... "index_N";
* Index (index.jl:4-22): "id" (UInt64), "dim", "dir", "plev" (Int64),
  "tags" (TagSet), attribute "space_type" = "Int" | "QNBlocks" and, for
  QNBlocks, "space";
* QNBlocks (qnindex.jl:4-15): "length", "dims" (Int64 vector), "QN[n]";
* QN (qn.jl:4-15): "names" (4 strings), "vals", "mods" (4 Int64 each),
  unused slots ("", 0, 0), named values sorted by name;
* TagSet (tagset.jl:4-9): "tags" = the tag string;
* Dense{T} (NDTensors/ext/NDTensorsHDF5Ext/dense.jl:4-16): "data";
* BlockSparse{T} (blocksparse.jl:37-52): "ndims", "data", "offsets" =
  per block its N 1-based block coordinates then its 0-based offset;
* Diag{T,Vector{T}} / Diag{T,T} (diag.jl:4-16, tensors without QNs): "data",
  the diagonal, or one value (delta) as a 0-d dataset. The reference has no
  HDF5 form for DiagBlockSparse (QN) storage, so neither has this module.

ComplexF64 data is an HDF5 compound {r, i} of doubles (HDF5.jl's complex
type); read also accepts the C++ ITensor form, a (n, 2) double array
flagged by the attribute "__complex__" (dense.jl:30-35).

The file format work runs on the host through libhdf5 (ctypes, the same C
library HDF5.jl calls); tensor data moves to and from the device with one
copy per tensor.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

from . import ndtensors as ND
from .itensors import ITensor, Index, QN, tagstring

__all__ = ["h5open", "read", "write", "TagSet", "IndexSet", "H5Error"]

_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5F_ACC_RDONLY, _H5F_ACC_RDWR, _H5F_ACC_TRUNC = 0, 1, 2
_H5S_SCALAR, _H5S_SIMPLE, _H5S_NULL = 0, 1, 2
_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING, _H5T_COMPOUND = 0, 1, 3, 6
_H5T_CSET_UTF8, _H5T_STR_NULLTERM = 1, 0
_H5T_VARIABLE = ctypes.c_size_t(-1).value


class H5Error(RuntimeError):
    pass


class _Lib:
    """The libhdf5 entry points used here, with their C signatures."""

    def __init__(self):
        path = os.environ.get("ITG_HDF5_LIB") or ctypes.util.find_library("hdf5")
        cands = [path] if path else []
        cands += ["/opt/conda/lib/libhdf5.so", "libhdf5.so"]
        err = None
        for c in cands:
            try:
                self.L = ctypes.CDLL(c)
                break
            except OSError as e:
                err = e
        else:
            raise H5Error(f"libhdf5 not found (set ITG_HDF5_LIB): {err}")
        L = self.L
        hid, herr, hsize = ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
        P, S = ctypes.c_void_p, ctypes.c_char_p
        HS = ctypes.POINTER(hsize)
        sig = {
            "H5open": (herr, []),
            "H5Fcreate": (hid, [S, ctypes.c_uint, hid, hid]),
            "H5Fopen": (hid, [S, ctypes.c_uint, hid]),
            "H5Fclose": (herr, [hid]),
            "H5Gcreate2": (hid, [hid, S, hid, hid, hid]),
            "H5Gopen2": (hid, [hid, S, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Lexists": (ctypes.c_int, [hid, S, hid]),
            "H5Screate": (hid, [ctypes.c_int]),
            "H5Screate_simple": (hid, [ctypes.c_int, HS, HS]),
            "H5Sclose": (herr, [hid]),
            "H5Sget_simple_extent_type": (ctypes.c_int, [hid]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [hid, HS, HS]),
            "H5Dcreate2": (hid, [hid, S, hid, hid, hid, hid, hid]),
            "H5Dopen2": (hid, [hid, S, hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, P]),
            "H5Dread": (herr, [hid, hid, hid, hid, hid, P]),
            "H5Dget_space": (hid, [hid]),
            "H5Dget_type": (hid, [hid]),
            "H5Dvlen_reclaim": (herr, [hid, hid, hid, P]),
            "H5Dclose": (herr, [hid]),
            "H5Acreate2": (hid, [hid, S, hid, hid, hid, hid]),
            "H5Aopen": (hid, [hid, S, hid]),
            "H5Awrite": (herr, [hid, hid, P]),
            "H5Aread": (herr, [hid, hid, P]),
            "H5Aexists": (ctypes.c_int, [hid, S]),
            "H5Aget_type": (hid, [hid]),
            "H5Aclose": (herr, [hid]),
            "H5Tcopy": (hid, [hid]),
            "H5Tcreate": (hid, [ctypes.c_int, ctypes.c_size_t]),
            "H5Tinsert": (herr, [hid, S, ctypes.c_size_t, hid]),
            "H5Tset_size": (herr, [hid, ctypes.c_size_t]),
            "H5Tset_cset": (herr, [hid, ctypes.c_int]),
            "H5Tset_strpad": (herr, [hid, ctypes.c_int]),
            "H5Tget_class": (ctypes.c_int, [hid]),
            "H5Tget_size": (ctypes.c_size_t, [hid]),
            "H5Tis_variable_str": (ctypes.c_int, [hid]),
            "H5Tget_nmembers": (ctypes.c_int, [hid]),
            "H5Tget_member_name": (P, [hid, ctypes.c_uint]),
            "H5Tget_sign": (ctypes.c_int, [hid]),
            "H5Tclose": (herr, [hid]),
            "H5Pcreate": (hid, [hid]),
            "H5Pset_chunk": (herr, [hid, ctypes.c_int, HS]),
            "H5Pset_deflate": (herr, [hid, ctypes.c_uint]),
            "H5Pclose": (herr, [hid]),
            "H5free_memory": (herr, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        if L.H5open() < 0:
            raise H5Error("H5open failed")

        def g(name):
            return ctypes.c_int64.in_dll(L, name).value

        self.INT64 = g("H5T_NATIVE_INT64_g")
        self.UINT64 = g("H5T_NATIVE_UINT64_g")
        self.DOUBLE = g("H5T_NATIVE_DOUBLE_g")
        self.C_S1 = g("H5T_C_S1_g")
        self.DCPL = g("H5P_CLS_DATASET_CREATE_ID_g")

    def __getattr__(self, name):
        f = getattr(self.L, name)

        def call(*a):
            r = f(*a)
            if isinstance(r, int) and r < 0:
                raise H5Error(f"{name} failed")
            return r
        return call


_lib = None


def _h5():
    global _lib
    if _lib is None:
        _lib = _Lib()
    return _lib


def _b(s):
    return s.encode("utf-8")


# ------------------------------------------------------------------ handles
class _Node:
    def __init__(self, hid, file):
        self.hid = hid
        self.file = file

    def __contains__(self, name):
        return _h5().H5Lexists(self.hid, _b(name), _H5P_DEFAULT) > 0

    def group(self, name):
        return _Group(_h5().H5Gopen2(self.hid, _b(name), _H5P_DEFAULT), self.file)

    def create_group(self, name):
        return _Group(_h5().H5Gcreate2(self.hid, _b(name), _H5P_DEFAULT, _H5P_DEFAULT,
                                       _H5P_DEFAULT), self.file)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _Group(_Node):
    def close(self):
        if self.hid is not None:
            _h5().H5Gclose(self.hid)
            self.hid = None


class H5File(_Node):
    def __init__(self, path, mode):
        H = _h5()
        if mode == "w":
            hid = H.H5Fcreate(_b(path), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        elif mode in ("r", "r+"):
            hid = H.H5Fopen(_b(path), _H5F_ACC_RDONLY if mode == "r" else _H5F_ACC_RDWR,
                            _H5P_DEFAULT)
        else:
            raise ValueError(f"mode {mode!r} (use 'r', 'r+' or 'w')")
        super().__init__(hid, self)
        self.path = path

    def close(self):
        if self.hid is not None:
            _h5().H5Fclose(self.hid)
            self.hid = None


def h5open(path, mode="r"):
    """Open an HDF5 file ('r', 'r+' or 'w'); use as a context manager."""
    return H5File(str(path), mode)


# ------------------------------------------------------- datatypes / values
def _str_type(s):
    H = _h5()
    t = H.H5Tcopy(H.C_S1)
    H.H5Tset_size(t, max(len(_b(s)), 1))
    H.H5Tset_cset(t, _H5T_CSET_UTF8)
    H.H5Tset_strpad(t, _H5T_STR_NULLTERM)
    return t


def _complex_type():
    H = _h5()
    t = H.H5Tcreate(_H5T_COMPOUND, 16)
    H.H5Tinsert(t, b"r", 0, H.DOUBLE)
    H.H5Tinsert(t, b"i", 8, H.DOUBLE)
    return t


def _write_attr_str(node, name, s):
    H = _h5()
    t, sp = _str_type(s), H.H5Screate(_H5S_SCALAR)
    a = H.H5Acreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, _H5P_DEFAULT)
    buf = ctypes.create_string_buffer(_b(s), max(len(_b(s)), 1))
    H.H5Awrite(a, t, buf)
    H.H5Aclose(a)
    H.H5Sclose(sp)
    H.H5Tclose(t)


def _write_attr_int(node, name, v):
    H = _h5()
    sp = H.H5Screate(_H5S_SCALAR)
    a = H.H5Acreate2(node.hid, _b(name), H.INT64, sp, _H5P_DEFAULT, _H5P_DEFAULT)
    x = ctypes.c_int64(int(v))
    H.H5Awrite(a, H.INT64, ctypes.byref(x))
    H.H5Aclose(a)
    H.H5Sclose(sp)


def _read_str(read_fn, get_type, hid):
    """A string attribute or dataset, fixed-length or variable-length."""
    H = _h5()
    ft = get_type(hid)
    try:
        if H.H5Tis_variable_str(ft) > 0:
            mt = H.H5Tcopy(H.C_S1)
            H.H5Tset_size(mt, _H5T_VARIABLE)
            H.H5Tset_cset(mt, _H5T_CSET_UTF8)
            p = ctypes.c_char_p()
            read_fn(hid, mt, ctypes.byref(p))
            s = p.value.decode("utf-8") if p.value is not None else ""
            H.H5free_memory(ctypes.cast(p, ctypes.c_void_p))
            H.H5Tclose(mt)
            return s
        n = H.H5Tget_size(ft)
        buf = ctypes.create_string_buffer(n + 1)
        read_fn(hid, ft, buf)
        return buf.raw[:n].split(b"\0", 1)[0].decode("utf-8")
    finally:
        H.H5Tclose(ft)


def _attr(node, name):
    """A scalar attribute (string or Int64); None when absent."""
    H = _h5()
    if H.H5Aexists(node.hid, _b(name)) <= 0:
        return None
    a = H.H5Aopen(node.hid, _b(name), _H5P_DEFAULT)
    try:
        ft = H.H5Aget_type(a)
        cls = H.H5Tget_class(ft)
        H.H5Tclose(ft)
        if cls == _H5T_STRING:
            return _read_str(lambda h, t, b: H.H5Aread(h, t, b), H.H5Aget_type, a)
        x = ctypes.c_int64()
        H.H5Aread(a, H.INT64, ctypes.byref(x))
        return x.value
    finally:
        H.H5Aclose(a)


def _tag(node, typename):
    _write_attr_str(node, "type", typename)
    _write_attr_int(node, "version", 1)


def _check_type(node, expected, what):
    t = _attr(node, "type")
    if t != expected:
        raise H5Error(f"HDF5 group does not contain {what} data (type {t!r}, expected {expected!r})")


def _write_scalar(node, name, v, kind="i64"):
    H = _h5()
    sp = H.H5Screate(_H5S_SCALAR)
    if kind == "str":
        t = _str_type(v)
        d = H.H5Dcreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT)
        H.H5Dwrite(d, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                   ctypes.create_string_buffer(_b(v), max(len(_b(v)), 1)))
        H.H5Tclose(t)
    elif kind == "c128":
        t = _complex_type()
        x = (ctypes.c_double * 2)(complex(v).real, complex(v).imag)
        d = H.H5Dcreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT)
        H.H5Dwrite(d, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, x)
        H.H5Tclose(t)
    else:
        t = {"u64": H.UINT64, "f64": H.DOUBLE}.get(kind, H.INT64)
        x = {"u64": ctypes.c_uint64, "f64": ctypes.c_double}.get(kind, ctypes.c_int64)(
            float(v) if kind == "f64" else int(v))
        d = H.H5Dcreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT)
        H.H5Dwrite(d, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, ctypes.byref(x))
    H.H5Dclose(d)
    H.H5Sclose(sp)


def _write_array(node, name, a, compress=None):
    """1-D array: int64, float64 or complex128 (compound {r, i})."""
    H = _h5()
    a = np.ascontiguousarray(a)
    n = a.shape[0]
    dims = (ctypes.c_uint64 * 1)(n)
    sp = H.H5Screate_simple(1, dims, None) if n > 0 else H.H5Screate(_H5S_NULL)
    if a.dtype == np.complex128:
        t, own = _complex_type(), True
    elif a.dtype == np.float64:
        t, own = H.DOUBLE, False
    elif a.dtype == np.int64:
        t, own = H.INT64, False
    else:
        raise TypeError(f"unsupported array dtype {a.dtype}")
    dcpl = _H5P_DEFAULT
    if compress and n > 0:
        dcpl = H.H5Pcreate(H.DCPL)
        H.H5Pset_chunk(dcpl, 1, (ctypes.c_uint64 * 1)(min(n, 1 << 16)))
        H.H5Pset_deflate(dcpl, int(compress))
    d = H.H5Dcreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, dcpl, _H5P_DEFAULT)
    if n > 0:
        H.H5Dwrite(d, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data_as(ctypes.c_void_p))
    H.H5Dclose(d)
    if dcpl != _H5P_DEFAULT:
        H.H5Pclose(dcpl)
    if own:
        H.H5Tclose(t)
    H.H5Sclose(sp)


def _write_str_vector(node, name, strs):
    """Vector{String}: variable-length UTF-8 strings (HDF5.jl's form)."""
    H = _h5()
    t = H.H5Tcopy(H.C_S1)
    H.H5Tset_size(t, _H5T_VARIABLE)
    H.H5Tset_cset(t, _H5T_CSET_UTF8)
    dims = (ctypes.c_uint64 * 1)(len(strs))
    sp = H.H5Screate_simple(1, dims, None)
    keep = [ctypes.create_string_buffer(_b(s)) for s in strs]
    ptrs = (ctypes.c_char_p * len(strs))(*[ctypes.cast(k, ctypes.c_char_p) for k in keep])
    d = H.H5Dcreate2(node.hid, _b(name), t, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT)
    H.H5Dwrite(d, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, ptrs)
    H.H5Dclose(d)
    H.H5Sclose(sp)
    H.H5Tclose(t)


def _dataset(node, name):
    """Read a dataset: scalar int / str, or a 1-D (or C++-complex 2-D) array."""
    H = _h5()
    d = H.H5Dopen2(node.hid, _b(name), _H5P_DEFAULT)
    try:
        sp = H.H5Dget_space(d)
        kind = H.H5Sget_simple_extent_type(sp)
        if kind == _H5S_NULL:
            H.H5Sclose(sp)
            return np.zeros(0)
        nd = H.H5Sget_simple_extent_ndims(sp)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        if nd > 0:
            H.H5Sget_simple_extent_dims(sp, dims, None)
        shape = tuple(int(dims[i]) for i in range(nd))
        ft = H.H5Dget_type(d)
        cls = H.H5Tget_class(ft)
        if cls == _H5T_STRING:
            if kind == _H5S_SCALAR:
                H.H5Tclose(ft)
                H.H5Sclose(sp)
                return _read_str(lambda h, t, b: H.H5Dread(h, t, _H5S_ALL, _H5S_ALL,
                                                           _H5P_DEFAULT, b), H.H5Dget_type, d)
            n = int(np.prod(shape))
            if H.H5Tis_variable_str(ft) > 0:
                mt = H.H5Tcopy(H.C_S1)
                H.H5Tset_size(mt, _H5T_VARIABLE)
                H.H5Tset_cset(mt, _H5T_CSET_UTF8)
                ptrs = (ctypes.c_char_p * n)()
                H.H5Dread(d, mt, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, ptrs)
                out = [(p or b"").decode("utf-8") for p in ptrs]
                H.H5Dvlen_reclaim(mt, sp, _H5P_DEFAULT, ptrs)
                H.H5Tclose(mt)
            else:
                w = H.H5Tget_size(ft)
                buf = ctypes.create_string_buffer(w * n)
                H.H5Dread(d, ft, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, buf)
                out = [buf.raw[k * w:(k + 1) * w].split(b"\0", 1)[0].decode("utf-8")
                       for k in range(n)]
            H.H5Tclose(ft)
            H.H5Sclose(sp)
            return out
        if cls == _H5T_COMPOUND:
            names = []
            for k in range(H.H5Tget_nmembers(ft)):
                p = H.H5Tget_member_name(ft, k)
                names.append(ctypes.cast(p, ctypes.c_char_p).value.decode())
                H.H5free_memory(p)
            if names != ["r", "i"]:
                raise H5Error(f"unsupported compound dataset {name!r}: members {names}")
            mt, dt = _complex_type(), np.complex128
        elif cls == _H5T_FLOAT:
            mt, dt = H.H5Tcopy(H.DOUBLE), np.float64
        elif cls == _H5T_INTEGER:
            unsigned = H.H5Tget_sign(ft) == 0
            mt = H.H5Tcopy(H.UINT64 if unsigned else H.INT64)
            dt = np.uint64 if unsigned else np.int64
        else:
            raise H5Error(f"unsupported dataset class {cls} for {name!r}")
        H.H5Tclose(ft)
        out = np.zeros(shape if shape else (), dtype=dt)
        H.H5Dread(d, mt, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data_as(ctypes.c_void_p))
        H.H5Tclose(mt)
        H.H5Sclose(sp)
        if kind == _H5S_SCALAR:
            return out.item()
        return out
    finally:
        H.H5Dclose(d)


# ------------------------------------------------------------- the objects
class TagSet:
    """A tag set (its canonical string, src/lib/TagSets/src/TagSets.jl)."""

    def __init__(self, tags=""):
        self.tags = tagstring(tags)

    def __eq__(self, other):
        return isinstance(other, TagSet) and self.tags == other.tags

    def __repr__(self):
        return f"TagSet({self.tags!r})"


class IndexSet(tuple):
    """A tuple of Index written / read as an IndexSet."""


def _qn_sorted_vals(q):
    """(name, val, mod) of the active values, sorted by name as QN stores them
    (src/lib/QuantumNumbers/src/qn.jl:60-72)."""
    from .itensors import _tag_key
    vals = [(n, v, m) for n, (v, m) in q.vals.items()]
    vals.sort(key=lambda x: _tag_key(x[0]))
    return vals


def _write_qn(parent, name, q):
    with parent.create_group(name) as g:
        _tag(g, "QN")
        vals = _qn_sorted_vals(q)
        pad = QN.maxQNs - len(vals)
        _write_str_vector(g, "names", [n for n, _v, _m in vals] + [""] * pad)
        _write_array(g, "vals", np.array([v for _n, v, _m in vals] + [0] * pad, dtype=np.int64))
        _write_array(g, "mods", np.array([m for _n, _v, m in vals] + [0] * pad, dtype=np.int64))


def _read_qn(parent, name):
    with parent.group(name) as g:
        _check_type(g, "QN", "QN")
        names, vals, mods = _dataset(g, "names"), _dataset(g, "vals"), _dataset(g, "mods")
        return QN(*[(n, int(v), int(m)) for n, v, m in zip(names, vals, mods) if int(m) != 0])


def _write_index(parent, name, i):
    with parent.create_group(name) as g:
        _tag(g, "Index")
        _write_scalar(g, "id", i.id, "u64")
        _write_scalar(g, "dim", i.dim)
        _write_scalar(g, "dir", int(i.dir))
        with g.create_group("tags") as t:
            _tag(t, "TagSet")
            _write_scalar(t, "tags", tagstring(i.tags), "str")
        _write_scalar(g, "plev", i.plev)
        if not i.hasqns:
            _write_attr_str(g, "space_type", "Int")
        else:
            _write_attr_str(g, "space_type", "QNBlocks")
            with g.create_group("space") as s:
                _tag(s, "QNBlocks")
                _write_scalar(s, "length", len(i.space))
                _write_array(s, "dims", np.array([d for _q, d in i.space], dtype=np.int64))
                for n, (q, _d) in enumerate(i.space, 1):
                    _write_qn(s, f"QN[{n}]", q)


def _read_index(parent, name):
    with parent.group(name) as g:
        _check_type(g, "Index", "Index")
        iid, dim = int(_dataset(g, "id")), int(_dataset(g, "dim"))
        dr, plev = int(_dataset(g, "dir")), int(_dataset(g, "plev"))
        with g.group("tags") as t:
            _check_type(t, "TagSet", "TagSet")
            tags = _dataset(t, "tags")
        st = _attr(g, "space_type") or "Int"
        if st == "Int":
            space = dim
        elif st == "QNBlocks":
            with g.group("space") as s:
                _check_type(s, "QNBlocks", "QNBlocks")
                n, dims = int(_dataset(s, "length")), _dataset(s, "dims")
                space = [(_read_qn(s, f"QN[{k}]"), int(dims[k - 1])) for k in range(1, n + 1)]
        else:
            raise H5Error(f"unknown Index space type {st!r}")
        return Index(space, tags=tags, plev=plev, dir=dr, id=iid)


def _write_indexset(parent, name, inds):
    with parent.create_group(name) as g:
        _tag(g, "IndexSet")
        _write_scalar(g, "length", len(inds))
        for n, i in enumerate(inds, 1):
            _write_index(g, f"index_{n}", i)


def _read_indexset(parent, name):
    with parent.group(name) as g:
        _check_type(g, "IndexSet", "IndexSet")
        n = int(_dataset(g, "length"))
        return IndexSet(_read_index(g, f"index_{k}") for k in range(1, n + 1))


def _eltype_name(dtype):
    import torch
    if dtype == torch.float64:
        return "Float64"
    if dtype == torch.complex128:
        return "ComplexF64"
    raise TypeError(f"unsupported element type {dtype}")


def _write_storage(parent, name, T, compress=None):
    if isinstance(T, ND.DiagBlockSparseTensor) and T.blockdims and any(
            len(bd) != 1 for bd in T.blockdims):
        raise TypeError("HDF5 write of DiagBlockSparse (QN diagonal) storage: the reference "
                        "defines no HDF5 form for it")
    host = None if T.data is None else T.data.detach().cpu().numpy()
    with parent.create_group(name) as g:
        if isinstance(T, ND.BlockSparseTensor):
            _tag(g, f"BlockSparse{{{_eltype_name(T.data.dtype)}}}")
            _write_scalar(g, "ndims", T.ndims)
            _write_array(g, "data", host, compress)
            off = []
            for b, o in T.blockoffsets.items():
                off.extend(int(c) for c in b)
                off.append(int(o))
            _write_array(g, "offsets", np.array(off, dtype=np.int64))
        elif isinstance(T, ND.DenseTensor):
            _tag(g, f"Dense{{{_eltype_name(T.data.dtype)}}}")
            _write_array(g, "data", host, compress)
        elif isinstance(T, ND.DiagBlockSparseTensor):
            el = _eltype_name(T.eltype)
            if T.uniform is not None:
                _tag(g, f"Diag{{{el},{el}}}")
                _write_scalar(g, "data", T.uniform, "c128" if el == "ComplexF64" else "f64")
            else:
                _tag(g, f"Diag{{{el},Vector{{{el}}}}}")
                _write_array(g, "data", host, compress)
        else:
            raise TypeError(f"HDF5 write of {type(T).__name__} storage is not supported")


def _has_dataset_attr(g, dname, aname):
    H = _h5()
    d = H.H5Dopen2(g.hid, _b(dname), _H5P_DEFAULT)
    try:
        return H.H5Aexists(d, _b(aname)) > 0
    finally:
        H.H5Dclose(d)


def _read_storage(parent, name, inds, device):
    import torch
    with parent.group(name) as g:
        t = _attr(g, "type")
        if t and t.startswith("Diag{"):
            return _read_diag(g, t, inds, device)
        for kind in ("Dense", "BlockSparse"):
            for el, dt in (("Float64", np.float64), ("ComplexF64", np.complex128)):
                if t == f"{kind}{{{el}}}":
                    break
            else:
                continue
            break
        else:
            raise H5Error(f"HDF5 storage type {t!r} is not supported (Dense / BlockSparse, "
                          "Float64 / ComplexF64)")
        data = _dataset(g, "data") if not _has_dataset_attr(g, "data", "__complex__") else None
        if data is None:  # C++ ITensor complex form: (n, 2) doubles
            data = np.ascontiguousarray(_dataset(g, "data"), dtype=np.float64).reshape(-1)
            data = data.view(np.complex128)
        data = np.ascontiguousarray(np.asarray(data).reshape(-1), dtype=dt)
        tdata = torch.from_numpy(data).to(device)
        if kind == "Dense":
            dims = tuple(i.dim for i in inds)
            if data.size != int(np.prod(dims)):
                raise H5Error(f"Dense data length {data.size} does not match the indices {dims}")
            return ND.DenseTensor(tdata, dims)
        N = int(_dataset(g, "ndims"))
        off = np.asarray(_dataset(g, "offsets"), dtype=np.int64).reshape(-1)
        if N != len(inds) or off.size % (N + 1):
            raise H5Error("BlockSparse offsets do not match the indices")
        boff = {}
        for k in range(off.size // (N + 1)):
            rec = off[k * (N + 1):(k + 1) * (N + 1)]
            boff[tuple(int(c) for c in rec[:N])] = int(rec[N])
        return ND.BlockSparseTensor(tdata, boff, tuple(i.blockdims() for i in inds))


def _read_diag(g, t, inds, device):
    """Diag{ElT,DataT} (diag.jl:18-44): one diagonal block over the whole
    index set; DataT = ElT means one value on every diagonal element."""
    import torch
    el, datat = t[len("Diag{"):-1].split(",", 1)
    if el not in ("Float64", "ComplexF64") or datat not in (el, f"Vector{{{el}}}"):
        raise H5Error(f"HDF5 storage type {t!r} is not supported")
    cplx = el == "ComplexF64"
    if any(i.hasqns for i in inds):
        raise H5Error("Diag storage over QN indices")
    offs = {(1,) * len(inds): 0}
    bdims = tuple((i.dim,) for i in inds)
    v = _dataset(g, "data")
    if datat == el:
        return ND.DiagBlockSparseTensor(None, offs, bdims, uniform=complex(v) if cplx else float(v),
                                        dtype=torch.complex128 if cplx else torch.float64)
    a = np.ascontiguousarray(np.asarray(v).reshape(-1), dtype=np.complex128 if cplx else np.float64)
    n = min(i.dim for i in inds) if inds else 1
    if a.size != n:
        raise H5Error(f"Diag data length {a.size} does not match the indices (diagonal {n})")
    return ND.DiagBlockSparseTensor(torch.from_numpy(a).to(device), offs, bdims)


# ---------------------------------------------------------------- dispatch
def write(parent, name, obj, compress=None):
    """write(parent, name, x) for an ITensor, Index, IndexSet (tuple of Index),
    TagSet or QN; `compress` (1-9) deflates tensor data, as HDF5.jl's
    `compress` keyword does."""
    if isinstance(obj, ITensor):
        with parent.create_group(name) as g:
            _tag(g, "ITensor")
            _write_indexset(g, "inds", obj.inds)
            _write_storage(g, "storage", obj.tensor, compress)
    elif isinstance(obj, Index):
        _write_index(parent, name, obj)
    elif isinstance(obj, (tuple, list)) and all(isinstance(i, Index) for i in obj):
        _write_indexset(parent, name, obj)
    elif isinstance(obj, TagSet):
        with parent.create_group(name) as g:
            _tag(g, "TagSet")
            _write_scalar(g, "tags", obj.tags, "str")
    elif isinstance(obj, QN):
        _write_qn(parent, name, obj)
    else:
        raise TypeError(f"HDF5 write of {type(obj).__name__} is not supported")


def read(parent, name, T, device=None):
    """read(parent, name, T) for T in ITensor, Index, IndexSet, TagSet, QN.
    ITensor data goes to `device` (default: the current HIP device when
    there is one, else the host)."""
    if T is ITensor:
        import torch
        if device is None:
            device = "cuda" if torch.cuda.is_available() else "cpu"
        with parent.group(name) as g:
            _check_type(g, "ITensor", "ITensor")
            inds = _read_indexset(g, "inds")
            for key in ("storage", "store"):
                if key in g:
                    return ITensor(_read_storage(g, key, inds, device), inds)
            raise H5Error("HDF5 ITensor group holds neither 'storage' nor 'store'")
    if T is Index:
        return _read_index(parent, name)
    if T is IndexSet or T is tuple:
        return _read_indexset(parent, name)
    if T is TagSet:
        with parent.group(name) as g:
            _check_type(g, "TagSet", "TagSet")
            return TagSet(_dataset(g, "tags"))
    if T is QN:
        return _read_qn(parent, name)
    raise TypeError(f"HDF5 read of {T} is not supported")

