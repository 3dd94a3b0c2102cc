"""A small HDF5 binding (ctypes over the HDF5 C library, 1.10 ABI) for the
reference's on-disk format: groups, and datasets of float64 / int32 /
int64 arrays and scalars and of fixed-length strings, optionally deflated.

The library is found through $UGKS_HDF5_LIB, the loader's search path, or a
conda installation; without it every call raises H5Error.  h5py is not
needed (and not present here)."""
from __future__ import annotations

import ctypes as C
import ctypes.util
import glob
import os

import numpy as np

hid_t = C.c_int64
hsize_t = C.c_uint64
H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5I_GROUP, H5I_DATASET = 2, 5


class H5Error(RuntimeError):
    pass


_lib = None


def _candidates():
    env = os.environ.get("UGKS_HDF5_LIB")
    if env:
        yield env
    name = ctypes.util.find_library("hdf5")
    if name:
        yield name
    for pat in ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*",
                "/usr/lib/x86_64-linux-gnu/libhdf5.so*", "/usr/local/lib/libhdf5.so*"):
        yield from sorted(glob.glob(pat))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for path in _candidates():
        try:
            L = C.CDLL(path)
            break
        except OSError as e:
            err = e
    else:
        raise H5Error(f"the HDF5 C library was not found (set UGKS_HDF5_LIB): {err}")
    maj, mi, rel = C.c_uint(), C.c_uint(), C.c_uint()
    L.H5get_libversion(C.byref(maj), C.byref(mi), C.byref(rel))
    if (maj.value, mi.value) < (1, 10):
        raise H5Error(f"HDF5 {maj.value}.{mi.value} found; 1.10 or later is needed (64-bit identifiers)")
    if L.H5open() < 0:
        raise H5Error("H5open failed")
    sig = {
        "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]), "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
        "H5Fclose": (C.c_int, [hid_t]), "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]), "H5Gclose": (C.c_int, [hid_t]),
        "H5Screate": (hid_t, [C.c_int]), "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(hsize_t), C.c_void_p]),
        "H5Sclose": (C.c_int, [hid_t]), "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(hsize_t), C.c_void_p]),
        "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]), "H5Dclose": (C.c_int, [hid_t]),
        "H5Dwrite": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dread": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]), "H5Dget_type": (hid_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]), "H5Tset_size": (C.c_int, [hid_t, C.c_size_t]), "H5Tclose": (C.c_int, [hid_t]),
        "H5Tget_class": (C.c_int, [hid_t]), "H5Tget_size": (C.c_size_t, [hid_t]), "H5Tget_sign": (C.c_int, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]), "H5Pclose": (C.c_int, [hid_t]),
        "H5Pset_chunk": (C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t)]), "H5Pset_deflate": (C.c_int, [hid_t, C.c_uint]),
        "H5Lexists": (C.c_int, [hid_t, C.c_char_p, hid_t]), "H5Oopen": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Oclose": (C.c_int, [hid_t]), "H5Iget_type": (C.c_int, [hid_t]),
        "H5Gget_info": (C.c_int, [hid_t, C.c_void_p]),
        "H5Lget_name_by_idx": (C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t,
                                             hid_t]),
        "H5Eset_auto2": (C.c_int, [hid_t, C.c_void_p, C.c_void_p]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    L.H5Eset_auto2(0, None, None)           # errors come back as return codes, not printed
    _lib = L
    return L


def _g(name):
    return hid_t.in_dll(lib(), name).value


def _check(v, what):
    if v < 0:
        raise H5Error(f"HDF5: {what} failed")
    return v


class _GroupInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_uint)]


class Group:
    """An open group (or the file's root group)."""

    def __init__(self, hid, path="/"):
        self.id, self.path = hid, path

    def __contains__(self, name: str) -> bool:
        L, cur = lib(), self.id
        parts = [p for p in name.split("/") if p]
        prefix = ""
        for p in parts:        # H5Lexists needs every intermediate link to exist
            prefix = p if not prefix else prefix + "/" + p
            if L.H5Lexists(cur, prefix.encode(), 0) <= 0:
                return False
        return True

    def keys(self) -> list:
        L = lib()
        info = _GroupInfo()
        _check(L.H5Gget_info(self.id, C.byref(info)), "H5Gget_info")
        out = []
        for k in range(info.nlinks):
            n = L.H5Lget_name_by_idx(self.id, b".", 0, 0, k, None, 0, 0)
            buf = C.create_string_buffer(n + 1)
            L.H5Lget_name_by_idx(self.id, b".", 0, 0, k, buf, n + 1, 0)
            out.append(buf.value.decode())
        return out

    def create_group(self, name: str) -> "Group":
        return Group(_check(lib().H5Gcreate2(self.id, name.encode(), 0, 0, 0), f"create group {name}"),
                     self.path.rstrip("/") + "/" + name)

    def require_group(self, name: str) -> "Group":
        if name in self:
            return Group(_check(lib().H5Gopen2(self.id, name.encode(), 0), f"open group {name}"),
                         self.path.rstrip("/") + "/" + name)
        return self.create_group(name)

    def close(self):
        if self.id is not None and self.path != "/":
            lib().H5Gclose(self.id)
        self.id = None

    def create_dataset(self, name: str, data, compression=None):
        """float64 / int32 / int64 / bool arrays or scalars, str, or sequences of str."""
        L = lib()
        if isinstance(data, (str, bytes)) or (isinstance(data, (list, tuple, np.ndarray)) and len(data) > 0 and
                                              isinstance(np.asarray(data).flat[0], (str, np.str_, bytes))):
            self._write_strings(name, data)
            return
        arr = np.asarray(data)
        if arr.dtype == np.bool_:
            arr = arr.astype(np.int64)
        # (np.array keeps 0-d arrays 0-d; ascontiguousarray would make them 1-d)
        if arr.dtype.kind == "f":
            arr, mt = np.array(arr, dtype=np.float64, order="C"), _g("H5T_NATIVE_DOUBLE_g")
        elif arr.dtype == np.int32:
            arr, mt = np.array(arr, order="C"), _g("H5T_NATIVE_INT_g")
        elif arr.dtype.kind in "iu":
            arr, mt = np.array(arr, dtype=np.int64, order="C"), _g("H5T_NATIVE_LLONG_g")
        else:
            raise H5Error(f"dataset {name}: unsupported dtype {arr.dtype}")
        if arr.ndim == 0:
            space = L.H5Screate(H5S_SCALAR)
        else:
            dims = (hsize_t * arr.ndim)(*arr.shape)
            space = L.H5Screate_simple(arr.ndim, dims, None)
        _check(space, f"dataspace for {name}")
        dcpl = H5P_DEFAULT
        if compression is not None and arr.ndim > 0 and arr.size > 0:
            level = 4 if compression in (True, "gzip") else int(compression)
            dcpl = _check(L.H5Pcreate(_g("H5P_CLS_DATASET_CREATE_ID_g")), "H5Pcreate")
            _check(L.H5Pset_chunk(dcpl, arr.ndim, (hsize_t * arr.ndim)(*arr.shape)), "H5Pset_chunk")
            _check(L.H5Pset_deflate(dcpl, level), "H5Pset_deflate")
        try:
            d = _check(L.H5Dcreate2(self.id, name.encode(), mt, space, 0, dcpl, 0), f"create dataset {name}")
            try:
                _check(L.H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), f"write {name}")
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Sclose(space)
            if dcpl != H5P_DEFAULT:
                L.H5Pclose(dcpl)

    def _write_strings(self, name, data):
        L = lib()
        if isinstance(data, (str, bytes)):
            items, shape = [data], ()
        else:
            a = np.asarray(data)
            items, shape = [x for x in a.flat], a.shape
        raw = [x.encode() if isinstance(x, str) else bytes(x) for x in items]
        size = max(1, max(len(r) for r in raw))
        buf = b"".join(r.ljust(size, b"\0") for r in raw)
        t = _check(L.H5Tcopy(_g("H5T_C_S1_g")), "H5Tcopy")
        try:
            _check(L.H5Tset_size(t, size), "H5Tset_size")
            space = L.H5Screate(H5S_SCALAR) if shape == () else \
                L.H5Screate_simple(len(shape), (hsize_t * len(shape))(*shape), None)
            _check(space, f"dataspace for {name}")
            try:
                d = _check(L.H5Dcreate2(self.id, name.encode(), t, space, 0, 0, 0), f"create dataset {name}")
                try:
                    _check(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), f"write {name}")
                finally:
                    L.H5Dclose(d)
            finally:
                L.H5Sclose(space)
        finally:
            L.H5Tclose(t)

    def __getitem__(self, name: str):
        """A group, or the dataset's data (numpy array, numpy scalar or str)."""
        L = lib()
        if name not in self:
            raise KeyError(f"{self.path}: no member {name}")
        o = _check(L.H5Oopen(self.id, name.encode(), 0), f"open {name}")
        kind = L.H5Iget_type(o)
        L.H5Oclose(o)
        if kind == H5I_GROUP:
            return Group(_check(L.H5Gopen2(self.id, name.encode(), 0), f"open group {name}"),
                         self.path.rstrip("/") + "/" + name)
        if kind != H5I_DATASET:
            raise H5Error(f"{name}: not a group or dataset")
        d = _check(L.H5Dopen2(self.id, name.encode(), 0), f"open dataset {name}")
        try:
            space = L.H5Dget_space(d)
            nd = L.H5Sget_simple_extent_ndims(space)
            dims = (hsize_t * max(nd, 1))()
            if nd > 0:
                L.H5Sget_simple_extent_dims(space, dims, None)
            L.H5Sclose(space)
            shape = tuple(int(dims[k]) for k in range(nd))
            t = L.H5Dget_type(d)
            cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
            try:
                if cls == H5T_STRING:
                    n = int(np.prod(shape)) if shape else 1
                    buf = C.create_string_buffer(n * size)
                    _check(L.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), f"read {name}")
                    vals = [buf.raw[k * size:(k + 1) * size].split(b"\0", 1)[0].decode() for k in range(n)]
                    return vals[0] if not shape else np.array(vals, dtype=object).reshape(shape)
                if cls == H5T_FLOAT:
                    out, mt = np.empty(shape, dtype=np.float64), _g("H5T_NATIVE_DOUBLE_g")
                elif cls == H5T_INTEGER and size <= 4:
                    out, mt = np.empty(shape, dtype=np.int32), _g("H5T_NATIVE_INT_g")
                elif cls == H5T_INTEGER:
                    out, mt = np.empty(shape, dtype=np.int64), _g("H5T_NATIVE_LLONG_g")
                else:
                    raise H5Error(f"{name}: unsupported stored type class {cls}")
                _check(L.H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), f"read {name}")
                return out[()] if not shape else out
            finally:
                L.H5Tclose(t)
        finally:
            L.H5Dclose(d)


class File(Group):
    """h5py-like file: File(path, "w") creates (truncating), File(path) reads."""

    def __init__(self, path: str, mode: str = "r"):
        L = lib()
        if mode == "w":
            hid = L.H5Fcreate(os.fspath(path).encode(), H5F_ACC_TRUNC, 0, 0)
        elif mode == "r":
            hid = L.H5Fopen(os.fspath(path).encode(), H5F_ACC_RDONLY, 0)
        else:
            raise ValueError("mode must be 'r' or 'w'")
        super().__init__(_check(hid, f"open {path}"), "/")
        self.filename = os.fspath(path)

    def close(self):
        if self.id is not None:
            lib().H5Fclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
