"""HDF5 output of the cube and beam (Martini.write_hdf5, martini/martini.py:1154-1366).

The reference writes through h5py; here the HDF5 C library is called directly
through ctypes (1.10 API), so that no Python HDF5 package is needed.  The
file holds what h5py would write for the same calls: datasets of the numpy
arrays' shape and type (C order), string attributes as variable-length UTF-8
strings, float attributes as float64 and integer attributes as int64 scalars.

The library is looked up as ``MARTINI_HDF5_LIB`` (a path), then by
``ctypes.util.find_library("hdf5")``, then in a few usual install places.
"""

from __future__ import annotations

import ctypes
import ctypes.util
import datetime
import glob
import os

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
hsize_t = ctypes.c_uint64

H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5F_ACC_EXCL = 0x0004
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3

_CANDIDATES = ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*",
               "/usr/lib/x86_64-linux-gnu/libhdf5.so*", "/usr/lib64/libhdf5.so*",
               "/usr/local/lib/libhdf5.so*")

_lib = None


def _find() -> str | None:
    path = os.environ.get("MARTINI_HDF5_LIB")
    if path:
        return path
    name = ctypes.util.find_library("hdf5")
    if name:
        return name
    for pattern in _CANDIDATES:
        hits = sorted(p for p in glob.glob(pattern) if "_hl" not in p and "fortran" not in p
                      and "cpp" not in p)
        if hits:
            return hits[0]
    return None


def lib():
    """The loaded HDF5 C library, with the signatures used here (ImportError
    when it cannot be found, as the reference's ``import h5py`` would raise)."""
    global _lib
    if _lib is not None:
        return _lib
    path = _find()
    if path is None:
        raise ImportError("write_hdf5 needs the HDF5 C library (set MARTINI_HDF5_LIB to its path)")
    L = ctypes.CDLL(path)
    vp, cp = ctypes.c_void_p, ctypes.c_char_p
    sig = {
        "H5open": (herr_t, []),
        "H5Fcreate": (hid_t, [cp, ctypes.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [cp, ctypes.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]),
        "H5Pclose": (herr_t, [hid_t]),
        "H5Pset_fapl_core": (herr_t, [hid_t, ctypes.c_size_t, ctypes.c_uint]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), vp]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), vp]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tset_cset": (herr_t, [hid_t, ctypes.c_int]),
        "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tget_size": (ctypes.c_size_t, [hid_t]),
        "H5Tis_variable_str": (ctypes.c_int, [hid_t]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, cp, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, cp, hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, vp]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, vp]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Acreate2": (hid_t, [hid_t, cp, hid_t, hid_t, hid_t, hid_t]),
        "H5Aopen": (hid_t, [hid_t, cp, hid_t]),
        "H5Awrite": (herr_t, [hid_t, hid_t, vp]),
        "H5Aread": (herr_t, [hid_t, hid_t, vp]),
        "H5Aget_type": (hid_t, [hid_t]),
        "H5Aclose": (herr_t, [hid_t]),
        "H5Dvlen_reclaim": (herr_t, [hid_t, hid_t, hid_t, vp]),
        "H5Eset_auto2": (herr_t, [hid_t, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    if L.H5open() < 0:
        raise OSError("H5open failed")
    L.H5Eset_auto2(0, None, None)  # failures come back as OSError, not as printed stacks
    _lib = L
    return L


def _global(name: str) -> int:
    """A predefined type or class id (the library's ``<name>_g`` variable)."""
    return hid_t.in_dll(lib(), name + "_g").value


def _check(value: int, what: str) -> int:
    if value < 0:
        raise OSError(f"HDF5 call {what} failed")
    return value


_NATIVE = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE", np.dtype(np.float32): "H5T_NATIVE_FLOAT",
           np.dtype(np.int64): "H5T_NATIVE_LLONG", np.dtype(np.int32): "H5T_NATIVE_INT"}


class Dataset:
    """A dataset read back: ``data`` (numpy array) and ``attrs`` (dict)."""

    def __init__(self, data: np.ndarray, attrs: dict) -> None:
        self.data = data
        self.attrs = attrs

    def __getitem__(self, key):
        return self.data[key]


class File:
    """An open HDF5 file (write, or read back with ``read``)."""

    def __init__(self, fid: int, names: list[str]) -> None:
        self.fid = fid
        self.names = names

    # ----------------------------------------------------------- writing
    def _attr(self, obj: int, name: str, value) -> None:
        L = lib()
        if isinstance(value, str):
            t = _check(L.H5Tcopy(_global("H5T_C_S1")), "H5Tcopy")
            L.H5Tset_size(t, H5T_VARIABLE)
            L.H5Tset_cset(t, H5T_CSET_UTF8)
            buf = ctypes.c_char_p(value.encode("utf-8"))
            ptr, close_t = ctypes.byref(buf), True
        else:
            arr = np.asarray(value)
            if arr.dtype.kind in "iub":
                arr = arr.astype(np.int64)
            elif arr.dtype.kind == "f":
                arr = arr.astype(np.float64)
            else:
                raise TypeError(f"attribute {name}: unsupported type {arr.dtype}")
            t = _global(_NATIVE[arr.dtype])
            arr = np.ascontiguousarray(arr)
            ptr, close_t = arr.ctypes.data_as(ctypes.c_void_p), False
        sp = _check(L.H5Screate(H5S_SCALAR), "H5Screate")
        a = _check(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        _check(L.H5Awrite(a, t, ptr), "H5Awrite")
        L.H5Aclose(a)
        L.H5Sclose(sp)
        if close_t:
            L.H5Tclose(t)

    def create_dataset(self, name: str, data, attrs: dict | None = None) -> None:
        """``f[name] = data`` and ``f[name].attrs[k] = v`` for the items of attrs."""
        L = lib()
        arr = np.ascontiguousarray(data)
        if arr.dtype not in _NATIVE:
            arr = np.ascontiguousarray(arr, np.float64)
        dims = (hsize_t * max(arr.ndim, 1))(*(arr.shape or (1,)))
        sp = (_check(L.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple") if arr.ndim
              else _check(L.H5Screate(H5S_SCALAR), "H5Screate"))
        t = _global(_NATIVE[arr.dtype])
        d = _check(L.H5Dcreate2(self.fid, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT,
                                H5P_DEFAULT), f"H5Dcreate2({name})")
        if arr.size:
            _check(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                              arr.ctypes.data_as(ctypes.c_void_p)), f"H5Dwrite({name})")
        for k, v in (attrs or {}).items():
            self._attr(d, k, v)
        L.H5Dclose(d)
        L.H5Sclose(sp)
        self.names.append(name)

    def set_attrs(self, name: str, attrs: dict) -> None:
        """Attributes added to an existing dataset."""
        L = lib()
        d = _check(L.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT), f"H5Dopen2({name})")
        for k, v in attrs.items():
            self._attr(d, k, v)
        L.H5Dclose(d)

    # ----------------------------------------------------------- reading
    def _read_attr(self, obj: int, name: str):
        L = lib()
        a = _check(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), f"H5Aopen({name})")
        t = L.H5Aget_type(a)
        cls = L.H5Tget_class(t)
        try:
            if cls == H5T_STRING:
                if L.H5Tis_variable_str(t) > 0:
                    mt = L.H5Tcopy(_global("H5T_C_S1"))
                    L.H5Tset_size(mt, H5T_VARIABLE)
                    L.H5Tset_cset(mt, H5T_CSET_UTF8)
                    buf = ctypes.c_char_p()
                    _check(L.H5Aread(a, mt, ctypes.byref(buf)), "H5Aread")
                    value = buf.value.decode("utf-8")
                    sp = L.H5Screate(H5S_SCALAR)
                    L.H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, ctypes.byref(buf))
                    L.H5Sclose(sp)
                    L.H5Tclose(mt)
                    return value
                n = L.H5Tget_size(t)
                buf = ctypes.create_string_buffer(n)
                _check(L.H5Aread(a, t, buf), "H5Aread")
                return buf.raw.rstrip(b"\0").decode("utf-8")
            out = np.zeros((), np.float64 if cls == H5T_FLOAT else np.int64)
            mt = _global(_NATIVE[out.dtype])
            _check(L.H5Aread(a, mt, out.ctypes.data_as(ctypes.c_void_p)), "H5Aread")
            return out[()]
        finally:
            L.H5Tclose(t)
            L.H5Aclose(a)

    def read(self, name: str, attr_names=()) -> Dataset:
        """The dataset `name` as a numpy array (float64 for floating data,
        int64 for integers) and the named attributes."""
        L = lib()
        d = _check(L.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT), f"H5Dopen2({name})")
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        t = L.H5Dget_type(d)
        kind = np.float64 if L.H5Tget_class(t) == H5T_FLOAT else np.int64
        L.H5Tclose(t)
        out = np.zeros(tuple(dims[i] for i in range(nd)), kind)
        if out.size:
            _check(L.H5Dread(d, _global(_NATIVE[out.dtype]), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                             out.ctypes.data_as(ctypes.c_void_p)), f"H5Dread({name})")
        attrs = {k: self._read_attr(d, k) for k in attr_names}
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return Dataset(out, attrs)

    def close(self) -> None:
        if self.fid >= 0:
            lib().H5Fclose(self.fid)
            self.fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def create(filename: str, overwrite: bool, memmap: bool = False) -> File:
    """A new file: truncating (overwrite) or failing if it exists (h5py's
    modes "w" and "x"); with memmap, an in-memory file never written to disk
    (the core driver without backing store)."""
    L = lib()
    fapl = H5P_DEFAULT
    if memmap:
        fapl = _check(L.H5Pcreate(_global("H5P_CLS_FILE_ACCESS_ID")), "H5Pcreate")
        _check(L.H5Pset_fapl_core(fapl, 1 << 20, 0), "H5Pset_fapl_core")
    if not overwrite and os.path.exists(filename) and not memmap:
        raise FileExistsError(f"Unable to create file (file exists): {filename}")
    flags = H5F_ACC_TRUNC if overwrite else H5F_ACC_EXCL
    fid = L.H5Fcreate(filename.encode(), flags, H5P_DEFAULT, fapl)
    if fapl != H5P_DEFAULT:
        L.H5Pclose(fapl)
    _check(fid, f"H5Fcreate({filename})")
    return File(fid, [])


def open_file(filename: str) -> File:
    """An existing file, read-only."""
    fid = _check(lib().H5Fopen(filename.encode(), H5F_ACC_RDONLY, H5P_DEFAULT),
                 f"H5Fopen({filename})")
    return File(fid, [])


def now_string() -> str:
    """The creation time as str(astropy.time.Time.now()) shows it (ISO, UTC)."""
    return datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M:%S.%f")
