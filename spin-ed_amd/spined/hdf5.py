"""Minimal HDF5 access through libhdf5's C API (ctypes) for SpinED's output
layout (src/SpinED.hs:267-368, README.md:55-81):

  /basis/representatives        uint64 LE, (D,)
  /hamiltonian/eigenvalues      real,      (k,)
  /hamiltonian/eigenvectors     real/complex, (k, D): row i is eigenvector i
                                (the bytes of the column-major D x k block, as
                                the KnownDataset (Block a) instance writes it,
                                SpinED.hs:362-368); read_eigenvectors also
                                accepts the (D, k) shape shown in README.md:70-73
  /hamiltonian/residuals        real,      (k,)
  /observables/<name>           complex double, (k,)
  /_workspace                   (created, unused -- SpinED.hs:276-277)

Complex values use the {r, i} compound type.  Files are opened in
read-write-append mode like H5.withFile' ... WriteAppend (SpinED.hs:284-287),
and an existing dataset is deleted (with a warning) before it is rewritten
(writeDataset, SpinED.hs:354-360).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import logging
import os
from pathlib import Path

import numpy as np

log = logging.getLogger("spined")

_CANDIDATES = [os.environ.get("SPINED_HDF5_LIBRARY", ""), "/opt/conda/lib/libhdf5.so",
               ctypes.util.find_library("hdf5") or ""]

_h5 = None
hid_t = C.c_int64
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5T_COMPOUND = 6


class HDF5Unavailable(RuntimeError):
    pass


def lib():
    global _h5
    if _h5 is not None:
        return _h5
    for c in _CANDIDATES:
        if c and Path(c).exists() or (c and not c.startswith("/")):
            try:
                L = C.CDLL(c)
                break
            except OSError:
                continue
    else:
        raise HDF5Unavailable("libhdf5 not found (set SPINED_HDF5_LIBRARY)")
    L.H5open()
    sig = {
        "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
        "H5Fclose": (C.c_int, [hid_t]),
        "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Gclose": (C.c_int, [hid_t]),
        "H5Lexists": (C.c_int, [hid_t, C.c_char_p, hid_t]),
        "H5Ldelete": (C.c_int, [hid_t, C.c_char_p, hid_t]),
        "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "H5Sclose": (C.c_int, [hid_t]),
        "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
        "H5Dwrite": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dread": (C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dclose": (C.c_int, [hid_t]),
        "H5Tcreate": (hid_t, [C.c_int, C.c_size_t]),
        "H5Tinsert": (C.c_int, [hid_t, C.c_char_p, C.c_size_t, hid_t]),
        "H5Tclose": (C.c_int, [hid_t]),
        "H5Tget_class": (C.c_int, [hid_t]),
        "H5Tget_size": (C.c_size_t, [hid_t]),
        "H5Eset_auto2": (C.c_int, [hid_t, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.H5Eset_auto2(0, None, None)  # H5.disableDiagOutput (app/Main.hs:16)
    _h5 = L
    return L


def available() -> bool:
    try:
        lib()
        return True
    except (HDF5Unavailable, OSError):
        return False


def _type(name):
    return hid_t.in_dll(lib(), name).value


def _mem_and_file_type(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.uint64:
        return _type("H5T_NATIVE_ULLONG_g"), _type("H5T_STD_U64LE_g"), None
    if dtype == np.int64:
        return _type("H5T_NATIVE_LLONG_g"), _type("H5T_STD_I64LE_g"), None
    if dtype == np.float64:
        return _type("H5T_NATIVE_DOUBLE_g"), _type("H5T_IEEE_F64LE_g"), None
    if dtype == np.float32:
        return _type("H5T_NATIVE_FLOAT_g"), _type("H5T_IEEE_F32LE_g"), None
    if dtype in (np.complex128, np.complex64):
        base = "H5T_NATIVE_DOUBLE_g" if dtype == np.complex128 else "H5T_NATIVE_FLOAT_g"
        sz = dtype.itemsize
        t = lib().H5Tcreate(H5T_COMPOUND, sz)
        lib().H5Tinsert(t, b"r", 0, _type(base))
        lib().H5Tinsert(t, b"i", sz // 2, _type(base))
        return t, t, t
    raise TypeError(f"unsupported dataset type {dtype}")


def _check(v, what):
    if v < 0:
        raise OSError(f"HDF5 call failed: {what}")
    return v


class File:
    """H5.withFile' filename WriteAppend (SpinED.hs:284-287)."""

    def __init__(self, path, mode="a"):
        L = lib()
        p = str(path).encode()
        if mode == "r":
            self.fid = _check(L.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT), "open")
        elif mode == "w" or not Path(path).exists():
            Path(path).parent.mkdir(parents=True, exist_ok=True)
            self.fid = _check(L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create")
        else:
            self.fid = _check(L.H5Fopen(p, H5F_ACC_RDWR, H5P_DEFAULT), "open rw")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.fid is not None:
            lib().H5Fclose(self.fid)
            self.fid = None

    def exists(self, path: str) -> bool:
        L = lib()
        parts = [q for q in path.strip("/").split("/") if q]
        cur = ""
        for q in parts:
            cur += "/" + q
            if L.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def create_group(self, path: str) -> None:
        """createGroupWhenNeeded (SpinED.hs:279-282), each missing level."""
        cur = ""
        for q in [q for q in path.strip("/").split("/") if q]:
            cur += "/" + q
            if not self.exists(cur):
                g = _check(lib().H5Gcreate2(self.fid, cur.encode(), H5P_DEFAULT, H5P_DEFAULT,
                                            H5P_DEFAULT), "create group")
                lib().H5Gclose(g)

    def delete(self, path: str) -> None:
        """Unlink a dataset or group when present."""
        if self.exists(path):
            lib().H5Ldelete(self.fid, path.encode(), H5P_DEFAULT)

    def write(self, path: str, data: np.ndarray, warn: bool = True) -> None:
        """writeDataset (SpinED.hs:354-360): overwrite with a warning (warn=False:
        silently, for the solver's periodic checkpoints)."""
        L = lib()
        data = np.ascontiguousarray(data)
        parent = path.rsplit("/", 1)[0]
        if parent:
            self.create_group(parent)
        if self.exists(path):
            if warn:
                log.warning("Overwriting %s...", path)
            L.H5Ldelete(self.fid, path.encode(), H5P_DEFAULT)
        mt, ft, owned = _mem_and_file_type(data.dtype)
        dims = (C.c_uint64 * max(data.ndim, 1))(*(data.shape or (1,)))
        sp = _check(L.H5Screate_simple(max(data.ndim, 1), dims, None), "dataspace")
        try:
            ds = _check(L.H5Dcreate2(self.fid, path.encode(), ft, sp, H5P_DEFAULT, H5P_DEFAULT,
                                     H5P_DEFAULT), f"create {path}")
            try:
                _check(L.H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                  data.ctypes.data if data.size else None), f"write {path}")
            finally:
                L.H5Dclose(ds)
        finally:
            L.H5Sclose(sp)
            if owned is not None:
                L.H5Tclose(owned)

    def read(self, path: str, dtype) -> np.ndarray:
        L = lib()
        ds = _check(L.H5Dopen2(self.fid, path.encode(), H5P_DEFAULT), f"open {path}")
        try:
            sp = L.H5Dget_space(ds)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_uint64 * max(nd, 1))()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            L.H5Sclose(sp)
            shape = tuple(int(dims[i]) for i in range(nd))
            out = np.empty(shape, dtype=dtype)
            mt, _, owned = _mem_and_file_type(dtype)
            try:
                if out.size:
                    _check(L.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data),
                           f"read {path}")
            finally:
                if owned is not None:
                    L.H5Tclose(owned)
            return out
        finally:
            L.H5Dclose(ds)


def read_eigenvectors(f: File, dim: int, dtype=np.float64) -> np.ndarray:
    """Eigenvectors as a (D, k) array from either on-disk shape: (k, D) as the
    reference code writes it, or (D, k) as README.md shows it."""
    a = f.read("/hamiltonian/eigenvectors", dtype)
    if a.ndim == 1:
        return a.reshape(-1, 1)
    if a.shape[1] == dim and a.shape[0] != dim:
        return np.ascontiguousarray(a.T)
    if a.shape[0] == dim and a.shape[1] != dim:
        return a
    return np.ascontiguousarray(a.T)  # square: (k, D) convention
