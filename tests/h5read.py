"""Read-back of HDF5 files in tests through the HDF5 1.10 C library (ctypes);
no Python HDF5 package is installed here."""
import ctypes as C
import os

import numpy as np

_L = None
H5F_ACC_RDONLY = 0
H5P_DEFAULT = 0
H5S_ALL = 0
H5T_INTEGER, H5T_FLOAT = 0, 1


def lib():
    global _L
    if _L is None:
        for name in (os.environ.get("EDUS_HDF5_LIB", ""), "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"):
            if not name:
                continue
            try:
                _L = C.CDLL(name)
                break
            except OSError:
                continue
        if _L is None:
            raise OSError("no HDF5 1.10 library")
        L = _L
        hid = C.c_int64
        L.H5open.restype = C.c_int
        L.H5Fopen.restype = hid
        L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
        L.H5Fclose.argtypes = [hid]
        L.H5Dopen2.restype = hid
        L.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
        L.H5Dclose.argtypes = [hid]
        L.H5Dget_space.restype = hid
        L.H5Dget_space.argtypes = [hid]
        L.H5Dget_type.restype = hid
        L.H5Dget_type.argtypes = [hid]
        L.H5Tget_class.restype = C.c_int
        L.H5Tget_class.argtypes = [hid]
        L.H5Tget_size.restype = C.c_size_t
        L.H5Tget_size.argtypes = [hid]
        L.H5Tclose.argtypes = [hid]
        L.H5Sget_simple_extent_npoints.restype = C.c_int64
        L.H5Sget_simple_extent_npoints.argtypes = [hid]
        L.H5Sclose.argtypes = [hid]
        L.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
        L.H5Lexists.restype = C.c_int
        L.H5Lexists.argtypes = [hid, C.c_char_p, hid]
        L.H5open()
    return _L


class File:
    def __init__(self, path):
        self.id = lib().H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.id < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.id >= 0:
            lib().H5Fclose(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, path: str) -> bool:
        """Every component of an absolute path exists."""
        parts = [p for p in path.split("/") if p]
        for i in range(1, len(parts) + 1):
            if lib().H5Lexists(self.id, ("/" + "/".join(parts[:i])).encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def read(self, path: str) -> np.ndarray:
        """A one-dimensional dataset as float64, int32 or uint8."""
        L = lib()
        d = L.H5Dopen2(self.id, path.encode(), H5P_DEFAULT)
        if d < 0:
            raise KeyError(path)
        try:
            sp, ty = L.H5Dget_space(d), L.H5Dget_type(d)
            n = L.H5Sget_simple_extent_npoints(sp)
            cls, size = L.H5Tget_class(ty), L.H5Tget_size(ty)
            L.H5Sclose(sp)
            L.H5Tclose(ty)
            if cls == H5T_FLOAT and size == 8:
                out, mem = np.empty(n, dtype=np.float64), C.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
            elif cls == H5T_INTEGER and size == 4:
                out, mem = np.empty(n, dtype=np.int32), C.c_int64.in_dll(L, "H5T_NATIVE_INT_g").value
            elif cls == H5T_INTEGER and size == 1:
                out, mem = np.empty(n, dtype=np.uint8), C.c_int64.in_dll(L, "H5T_NATIVE_UCHAR_g").value
            else:
                raise TypeError(f"{path}: class {cls}, size {size}")
            if n and L.H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data_as(C.c_void_p)) < 0:
                raise OSError(f"H5Dread failed for {path}")
            return out
        finally:
            L.H5Dclose(d)
