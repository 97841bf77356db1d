"""Minimal reader of HDF5 files through the HDF5 C library (ctypes), for
checking the integral file the library writes (there is no h5py here)."""
import ctypes as C
import os

import numpy as np

H5LIB = "/opt/conda/lib/libhdf5.so"
_h5 = None


def available():
    return os.path.exists(H5LIB)


def lib():
    global _h5
    if _h5 is None:
        _h5 = C.CDLL(H5LIB)
        _h5.H5open()
        for f in ("H5Fopen", "H5Dopen2", "H5Gopen2", "H5Dget_space", "H5Aopen", "H5Dget_create_plist"):
            getattr(_h5, f).restype = C.c_int64
        _h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        _h5.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        _h5.H5Gopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        _h5.H5Aopen.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        _h5.H5Dget_space.argtypes = [C.c_int64]
        _h5.H5Dget_create_plist.argtypes = [C.c_int64]
        _h5.H5Sget_simple_extent_dims.argtypes = [C.c_int64, C.c_void_p, C.c_void_p]
        _h5.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        _h5.H5Aread.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
        _h5.H5Pget_nfilters.argtypes = [C.c_int64]
        _h5.H5Pget_layout.argtypes = [C.c_int64]
        _h5.H5Lexists.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        for f in ("H5Dclose", "H5Sclose", "H5Aclose", "H5Gclose", "H5Fclose", "H5Pclose"):
            getattr(_h5, f).argtypes = [C.c_int64]
    return _h5


class File:
    def __init__(self, path):
        self.h = lib()
        self.fid = self.h.H5Fopen(path.encode(), 0, 0)   # H5F_ACC_RDONLY
        assert self.fid >= 0, path

    def close(self):
        self.h.H5Fclose(self.fid)

    def exists(self, name):
        parts = name.strip("/").split("/")
        for i in range(1, len(parts) + 1):
            if self.h.H5Lexists(self.fid, ("/" + "/".join(parts[:i])).encode(), 0) <= 0:
                return False
        return True

    def read(self, name):
        """(array in C order, chunked?, number of filters)."""
        h = self.h
        d = h.H5Dopen2(self.fid, name.encode(), 0)
        assert d >= 0, name
        s = h.H5Dget_space(d)
        dims = (C.c_uint64 * 2)()
        rank = h.H5Sget_simple_extent_dims(s, dims, None)
        shape = tuple(int(dims[i]) for i in range(rank))
        out = np.zeros(shape)
        dt = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
        assert h.H5Dread(d, dt, 0, 0, 0, out.ctypes.data) >= 0
        p = h.H5Dget_create_plist(d)
        chunked = h.H5Pget_layout(p) == 2          # H5D_CHUNKED
        nfilt = h.H5Pget_nfilters(p)
        h.H5Pclose(p)
        h.H5Sclose(s)
        h.H5Dclose(d)
        return out, chunked, nfilt

    def status(self, group):
        h = self.h
        g = h.H5Gopen2(self.fid, group.encode(), 0)
        assert g >= 0, group
        a = h.H5Aopen(g, b"status", 0)
        assert a >= 0, group
        v = C.c_int(0)
        it = C.c_int64.in_dll(h, "H5T_NATIVE_INT_g").value
        assert h.H5Aread(a, it, C.byref(v)) >= 0
        h.H5Aclose(a)
        h.H5Gclose(g)
        return v.value

    def dataset_status(self, name):
        h = self.h
        d = h.H5Dopen2(self.fid, name.encode(), 0)
        assert d >= 0, name
        a = h.H5Aopen(d, b"status", 0)
        assert a >= 0, name
        v = C.c_int(-1)
        it = C.c_int64.in_dll(h, "H5T_NATIVE_INT_g").value
        assert h.H5Aread(a, it, C.byref(v)) >= 0
        h.H5Aclose(a)
        h.H5Dclose(d)
        return v.value
