"""Minimal HDF5 reader over the HDF5 C library (ctypes), for the output
tests: h5py is not installed.  Returns numpy arrays of whole datasets."""
import ctypes
import os

import numpy as np

HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
_lib = None


def _h5():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))
        hid = ctypes.c_int64
        lib.H5open.restype = ctypes.c_int
        lib.H5open()
        lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
        lib.H5Fopen.restype = hid
        lib.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
        lib.H5Dopen2.restype = hid
        lib.H5Dget_space.argtypes = [hid]
        lib.H5Dget_space.restype = hid
        lib.H5Sget_simple_extent_ndims.argtypes = [hid]
        lib.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(ctypes.c_uint64),
                                                  ctypes.c_void_p]
        lib.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
            getattr(lib, f).argtypes = [hid]
        _lib = lib
    return _lib


def available():
    try:
        _h5()
        return True
    except OSError:
        return False


def read(filename, name):
    lib = _h5()
    hid = ctypes.c_int64
    types = {np.float32: "H5T_NATIVE_FLOAT_g", np.float64: "H5T_NATIVE_DOUBLE_g",
             np.uint64: "H5T_NATIVE_UINT64_g", np.int32: "H5T_NATIVE_INT_g"}
    f = lib.H5Fopen(filename.encode(), 0, 0)
    if f < 0:
        raise IOError(filename)
    try:
        d = lib.H5Dopen2(f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        s = lib.H5Dget_space(d)
        nd = lib.H5Sget_simple_extent_ndims(s)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        lib.H5Sget_simple_extent_dims(s, dims, None)
        shape = tuple(dims[i] for i in range(nd))
        lib.H5Sclose(s)
        if name.split("/")[-1] in ("id",):
            dtype = np.uint64
        elif name.split("/")[-1] == "nc":
            dtype = np.int32
        elif name.startswith("parameters"):
            dtype = np.float64
        else:
            dtype = np.float32
        out = np.empty(shape, dtype=dtype)
        mem = hid.in_dll(lib, types[dtype]).value
        if lib.H5Dread(d, mem, 0, 0, 0, out.ctypes.data) < 0:
            raise IOError(name)
        lib.H5Dclose(d)
        return out
    finally:
        lib.H5Fclose(f)
