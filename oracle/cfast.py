"""ctypes front end of oracle/cic_fp64.c (test oracle only).

paint() and readout() give the same FP64 results as oracle/cic.py's paint()
and readout() with wdtype float64 (to summation-order rounding; pinned by
tests/test_oracle_pins.py), but run in compiled C: at 134M particles the numpy
versions take minutes and tens of GB of temporaries.

The library is built by build() into oracle/_build/ (gcc -O2, no fast-math
and no contraction, so every FP64 operation rounds as written; the readout's
particle loop runs on OpenMP threads, which changes no result bit).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cic_fp64.c')
LIB = os.path.join(HERE, '_build', 'liboracle_cic.so')

_lib = None


def build():
    """Compiles oracle/cic_fp64.c into oracle/_build/liboracle_cic.so."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + '.tmp'
    subprocess.run(['gcc', '-O2', '-std=c99', '-fPIC', '-shared', '-ffp-contract=off', '-fopenmp',
                    SRC, '-o', tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        lib = ctypes.CDLL(LIB)
        d = ctypes.POINTER(ctypes.c_double)
        lib.oracle_cic_paint.argtypes = [d, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, d]
        lib.oracle_cic_paint.restype = None
        lib.oracle_cic_readout3.argtypes = [d, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                            d, d, d, d]
        lib.oracle_cic_readout3.restype = None
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def paint(x, nc, boxsize, pm_factor=1.0):
    """delta on an (nc, nc, nc) float64 grid, -1 applied (pm.cpp:57-139, 414-426)."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    rho = np.zeros((nc, nc, nc), dtype=np.float64)
    lib.oracle_cic_paint(_ptr(x), len(x), nc, float(boxsize), float(pm_factor) ** 3, _ptr(rho))
    rho -= 1.0
    return rho


def readout(x, meshes, boxsize):
    """(N, 3) float64: CIC interpolation of three (nc, nc, nc) meshes at x
    (pm.cpp:142-220); a mesh given as None reads as zero."""
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    ms = [None if m is None else np.ascontiguousarray(m, dtype=np.float64) for m in meshes]
    nc = next(m for m in ms if m is not None).shape[0]
    while len(ms) < 3:
        ms.append(None)
    out = np.zeros((len(x), 3), dtype=np.float64)
    null = ctypes.POINTER(ctypes.c_double)()
    lib.oracle_cic_readout3(_ptr(x), len(x), nc, float(boxsize),
                            *[null if m is None else _ptr(m) for m in ms], _ptr(out))
    return out
