"""The reference's CPU step, timed as the CPU baseline (bench only).

Measurement infrastructure: bench.py's CPU-baseline leg drives this; the
product path never imports it.  oracle/cpu_step.cpp holds the particle and
mesh loops of py/pm.cpp and py/cola.cpp on the 56-byte AoS records, with
OpenMP at the reference's `#pragma omp` sites; the transforms are scipy's
pocketfft in FP32 (the reference calls FFTW, absent here and on the GPU
host) with as many worker threads as the loops.

One step follows pm.force (py/fs/pm.py:17-31: send_positions, density,
force) then cola.kick and cola.drift (script/cola.py:37-44):
    wrap, clear to -1, CIC paint, r2c, copy into delta_k,
    3 x (Green x gradient, c2r, readout of one axis), kick, drift.
"""
import ctypes
import os
import subprocess
import time

import numpy as np
from scipy import fft as sfft

from . import cola as ocola
from . import cosmology as ocosmo

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cpu_step.cpp')
LIB = os.path.join(HERE, '_build', 'libcpu_step.so')

RECORD = np.dtype([('x', '<f4', 3), ('v', '<f4', 3), ('dx1', '<f4', 3), ('dx2', '<f4', 3),
                   ('id', '<u8')])
assert RECORD.itemsize == 56

_lib = None


def build():
    """Compiles oracle/cpu_step.cpp (g++ -O3 -fopenmp)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + '.tmp'
    subprocess.run(['g++', '-O3', '-std=c++17', '-fPIC', '-shared', '-fopenmp',
                    '-ffp-contract=off', SRC, '-o', tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        lib = ctypes.CDLL(LIB)
        vp, i32, i64, f32, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_float, ctypes.c_double)
        sig = {
            'cpu_step_threads': ([i32], i32),
            'cpu_step_record_size': ([], i32),
            'cpu_clear_density': ([vp, i32], None),
            'cpu_wrap': ([vp, i64, f32], None),
            'cpu_paint': ([vp, i64, vp, i32, f32, f64], None),
            'cpu_copy_delta_k': ([vp, vp, i32], None),
            'cpu_force_k': ([vp, vp, i32, f32, i32], None),
            'cpu_readout': ([vp, i64, vp, i32, f32, i32, vp], None),
            'cpu_kick': ([vp, vp, i64, f64, f32, f32, f32], None),
            'cpu_drift': ([vp, i64, f64, f32, f32], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        assert lib.cpu_step_record_size() == RECORD.itemsize
        _lib = lib
    return _lib


def host_info():
    """nproc, the cores this process may use, and the CPU model."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {'nproc': os.cpu_count(), 'usable_cores': usable, 'cpu_model': model}


def records(x, v, dx1, dx2, ids):
    """(N,3) float32 fields and u64 ids -> one array of 56-byte records."""
    r = np.empty(len(x), dtype=RECORD)
    r['x'], r['v'], r['dx1'], r['dx2'], r['id'] = x, v, dx1, dx2, ids
    return r


class CpuCola:
    """The reference's PM + COLA state and step on the host, FP32."""

    def __init__(self, recs, nc, boxsize, pm_factor, omega_m, a, threads, a_v=None):
        self.lib = _load()
        self.p = np.ascontiguousarray(recs)
        self.np = len(self.p)
        self.nc, self.L, self.pm_factor = int(nc), float(boxsize), float(pm_factor)
        self.cosmo = ocosmo.Cosmology(omega_m)
        self.a_x = float(a)
        self.a_v = float(a if a_v is None else a_v)
        self.threads = self.lib.cpu_step_threads(int(threads))
        n = self.nc
        self.mesh = np.empty((n, n, n), dtype=np.float32)
        self.delta_k = np.empty((n, n, n // 2 + 1), dtype=np.complex64)
        self.fk = np.empty_like(self.delta_k)
        self.force = np.empty((self.np, 3), dtype=np.float32)

    @staticmethod
    def _ptr(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    def step(self, a_vel, a_pos):
        """One COLA step; returns seconds per phase."""
        lib, P, n, L = self.lib, self._ptr(self.p), self.nc, self.L
        t = {}
        t0 = time.perf_counter()
        lib.cpu_wrap(P, self.np, L)
        lib.cpu_clear_density(self._ptr(self.mesh), n)
        lib.cpu_paint(P, self.np, self._ptr(self.mesh), n, L, self.pm_factor)
        t['paint'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        fk = sfft.rfftn(self.mesh, workers=self.threads)
        t['fft'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        lib.cpu_copy_delta_k(self._ptr(fk), self._ptr(self.delta_k), n)
        t['green'] = time.perf_counter() - t0
        del fk
        t['readout'] = 0.0
        for axis in range(3):
            t0 = time.perf_counter()
            lib.cpu_force_k(self._ptr(self.delta_k), self._ptr(self.fk), n, L, axis)
            t['green'] += time.perf_counter() - t0
            t0 = time.perf_counter()
            mesh = sfft.irfftn(self.fk, s=(n, n, n), norm='forward', workers=self.threads)
            t['fft'] += time.perf_counter() - t0
            t0 = time.perf_counter()
            lib.cpu_readout(P, self.np, self._ptr(mesh), n, L, axis, self._ptr(self.force))
            t['readout'] += time.perf_counter() - t0
            del mesh
        t0 = time.perf_counter()
        kf, q1, q2 = ocola.kick_coefficients(self.cosmo, self.a_v, self.a_x, a_vel)
        lib.cpu_kick(P, self._ptr(self.force), self.np, self.cosmo.om, kf, q1, q2)
        self.a_v = a_vel
        dt, da1, da2 = ocola.drift_coefficients(self.cosmo, self.a_x, self.a_v, a_pos)
        lib.cpu_drift(P, self.np, dt, da1, da2)
        self.a_x = a_pos
        t['kick_drift'] = time.perf_counter() - t0
        return t

    def positions(self):
        return self.p['x'].copy()
