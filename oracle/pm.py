"""PM force restated from py/pm.cpp:306-508 in FP64 (test oracle only).

Conventions (py/fft.cpp:80-100): the forward r2c is unnormalised with sign -1
(numpy.fft.rfftn), the inverse c2r is unnormalised (numpy irfftn times n^3).
numpy's irfftn runs the complex inverse over the two outer axes first and
the c2r over the last axis last, discarding the imaginary parts of the kz=0
and kz=n/2 planes -- the same result FFTW's c2r gives on the non-Hermitian
planes that multiplying by an odd function of k produces.

Green x gradient (pm.cpp:460-504):
    f1 = -1/n^3 / (2 pi / L),  f2 = f1 * k_axis / |k|^2 (integer k),
    fk = i f2 delta_k,  k = 0 -> 0,
with indices <= n/2 mapped to +index (Nyquist is +n/2, pm.cpp:483, 489) and
kz running over 0..n/2.
"""
import os

import numpy as np
from scipy import fft as sfft

from . import cic, cfast


def _workers():
    """Threads for scipy's pocketfft (FP64, the routine numpy.fft wraps)."""
    return max(1, min(16, int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))))


def int_kvectors_pm(n):
    """Integer wave numbers along x/y (Nyquist -> +n/2) and z (0..n/2)."""
    k = np.arange(n)
    kxy = np.where(k <= n // 2, k, k - n).astype(np.float64)
    kz = np.arange(n // 2 + 1).astype(np.float64)
    return kxy, kz


def density(x, nc, boxsize, pm_factor):
    return cic.paint(x, nc, boxsize, pm_factor)


def force_meshes(delta, boxsize, progress=None):
    """Three real force meshes f_a = irfft(i f2 k_a delta_k) (pm.cpp:460-508).
    progress(msg), if given, is called after each transform."""
    n = delta.shape[0]
    dk = sfft.rfftn(delta, workers=_workers())
    if progress:
        progress("r2c done")
    kxy, kz = int_kvectors_pm(n)
    KX = kxy[:, None, None]
    KY = kxy[None, :, None]
    KZ = kz[None, None, :]
    k2 = KX**2 + KY**2 + KZ**2
    k2[0, 0, 0] = 1.0
    f1 = -1.0 / n**3 / (2.0 * np.pi / boxsize)
    out = []
    for K in (KX, KY, KZ):
        f2 = f1 / k2 * K
        fk = 1j * f2 * dk
        fk[0, 0, 0] = 0.0
        out.append(sfft.irfftn(fk, s=(n, n, n), axes=(0, 1, 2), workers=_workers()) * n**3)
        del fk
        if progress:
            progress(f"c2r {len(out)} of 3 done")
    return out


def force_meshes_fp32(delta, boxsize, workers=1):
    """The same three meshes in FP32 with scipy.fft's multithreaded
    pocketfft; the CPU-baseline port of pm.cpp:431-508 (reference Float)."""
    from scipy import fft as sfft
    n = delta.shape[0]
    dk = sfft.rfftn(delta.astype(np.float32, copy=False), workers=workers)
    kxy, kz = int_kvectors_pm(n)
    f1 = np.float32(-1.0 / n**3 / (2.0 * np.pi / boxsize))
    kx = kxy.astype(np.float32)[:, None, None]
    ky = kxy.astype(np.float32)[None, :, None]
    kzz = kz.astype(np.float32)[None, None, :]
    out = []
    for axis in range(3):
        fk = np.empty_like(dk)
        for ix in range(n):   # slab by slab keeps temporaries small
            k2 = kx[ix] ** 2 + ky[0] ** 2 + kzz[0] ** 2
            if ix == 0:
                k2[0, 0] = 1.0
            K = (kx[ix] + 0 * k2, ky[0] + 0 * k2, kzz[0] + 0 * k2)[axis]
            f2 = (f1 / k2 * K).astype(np.float32)
            fk[ix] = 1j * f2 * dk[ix]
            if ix == 0:
                fk[0, 0, 0] = 0
        out.append(sfft.irfftn(fk, s=(n, n, n), workers=workers) * np.float32(n**3))
        del fk
    return out


def force(x, nc, boxsize, pm_factor):
    """Forces at (wrapped) positions x: paint -> r2c -> Green -> c2r -> readout."""
    delta = density(x, nc, boxsize, pm_factor)
    meshes = force_meshes(delta, boxsize)
    return cic.readout(x, meshes, boxsize)


def force_large(x, nc, boxsize, pm_factor, progress=None, delta=None):
    """force() for large meshes: the C restatement of the CIC paint and
    readout (oracle/cic_fp64.c), threaded FP64 FFTs, and one force mesh
    alive at a time.  delta, if given, is the density of x (from
    cfast.paint).  Also returns the FP64 sum of delta."""
    n = nc
    if delta is None:
        delta = cfast.paint(x, nc, boxsize, pm_factor)
        if progress:
            progress("paint done")
    total = float(delta.sum())
    dk = sfft.rfftn(delta, workers=_workers())
    del delta
    if progress:
        progress("r2c done")
    kxy, kz = int_kvectors_pm(n)
    f1 = -1.0 / n**3 / (2.0 * np.pi / boxsize)
    out = np.zeros((len(x), 3), dtype=np.float64)
    for axis in range(3):
        fk = np.empty_like(dk)
        for ix in range(n):   # slab by slab keeps the temporaries small
            KX = kxy[ix]
            KY = kxy[:, None]
            KZ = kz[None, :]
            k2 = KX * KX + KY * KY + KZ * KZ
            if ix == 0:
                k2[0, 0] = 1.0
            K = (KX + 0 * k2, KY + 0 * k2, KZ + 0 * k2)[axis]
            fk[ix] = 1j * (f1 / k2 * K) * dk[ix]
            if ix == 0:
                fk[0, 0, 0] = 0.0
        mesh = sfft.irfftn(fk, s=(n, n, n), axes=(0, 1, 2), workers=_workers())
        del fk
        mesh *= n**3
        ms = [None, None, None]
        ms[axis] = mesh
        out[:, axis] = cfast.readout(x, ms, boxsize)[:, axis]
        del mesh, ms
        if progress:
            progress(f"axis {axis} done")
    return out, total
