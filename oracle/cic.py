"""CIC mass assignment and readout restated from py/pm.cpp (test oracle only).

paint()   follows pm_assign_cic_density, py/pm.cpp:57-139, plus clear_density
          (pm.cpp:414-426): delta starts at -1 and every particle adds
          fac*wx*wy*wz with fac = pm_factor^3 (pm.cpp:78).
readout() follows force_at_particle_locations, py/pm.cpp:142-220.
wrap()    follows periodic_wrapup_p, py/util.h:14-24.

The cell index is the C (int) cast -- truncation toward zero, not floor
(pm.cpp:95-97) -- with an index equal to nc reset to 0 (pm.cpp:109-111).
``wdtype`` selects the precision of x*dx_inv and the weights: float32
reproduces the reference's Float arithmetic, float64 is the exact-arithmetic
oracle.  Sums are always accumulated in float64.
"""
import numpy as np


def wrap(x, boxsize, dtype=np.float32):
    """py/util.h:14-24: one +L for x<0, then one -L for x>=L, per axis."""
    x = np.array(x, dtype=dtype, copy=True)
    L = dtype(boxsize)
    x = np.where(x < 0, x + L, x)
    x = np.where(x >= L, x - L, x)
    return x


def _cic_index_weights(x, nc, boxsize, wdtype):
    x = np.asarray(x)
    if wdtype is np.float32:
        # pm.cpp:26,77 -- Float dx_inv = nc/boxsize, boxsize held as a Float
        dx_inv = np.float32(np.float32(nc) / np.float32(boxsize))
    else:
        dx_inv = nc / boxsize
    X = x.astype(wdtype) * wdtype(dx_inv)
    i0 = np.trunc(X).astype(np.int64)          # (int) cast, pm.cpp:95-97
    w1 = (X - i0.astype(wdtype)).astype(wdtype)
    w0 = (wdtype(1) - w1).astype(wdtype)
    i0 = np.where(i0 >= nc, 0, i0)             # pm.cpp:109-111
    i1 = i0 + 1
    i1 = np.where(i1 >= nc, i1 - nc, i1)       # pm.cpp:113-115
    return i0, i1, w0, w1


def paint(x, nc, boxsize, pm_factor=1.0, wdtype=np.float64):
    """Overdensity delta on an (nc,nc,nc) float64 grid (unpadded).

    x: (N,3) positions.  Returns delta with the -1 offset already applied.
    """
    ncz = nc
    i0, i1, w0, w1 = _cic_index_weights(x, nc, boxsize, wdtype)
    fac = wdtype(pm_factor ** 3)
    rho = np.zeros(nc * nc * ncz, dtype=np.float64)
    for cx in (0, 1):
        ix = i0[:, 0] if cx == 0 else i1[:, 0]
        wx = w0[:, 0] if cx == 0 else w1[:, 0]
        for cy in (0, 1):
            iy = i0[:, 1] if cy == 0 else i1[:, 1]
            wy = w0[:, 1] if cy == 0 else w1[:, 1]
            for cz in (0, 1):
                iz = i0[:, 2] if cz == 0 else i1[:, 2]
                wz = w0[:, 2] if cz == 0 else w1[:, 2]
                # pm.cpp:122-132: fac*wx*wy*wz evaluated left to right
                w = ((fac * wx).astype(wdtype) * wy).astype(wdtype) * wz
                idx = (ix * nc + iy) * ncz + iz
                rho += np.bincount(idx, weights=w.astype(np.float64),
                                   minlength=rho.size)
    return rho.reshape(nc, nc, nc) - 1.0


def readout(x, meshes, boxsize, wdtype=np.float64):
    """CIC interpolation of each (nc,nc,nc) mesh in ``meshes`` at x.

    Returns an (N, len(meshes)) float64 array; column a is
    force_at_particle_locations(axis=a) of pm.cpp:142-220.
    """
    nc = meshes[0].shape[0]
    i0, i1, w0, w1 = _cic_index_weights(x, nc, boxsize, wdtype)
    out = np.zeros((len(x), len(meshes)), dtype=np.float64)
    for cx in (0, 1):
        ix = i0[:, 0] if cx == 0 else i1[:, 0]
        wx = (w0[:, 0] if cx == 0 else w1[:, 0]).astype(np.float64)
        for cy in (0, 1):
            iy = i0[:, 1] if cy == 0 else i1[:, 1]
            wy = (w0[:, 1] if cy == 0 else w1[:, 1]).astype(np.float64)
            for cz in (0, 1):
                iz = i0[:, 2] if cz == 0 else i1[:, 2]
                wz = (w0[:, 2] if cz == 0 else w1[:, 2]).astype(np.float64)
                w = wx * wy * wz
                for a, m in enumerate(meshes):
                    out[:, a] += m[ix, iy, iz] * w
    return out
