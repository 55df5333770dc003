"""Binned power spectrum of a real mesh (test oracle only).

No reference counterpart (fs2 has no estimator).  Definition used by the
build: with delta_k = rfftn(delta) (unnormalised), P(k) = V |delta_k|^2 / Ng^2
per mode, |k| = 2 pi / L * sqrt(kx^2+ky^2+kz^2) with integer kx, ky in
(-n/2, n/2] and kz in [0, n/2].  Bin b collects modes with
b <= |k_int| / dk_bin < b+1 (dk_bin in units of the fundamental mode); the
kz = 0 and kz = n/2 planes count once, other kz twice (the conjugate half);
k = 0 is excluded.  Sums are FP64.  Returns (k_mean, P_mean, n_modes).
"""
import numpy as np


def power_spectrum(delta, boxsize, nbins=None, dk_bin=1.0):
    n = delta.shape[0]
    dk = np.fft.rfftn(delta)
    i = np.arange(n)
    kxy = np.where(i <= n // 2, i, i - n).astype(np.float64)
    kz = np.arange(n // 2 + 1).astype(np.float64)
    kk = np.sqrt(kxy[:, None, None]**2 + kxy[None, :, None]**2 + kz[None, None, :]**2)
    w = np.full(kk.shape, 2.0)
    w[:, :, 0] = 1.0
    w[:, :, n // 2] = 1.0
    if nbins is None:
        nbins = int(n // 2 / dk_bin)
    b = np.floor(kk / dk_bin).astype(np.int64)
    p = (np.abs(dk)**2) * boxsize**3 / float(n**3)**2
    sel = (kk > 0) & (b < nbins)
    cnt = np.bincount(b[sel], weights=w[sel], minlength=nbins)
    ks = np.bincount(b[sel], weights=(w * kk)[sel], minlength=nbins)
    ps = np.bincount(b[sel], weights=(w * p)[sel], minlength=nbins)
    kf = 2.0 * np.pi / boxsize
    with np.errstate(invalid='ignore', divide='ignore'):
        return ks / cnt * kf, ps / cnt, cnt
