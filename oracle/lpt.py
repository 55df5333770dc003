"""2LPT initial conditions from a white-noise field (test oracle only).

Follows py/lpt.cpp with the GSL/N-GenIC random draw replaced by a given
white-noise field w (n^3, N(0,1)); the mode amplitudes keep the reference's
normalisation E|delta_k|^2 = P(k)/V (lpt.cpp:364-366 with fac^2 fac_2pi3 =
1/L^3):
    delta_k = rfftn(w) * sqrt(P(|k|)/V) / n^{3/2}
    modes with any index == n/2, and k = 0, are zero        (lpt.cpp:328-331)
    Psi1_k  = i k / k^2 delta_k                              (lpt.cpp:374-375)
    Psi_ij  = c2r(i k_j Psi1_i,k)                             (lpt.cpp:486-516)
    div2    = P11(P22+P33) + P22 P33 - P12^2 - P13^2 - P23^2 (lpt.cpp:527-532)
    Psi2_k  = -i k / k^2 r2c(div2),  k = 0 -> 0              (lpt.cpp:546-583)
    dx1 = c2r(Psi1_k), dx2 = c2r(Psi2_k) / n^3               (lpt.cpp:132-147,191-192)
    x = q + D1 dx1 + D2 dx2, q = (i + offset) L/n            (lpt.cpp:182-199)
    v = Dv dx1 + D2v dx2 with (Dv, D2v) = (0,0) for 'cola',
        (Dv(a,D1), 0) for 'zeldovich', (Dv(a,D1), D2v(a,D1)) for '2lpt'
        (the reference passes D1 where D2v expects D2, lpt.cpp:172; kept),
        and D2 = 0 for 'zeldovich'                            (lpt.cpp:156-175)
    id = 1..n^3 in ix-major order                            (lpt.cpp:148, 206)
The lpt.cpp wave-number convention maps index n/2 to -n/2 (lpt.cpp:469-482,
558-571), which matters only for Psi2 on the Nyquist planes.
"""
import numpy as np


def kvectors_lpt(n, boxsize):
    dk = 2.0 * np.pi / boxsize
    i = np.arange(n)
    kxy = np.where(i < n // 2, i, i - n) * dk
    iz = np.arange(n // 2 + 1)
    kz = np.where(iz < n // 2, iz, iz - n) * dk
    return kxy, kz


def c2r(fk, n):
    return np.fft.irfftn(fk, s=(n, n, n), axes=(0, 1, 2)) * n**3


def delta_k_from_noise(noise, boxsize, Pfun):
    n = noise.shape[0]
    W = np.fft.rfftn(noise)
    kxy, kz = kvectors_lpt(n, boxsize)
    KX, KY, KZ = kxy[:, None, None], kxy[None, :, None], kz[None, None, :]
    k2 = KX**2 + KY**2 + KZ**2
    kmag = np.sqrt(k2)
    V = boxsize**3
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = np.sqrt(Pfun(kmag[nz]) / V) / n**1.5
    h = n // 2
    mask = np.ones_like(k2, dtype=bool)
    mask[h, :, :] = False
    mask[:, h, :] = False
    mask[:, :, h] = False
    mask[0, 0, 0] = False
    return np.where(mask, W * amp, 0.0), (KX, KY, KZ, k2)


def displacements(noise, boxsize, Pfun):
    """Returns (dx1, dx2) as (n,n,n,3) float64 fields."""
    n = noise.shape[0]
    dk, (KX, KY, KZ, k2) = delta_k_from_noise(noise, boxsize, Pfun)
    K = (KX, KY, KZ)
    k2s = np.where(k2 > 0, k2, 1.0)
    psi_k = [1j * Ki / k2s * dk for Ki in K]
    for p in psi_k:
        p[0, 0, 0] = 0.0
    # Psi_i,j = i k_j Psi_i   (lpt.cpp:486-502: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2))
    pij = {}
    for (i, j) in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
        pij[(i, j)] = c2r(1j * K[j] * psi_k[i], n)
    div2 = (pij[(0, 0)] * (pij[(1, 1)] + pij[(2, 2)]) + pij[(1, 1)] * pij[(2, 2)]
            - pij[(0, 1)]**2 - pij[(0, 2)]**2 - pij[(1, 2)]**2)
    Dk = np.fft.rfftn(div2)
    psi2_k = [-1j * Ki / k2s * Dk for Ki in K]
    for p in psi2_k:
        p[0, 0, 0] = 0.0
    dx1 = np.stack([c2r(p, n) for p in psi_k], axis=-1)
    dx2 = np.stack([c2r(p, n) for p in psi2_k], axis=-1) / n**3
    return dx1, dx2


def particles(noise, boxsize, a, Pfun, cosmo, kind='cola', offset=0.5):
    n = noise.shape[0]
    dx1, dx2 = displacements(noise, boxsize, Pfun)
    D1 = cosmo.D(a)
    D2 = 0.0 if kind == 'zeldovich' else cosmo.D2(a, D1)
    if kind == 'cola':
        Dv = D2v = 0.0
    elif kind == 'zeldovich':
        Dv, D2v = cosmo.Dv(a, D1), 0.0
    elif kind == '2lpt':
        Dv, D2v = cosmo.Dv(a, D1), cosmo.D2v(a, D1)
    else:
        raise ValueError(kind)
    dx = boxsize / n
    i = np.arange(n)
    Q = np.stack(np.meshgrid(i, i, i, indexing='ij'), axis=-1).astype(np.float64)
    q = (Q + offset) * dx
    x = q + D1 * dx1 + D2 * dx2
    v = Dv * dx1 + D2v * dx2
    ids = np.arange(1, n**3 + 1, dtype=np.uint64)
    r = lambda f: f.reshape(-1, 3)
    return dict(x=r(x), v=r(v), dx1=r(dx1), dx2=r(dx2), id=ids, a_x=a, a_v=a)
