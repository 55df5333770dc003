"""Float64 autograd reference of the differentiable COLA run (test oracle only).

No reference counterpart: fs2 has no adjoint.  This restates, in torch
float64 on the CPU, the same forward the GPU path differentiates --

  white noise w -> 2LPT (oracle/lpt.py conventions) -> x0 = q + D1 dx1 + D2 dx2,
  v0 = 0 ('cola' kind) -> nstep COLA steps (oracle/cola.py; forces by the
  PM chain of oracle/pm.py) -> final CIC density delta (oracle/cic.py)
  -> loss = mean((delta - target)^2)

-- and lets torch.autograd produce d loss / d w.  The periodic wrap is a
translation by a multiple of the box, so it carries unit derivative; CIC
weights are piecewise linear in x (torch differentiates them a.e.).
"""
import numpy as np
import torch

from . import cola as ocola

T = torch.float64


def _kvec(n, boxsize, lpt_convention):
    i = torch.arange(n, dtype=T)
    if lpt_convention:
        kxy = torch.where(i < n // 2, i, i - n)
        iz = torch.arange(n // 2 + 1, dtype=T)
        kz = torch.where(iz < n // 2, iz, iz - n)
        s = 2 * np.pi / boxsize
        return kxy * s, kz * s
    kxy = torch.where(i <= n // 2, i, i - n)
    kz = torch.arange(n // 2 + 1, dtype=T)
    return kxy, kz


def _c2r(fk, n):
    return torch.fft.irfftn(fk, s=(n, n, n), dim=(0, 1, 2)) * n**3


def lpt_fields(noise, boxsize, Pfun):
    """(dx1, dx2) as (n^3, 3) tensors from the noise (oracle/lpt.py)."""
    n = noise.shape[0]
    W = torch.fft.rfftn(noise, dim=(0, 1, 2))
    kxy, kz = _kvec(n, boxsize, True)
    KX, KY, KZ = kxy[:, None, None], kxy[None, :, None], kz[None, None, :]
    k2 = KX**2 + KY**2 + KZ**2
    kmag = torch.sqrt(k2).numpy()
    amp = np.zeros_like(kmag)
    nz = kmag > 0
    amp[nz] = np.sqrt(Pfun(kmag[nz]) / boxsize**3) / n**1.5
    h = n // 2
    mask = np.ones(kmag.shape, dtype=bool)
    mask[h, :, :] = False
    mask[:, h, :] = False
    mask[:, :, h] = False
    mask[0, 0, 0] = False
    amp = torch.tensor(np.where(mask, amp, 0.0), dtype=T)
    dk = W * amp
    K = (KX, KY, KZ)
    k2s = torch.where(k2 > 0, k2, torch.ones_like(k2))
    psi_k = [1j * Ki / k2s * dk for Ki in K]
    pij = {}
    for (i, j) in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
        pij[(i, j)] = _c2r(1j * K[j] * psi_k[i], n)
    div2 = (pij[(0, 0)] * (pij[(1, 1)] + pij[(2, 2)]) + pij[(1, 1)] * pij[(2, 2)]
            - pij[(0, 1)]**2 - pij[(0, 2)]**2 - pij[(1, 2)]**2)
    Dk = torch.fft.rfftn(div2, dim=(0, 1, 2))
    zero = (k2 == 0)
    psi2_k = [torch.where(zero, torch.zeros_like(Dk), -1j * Ki / k2s * Dk) for Ki in K]
    dx1 = torch.stack([_c2r(p, n) for p in psi_k], -1).reshape(-1, 3)
    dx2 = (torch.stack([_c2r(p, n) for p in psi2_k], -1) / n**3).reshape(-1, 3)
    return dx1, dx2


def paint(x, nc, boxsize, fac):
    """CIC density (pm.cpp:57-139) of wrapped positions x (N,3)."""
    X = x * (nc / boxsize)
    i0 = torch.floor(X).detach()
    w1 = X - i0
    w0 = 1.0 - w1
    i0 = i0.long() % nc
    i1 = (i0 + 1) % nc
    rho = torch.zeros(nc**3, dtype=T)
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                ix = (i0, i1)[cx][:, 0]
                iy = (i0, i1)[cy][:, 1]
                iz = (i0, i1)[cz][:, 2]
                w = (w0, w1)[cx][:, 0] * (w0, w1)[cy][:, 1] * (w0, w1)[cz][:, 2]
                rho = rho.index_add(0, (ix * nc + iy) * nc + iz, fac * w)
    return rho.reshape(nc, nc, nc) - 1.0


def readout(x, meshes, boxsize):
    nc = meshes[0].shape[0]
    X = x * (nc / boxsize)
    i0 = torch.floor(X).detach()
    w1 = X - i0
    w0 = 1.0 - w1
    i0 = i0.long() % nc
    i1 = (i0 + 1) % nc
    out = [0.0, 0.0, 0.0]
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                ix = (i0, i1)[cx][:, 0]
                iy = (i0, i1)[cy][:, 1]
                iz = (i0, i1)[cz][:, 2]
                w = (w0, w1)[cx][:, 0] * (w0, w1)[cy][:, 1] * (w0, w1)[cz][:, 2]
                flat = (ix * nc + iy) * nc + iz
                for a, m in enumerate(meshes):
                    out[a] = out[a] + m.reshape(-1)[flat] * w
    return torch.stack(out, -1)


def force_meshes(delta, boxsize):
    n = delta.shape[0]
    dk = torch.fft.rfftn(delta, dim=(0, 1, 2))
    kxy, kz = _kvec(n, boxsize, False)
    KX, KY, KZ = kxy[:, None, None], kxy[None, :, None], kz[None, None, :]
    k2 = KX**2 + KY**2 + KZ**2
    k2s = torch.where(k2 > 0, k2, torch.ones_like(k2))
    f1 = -1.0 / n**3 / (2.0 * np.pi / boxsize)
    out = []
    for K in (KX, KY, KZ):
        g = torch.where(k2 > 0, 1j * f1 / k2s * K, torch.zeros_like(k2s) * 1j)
        out.append(_c2r(g * dk, n))
    return out


def cola_density(noise, boxsize, nc, a0, a1, nstep, Pfun, cosmo):
    """Final CIC density of a 'cola' run started from `noise` (n_p^3)."""
    n = noise.shape[0]
    fac = (nc / n) ** 3
    dx1, dx2 = lpt_fields(noise, boxsize, Pfun)
    i = torch.arange(n, dtype=T)
    q = (torch.stack(torch.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5) * (boxsize / n)
    D1 = cosmo.D(a0)
    D2 = cosmo.D2(a0, D1)
    x = q + D1 * dx1 + D2 * dx2
    v = torch.zeros_like(x)
    a_x = a_v = a0
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        xw = torch.remainder(x, boxsize)
        delta = paint(xw, nc, boxsize, fac)
        f = readout(xw, force_meshes(delta, boxsize), boxsize)
        kf, q1, q2 = ocola.kick_coefficients(cosmo, a_v, a_x, a_vel)
        v = v + (-1.5 * cosmo.om) * (f + q1 * dx1 + q2 * dx2) * kf
        a_v = a_vel
        dt, da1, da2 = ocola.drift_coefficients(cosmo, a_x, a_v, a_pos)
        x = xw + v * dt + (dx1 * da1 + dx2 * da2)
        a_x = a_pos
    return paint(torch.remainder(x, boxsize), nc, boxsize, fac)


def loss_and_grad(noise_np, target_np, boxsize, nc, a0, a1, nstep, Pfun, cosmo):
    noise = torch.tensor(noise_np, dtype=T, requires_grad=True)
    target = torch.tensor(target_np, dtype=T)
    delta = cola_density(noise, boxsize, nc, a0, a1, nstep, Pfun, cosmo)
    loss = torch.mean((delta - target) ** 2)
    loss.backward()
    return float(loss.detach()), noise.grad.numpy().copy(), delta.detach().numpy()
