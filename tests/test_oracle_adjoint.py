"""CPU checks of the identities the hand-written adjoint relies on, and of
the float64 autograd reference (oracle/diff.py) against finite differences."""
import numpy as np
import pytest

from oracle import cosmology, lpt, pm, power, rng


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("n", [8, 12, 16])
def test_green_gradient_operator_is_antisymmetric(n):
    """M_a = c2r . (i f2 k_a) . r2c (pm.cpp:460-508, Nyquist +n/2, c2r
    dropping Im of the kz = 0, n/2 planes) satisfies M_a^T = -M_a, so the
    adjoint of the force-mesh operator is minus itself."""
    u, v = _rand((n, n, n), 1), _rand((n, n, n), 2)
    Mu, Mv = pm.force_meshes(u, 7.0), pm.force_meshes(v, 7.0)
    for a in range(3):
        lhs = np.sum(Mu[a] * v)
        rhs = -np.sum(u * Mv[a])
        assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1e-300) + 1e-12


def _lpt_ops(n, L):
    """The four LPT linear maps as functions of a real field."""
    kxy, kz = lpt.kvectors_lpt(n, L)
    KX, KY, KZ = kxy[:, None, None], kxy[None, :, None], kz[None, None, :]
    K = (KX, KY, KZ)
    k2 = KX**2 + KY**2 + KZ**2
    k2s = np.where(k2 > 0, k2, 1.0)
    h = n // 2
    mask = np.ones(k2.shape, bool)
    mask[h], mask[:, h], mask[:, :, h], mask[0, 0, 0] = False, False, False, False
    amp = np.where(mask, 1.0 / (1.0 + np.sqrt(k2)), 0.0)

    def op(mult):
        return lambda f: lpt.c2r(mult * np.fft.rfftn(f), n)

    T = [op(1j * Ki / k2s * amp) for Ki in K]                       # w -> dx1_a
    S = [op(-K[i] * K[j] / k2s * amp) for (i, j) in [(0, 0), (0, 1), (1, 2)]]   # w -> psi_ij
    B = [op(np.where(k2 > 0, -1j * Ki / k2s, 0.0)) for Ki in K]     # div2 -> n^3 dx2_a
    return T, S, B


@pytest.mark.parametrize("n", [8, 16])
def test_lpt_operator_symmetries(n):
    """dx1 = T_a w with T_a^T = -T_a; psi_ij = S_ij w with S_ij^T = S_ij;
    dx2 = B_a div2 / n^3 with B_a^T = -B_a (lpt.cpp conventions, Nyquist
    index mapped to -n/2)."""
    T, S, B = _lpt_ops(n, 11.0)
    u, v = _rand((n, n, n), 3), _rand((n, n, n), 4)
    for op, sign in [(t, -1) for t in T] + [(s, 1) for s in S] + [(b, -1) for b in B]:
        lhs = np.sum(op(u) * v)
        rhs = sign * np.sum(u * op(v))
        assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))


def test_float64_autograd_reference_matches_finite_differences():
    torch = pytest.importorskip("torch")
    from oracle import diff
    n, nc, L = 8, 16, 32.0
    cosmo = cosmology.Cosmology(0.31)
    eh = power.EH98NoWiggle()
    noise = rng.white_noise(n, 3)
    target = np.asarray(diff.cola_density(torch.tensor(rng.white_noise(n, 4)), L, nc, 0.1, 1.0,
                                          3, eh.P, cosmo))
    loss, grad, _ = diff.loss_and_grad(noise, target, L, nc, 0.1, 1.0, 3, eh.P, cosmo)
    d = np.random.default_rng(5).standard_normal(noise.shape)
    d /= np.linalg.norm(d)
    eps = 1e-4

    def f(w):
        with torch.no_grad():
            delta = diff.cola_density(torch.tensor(w), L, nc, 0.1, 1.0, 3, eh.P, cosmo)
            return float(torch.mean((delta - torch.tensor(target)) ** 2))
    fd = (f(noise + eps * d) - f(noise - eps * d)) / (2 * eps)
    an = float(np.sum(grad * d))
    assert abs(fd - an) < 2e-3 * abs(an) + 1e-12, (fd, an)
