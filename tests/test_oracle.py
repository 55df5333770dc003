"""CPU checks that pin the oracle to the reference's known answers and to
closed forms (no GPU needed)."""
import os

import numpy as np
import pytest

from oracle import cic, cola, cosmology, lpt, pk, pm, power, rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# ---- Philox4x32-10: published known-answer vectors (Random123 kat_vectors)
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, expect):
    out = rng.philox4x32_10([np.array([c], dtype=np.uint64) for c in ctr],
                            (np.uint64(key[0]), np.uint64(key[1])))
    assert [int(o[0]) for o in out] == expect


def test_white_noise_statistics():
    w = rng.white_noise(32, 99)
    assert abs(w.mean()) < 0.02 and abs(w.std() - 1.0) < 0.02
    # distinct seeds give independent fields
    w2 = rng.white_noise(32, 100)
    assert abs(np.corrcoef(w.ravel(), w2.ravel())[0, 1]) < 0.02


# ---- cosmology: quadrature vs closed forms
@pytest.mark.parametrize("om", [0.308, 0.31, 0.5])
def test_growth_closed_form(om):
    c = cosmology.Cosmology(om)
    for a in (0.05, 0.1, 0.3, 0.7, 1.0):
        assert abs(c.D(a) - c.D_closed_form(a)) < 1e-12 * c.D(a)


def test_growth_known_value():
    c = cosmology.Cosmology(0.31)
    assert abs(c.D(0.1) - 0.127351097784) < 1e-11
    assert abs(c.f(1.0) - 0.52237017) < 1e-7


def test_eds_limits():
    """Omega_m = 1: D = a, f = 1, D2 = -3/7 a^2."""
    c = cosmology.Cosmology(1.0)
    for a in (0.1, 0.5, 1.0):
        assert abs(c.D(a) - a) < 1e-12
        assert abs(c.f(a) - 1.0) < 1e-10
        assert abs(c.D2(a) + 3.0 / 7.0 * a * a) < 1e-12


def test_cola_coefficients_eds_closed_form():
    """EdS: kf = (af^n - ai^n)/(n a^n a^-1/2); dt = int a^(n-3/2) da / av^n."""
    c = cosmology.Cosmology(1.0)
    n = cola.NLPT
    ai, a, af = 0.2, 0.25, 0.3
    kf, q1, q2 = cola.kick_coefficients(c, ai, a, af)
    assert abs(kf - (af**n - ai**n) / (n * a**n * a**-0.5)) < 1e-13
    assert abs(q1 - a) < 1e-12 and abs(q2 - (-3 / 7 * a * a - a * a)) < 1e-12
    dt = cola.drift_dt(c, 0.2, 0.3, 0.25)
    e = n - 1.5
    exact = (0.3**(e + 1) - 0.2**(e + 1)) / (e + 1) / 0.25**n
    assert abs(dt - exact) < 1e-12 * abs(exact)


# ---- standard leapfrog factors
class _Om:
    def __init__(self, om):
        self.om = om


@pytest.mark.parametrize("ai,af", [(0.1, 0.2), (0.3, 1.0)])
def test_leapfrog_factors_closed_forms(ai, af):
    """py/leapfrog.cpp:96-141 integrals: Om = 1 (EdS, H = a^-3/2) gives
    Sphi = 2(sqrt af - sqrt ai), Sq = 2(ai^-1/2 - af^-1/2); Om = 0 (de
    Sitter, H = 1) gives Sphi = 1/ai - 1/af, Sq = (ai^-2 - af^-2)/2."""
    from oracle import leapfrog as olf
    eds, ds = _Om(1.0), _Om(0.0)
    assert olf.kick_factor(eds, ai, af) == pytest.approx(2 * (af**0.5 - ai**0.5), rel=1e-12)
    assert olf.drift_factor(eds, ai, af) == pytest.approx(2 * (ai**-0.5 - af**-0.5), rel=1e-12)
    assert olf.kick_factor(ds, ai, af) == pytest.approx(1 / ai - 1 / af, rel=1e-12)
    assert olf.drift_factor(ds, ai, af) == pytest.approx((ai**-2 - af**-2) / 2, rel=1e-12)


# ---- CIC: test/test_pm_cic.py known answers and closed-form weights
@pytest.mark.parametrize("xc", [0, 1, 1.5, 3, 3.5, 4])
def test_cic_one_particle_known_answer(xc):
    nc, L = 4, 64.0
    dx = L / nc
    x = cic.wrap(np.array([[xc * dx, 2 * dx, 2 * dx]]), L)
    d = cic.paint(x, nc, L, 1.0, wdtype=np.float32)
    assert abs(np.sum(d + 1.0) - 1.0) < 1e-15
    # closed form: weight (1-t) at floor cell, t at the next (periodic)
    X = (xc % nc)
    i0 = int(np.trunc(X)) % nc
    t = X - np.trunc(X)
    expect = np.zeros((nc, nc, nc))
    expect[i0, 2, 2] += 1 - t
    expect[(i0 + 1) % nc, 2, 2] += t
    assert np.allclose(d + 1.0, expect, atol=0)


def test_cic_truncation_not_floor():
    """pm.cpp:95: (int) truncates toward zero: x in (-1, 0) cells -> cell 0."""
    nc, L = 8, 8.0
    d = cic.paint(np.array([[-0.25, 0.5, 0.5]]), nc, L, 1.0, wdtype=np.float32)
    assert d[0, 0, 0] + 1 == pytest.approx(1.25 * 0.5 * 0.5)


def test_paint_readout_adjoint():
    """<paint(m), g> = <m, readout(g)> for the CIC kernel."""
    r = np.random.default_rng(1)
    nc, L, N = 16, 10.0, 500
    x = r.uniform(0, L, (N, 3))
    g = r.standard_normal((nc, nc, nc))
    lhs = np.sum((cic.paint(x, nc, L, 1.0) + 1.0) * g)
    rhs = np.sum(cic.readout(x, [g], L)[:, 0])
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_mass_conservation():
    r = np.random.default_rng(2)
    nc, L, N = 16, 10.0, 4096
    d = cic.paint(r.uniform(0, L, (N, 3)), nc, L, 2.0)
    assert abs(d.sum() - (N * 8 - nc**3)) < 1e-9


# ---- PM force: plane wave closed form, uniform lattice, momentum
def test_force_plane_wave():
    nc, L, m, A = 32, 50.0, 2, 0.3
    i = np.arange(nc)
    Y = (i * L / nc)[None, :, None] * np.ones((nc, 1, nc))
    delta = A * np.cos(2 * np.pi * m * Y / L)
    f = pm.force_meshes(delta, L)
    k = 2 * np.pi * m / L
    assert np.abs(f[1] - A / k * np.sin(k * Y)).max() < 1e-12
    assert np.abs(f[0]).max() < 1e-12 and np.abs(f[2]).max() < 1e-12


def test_force_uniform_lattice_is_zero():
    n_p, nc, L = 8, 16, 16.0
    i = np.arange(n_p)
    q = (np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5) * L / n_p
    f = pm.force(q, nc, L, 2.0)
    assert np.abs(f).max() < 1e-12


def test_force_momentum_conservation():
    r = np.random.default_rng(3)
    x = r.uniform(0, 20.0, (2000, 3))
    f = pm.force(x, 16, 20.0, 1.0)
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).sum()


def test_force_fp32_port_matches_fp64():
    r = np.random.default_rng(4)
    x = r.uniform(0, 20.0, (4000, 3))
    d = cic.paint(x, 16, 20.0, 1.0)
    m64 = pm.force_meshes(d, 20.0)
    m32 = pm.force_meshes_fp32(d, 20.0, workers=2)
    for a, b in zip(m64, m32):
        assert np.abs(a - b).max() < 1e-5 * np.abs(a).max()


# ---- 2LPT identities
def _eh():
    return power.EH98NoWiggle(0.31, 0.049, 0.67, 0.96, 0.81)


def test_lpt_div_psi1_is_minus_delta():
    n, L = 16, 64.0
    noise = rng.white_noise(n, 3)
    eh = _eh()
    dk, (KX, KY, KZ, k2) = lpt.delta_k_from_noise(noise, L, eh.P)
    dx1, _ = lpt.displacements(noise, L, eh.P)
    div = sum(np.fft.rfftn(dx1[..., a]) / n**3 * 1j * K for a, K in enumerate((KX, KY, KZ)))
    assert np.abs(div + dk).max() < 1e-10 * np.abs(dk).max()


def test_lpt_2lpt_source_single_plane_wave_vanishes():
    """A single plane wave has Psi_ij of rank one, so div Psi2 = 0."""
    n, L = 16, 16.0
    x = (np.arange(n) * L / n)[:, None, None] * np.ones((1, n, n))
    noise = np.cos(2 * np.pi * 2 * x / L) / n**1.5
    dx1, dx2 = lpt.displacements(noise, L, lambda k: np.ones_like(k) * L**3)
    assert np.abs(dx1).max() > 0.1
    assert np.abs(dx2).max() < 1e-12 * max(1.0, np.abs(dx1).max())


def test_lpt_2lpt_two_crossed_waves_closed_form():
    """Psi1 = (a sin(kx), b sin(ky), 0)/k: div Psi2 = Psi11 Psi22 = a b cos cos;
    Psi2 = -grad lap^-1 div -> Psi2_x = a b/(2k) sin(kx) cos(ky)."""
    n, L, m = 16, 16.0, 1
    k = 2 * np.pi * m / L
    i = np.arange(n) * L / n
    X, Y = np.meshgrid(i, i, indexing="ij")
    X = X[:, :, None] * np.ones((1, 1, n))
    Y = Y[:, :, None] * np.ones((1, 1, n))
    # delta = -div Psi1; with P = 1 the amplitude is sqrt(1/V)/n^1.5 and
    # delta(x) = c2r(delta_k) = rfftn(noise) amp, so noise = delta/(n^1.5 amp)
    a, b = 0.7, 0.4
    delta = -(a * np.cos(k * X) + b * np.cos(k * Y))
    noise = delta / (n**1.5 * np.sqrt(1.0 / L**3))
    dx1, dx2 = lpt.displacements(noise, L, lambda kk: np.ones_like(kk))
    assert np.abs(dx1[..., 0] - a / k * np.sin(k * X)).max() < 1e-10
    assert np.abs(dx1[..., 1] - b / k * np.sin(k * Y)).max() < 1e-10
    # div Psi2 = Psi_11 Psi_22 = a b cos(kX) cos(kY); Psi2_k = -i k/k^2 D_k
    # => Psi2_x = a b / (2k) sin(kX) cos(kY)
    assert np.abs(dx2[..., 0] - a * b / (2 * k) * np.sin(k * X) * np.cos(k * Y)).max() < 1e-10
    assert np.abs(dx2[..., 1] - a * b / (2 * k) * np.cos(k * X) * np.sin(k * Y)).max() < 1e-10


def test_lpt_lattice_ids_and_kinds():
    n, L = 8, 32.0
    c = cosmology.Cosmology(0.31)
    noise = rng.white_noise(n, 5)
    p0 = lpt.particles(noise, L, 0.0, _eh().P, c, "cola")
    i = np.arange(n)
    q = (np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5) * L / n
    assert np.allclose(p0["x"], q, atol=0)
    assert np.array_equal(p0["id"], np.arange(1, n**3 + 1))
    pz = lpt.particles(noise, L, 0.5, _eh().P, c, "zeldovich")
    D1 = c.D(0.5)
    assert np.allclose(pz["x"], q + D1 * pz["dx1"])
    assert np.allclose(pz["v"], c.Dv(0.5, D1) * pz["dx1"])


# ---- power spectra
def test_table_spline_reproduces_nodes_and_sigma8():
    path = os.path.join(GOLDEN, "planck_matterpower.dat")
    t = power.TablePower(path)
    assert np.allclose(t.P(t.k[5:-5]), t.Ptab[5:-5], rtol=1e-12)
    s8 = t.sigma(8.0)
    assert 0.7 < s8 < 0.9


def test_eh98_sigma8_normalisation():
    eh = _eh()
    assert abs(np.sqrt(power.tophat_sigma2(eh.P, 8.0)) - 0.81) < 1e-9


def test_pk_single_mode_amplitude():
    """delta = A cos(k.x): P at that k = V A^2 / 4 per conjugate pair."""
    n, L, m, A = 32, 100.0, 3, 0.2
    x = (np.arange(n) * L / n)[:, None, None] * np.ones((1, n, n))
    delta = A * np.cos(2 * np.pi * m * x / L)
    k, P, cnt = pk.power_spectrum(delta, L)
    assert cnt[m] > 0
    # only the (m,0,0) and (-m,0,0) modes are non-zero; bin m averages them
    # with the rest of its shell
    total = P[m] * cnt[m]
    assert abs(total - 2 * L**3 * A**2 / 4) < 1e-9 * total
