"""The committed golden fixtures (tests/golden/*.npz) on the CPU: the oracle
still reproduces them, they satisfy what the physics fixes (mass and
momentum conservation, the plane wave's symmetry, the growth closed form and
the reference's known growth values), and the product's FP64 host numerics
(cosmology, COLA coefficients) match the coefficient table."""
import os
import sys

import numpy as np
import pytest

from oracle import cosmology as ocosmo

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

NAMES = ("cic", "pm_force", "lpt", "cosmology", "cola_2step")


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def rebuilt():
    return make_golden.build_all()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(rebuilt, name):
    """Same keys, dtypes and shapes; float64 values to 1e-12 of the array's
    scale, float32 ones to one rounding of it."""
    ref, got = _load(name), rebuilt[name]
    assert sorted(ref) == sorted(got)
    for k, r in ref.items():
        g = np.asarray(got[k])
        assert g.dtype == r.dtype and g.shape == r.shape, k
        if r.dtype.kind in "iu":
            assert np.array_equal(g, r), k
            continue
        scale = max(float(np.abs(r).max()), 1e-300)
        tol = 1e-12 if r.dtype == np.float64 else 2.4e-7
        assert np.abs(g.astype(np.float64) - r).max() <= tol * scale, k


def test_cic_fixture_conserves_mass():
    """Sum(delta) = Np pm^3 - Ng = 0 for 512 / 4096 particles at pm_factor 2
    (py/pm.cpp:357-390 tolerance 10 FLT_EPS Ng); every cell has delta >= -1."""
    z = _load("cic")
    for case in ("uniform16", "clustered32"):
        d, n = z[f"{case}_delta"], int(z[f"{case}_n"])
        assert len(z[f"{case}_x"]) * 8 == n**3
        assert abs(d.sum()) < 10 * 1.2e-7 * n**3
        assert d.min() >= -1.0
    # the clumps: a few cells carry hundreds of particles
    assert z["clustered32_delta"].max() > 50


def test_force_fixture_conservation_and_plane_wave_symmetry():
    z = _load("pm_force")
    for case in ("lattice8", "planewave8"):
        f = z[f"{case}_force"]
        rms = np.sqrt(np.mean(f**2))
        assert np.abs(f.sum(axis=0)).max() < 1e-10 * rms * len(f)   # momentum
    # a displacement along x only: f_y = f_z = 0 and f_x a function of the
    # lattice plane alone, odd about the planes where sin(2 pi q / L) = 0
    f, x, L = z["planewave8_force"], z["planewave8_x"].astype(np.float64), float(z["planewave8_boxsize"])
    rms = np.sqrt(np.mean(f[:, 0] ** 2))
    assert rms > 0 and np.abs(f[:, 1:]).max() < 1e-10 * rms
    fx = f[:, 0].reshape(8, 64)
    assert np.abs(fx - fx[:, :1]).max() < 1e-10 * rms
    # odd up to the float32 rounding of the displaced positions (~1e-6)
    assert np.abs(fx[:, 0] + fx[::-1, 0]).max() < 1e-5 * rms
    # each x plane holds 64 particles at distinct (y, z)
    yz = x[:, 1:].reshape(8, 64, 2)
    assert all(len({tuple(r) for r in plane}) == 64 for plane in yz)


def test_lpt_fixture_properties():
    z = _load("lpt")
    assert np.array_equal(z["id"], np.arange(1, 16**3 + 1, dtype=np.uint64))
    w = z["noise"].astype(np.float64)
    assert abs(w.mean()) < 0.05 and abs(w.std() - 1) < 0.05
    c = ocosmo.Cosmology(float(z["eh_omega_m"]))
    a = float(z["a"])
    D1, D2 = c.D(a), c.D2(a, c.D(a))
    L, n = float(z["boxsize"]), int(z["n"])
    i = np.arange(n)
    q = ((np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5) * (L / n))
    x = q + D1 * z["dx1"].astype(np.float64) + D2 * z["dx2"].astype(np.float64)
    assert np.abs(x - z["x"]).max() < 1e-5 * L
    v = c.Dv(a, D1) * z["dx1"].astype(np.float64) + c.D2v(a, D1) * z["dx2"].astype(np.float64)
    assert np.abs(v - z["v"]).max() < 1e-5 * np.abs(z["v"]).max()
    # the displacement fields have zero mean (no k = 0 mode)
    for k in ("dx1", "dx2"):
        f = z[k].astype(np.float64)
        assert np.abs(f.mean(axis=0)).max() < 1e-5 * np.sqrt(np.mean(f**2))


def test_cosmology_fixture_known_values():
    """D1(1) = 1; D1(0.1) = 0.127351097784 and f(1) = 0.52237017 for
    Omega_m = 0.31 (py/cosmology.cpp:153-171 with an independent
    quadrature); D1 against the hypergeometric closed form; D2 against
    -3/7 D1^2 Omega(a)^(-1/143)."""
    z = _load("cosmology")
    a = z["a"]
    assert abs(z["om310_D1"][a == 0.1][0] - 0.127351097784) < 1e-11
    assert abs(z["om310_f"][a == 1.0][0] - 0.52237017) < 1e-8
    for om in (0.308, 0.31):
        tag = f"om{int(round(om * 1000))}"
        c = ocosmo.Cosmology(om)
        D = z[f"{tag}_D1"]
        assert abs(D[-1] - 1.0) < 1e-14
        closed = np.array([c.D_closed_form(x) for x in a])
        assert np.abs(D - closed).max() < 1e-12
        omega = om / (om + (1 - om) * a**3)
        assert np.abs(z[f"{tag}_D2"] + 3 / 7 * D**2 * omega ** (-1 / 143)).max() < 1e-14
        # kick factor (py/cola.cpp:58-60): closed form in a_v, a_x, a_vel
        sched = z["schedule"]
        a_v = np.concatenate([[0.1], sched[:-1, 0]])
        a_x = np.concatenate([[0.1], sched[:-1, 1]])
        n = -2.5
        kf = (sched[:, 0] ** n - a_v**n) / (n * a_x**n * np.sqrt(om / a_x + (1 - om) * a_x**2))
        assert np.abs(z[f"{tag}_kick"][:, 0] - kf).max() < 1e-14 * np.abs(kf).max()


@pytest.mark.parametrize("om", [0.308, 0.31])
def test_host_numerics_match_cosmology_fixture(fs, om):
    """The library's FP64 host cosmology and COLA coefficients against the
    table (no device needed)."""
    z = _load("cosmology")
    tag = f"om{int(round(om * 1000))}"
    fs.cosmology.init(om)
    for k, a in enumerate(z["a"]):
        assert abs(fs.cosmology.D_growth(a) - z[f"{tag}_D1"][k]) < 1e-12 * z[f"{tag}_D1"][k]
        assert abs(fs.cosmology.D2_growth(a) - z[f"{tag}_D2"][k]) < 1e-12 * abs(z[f"{tag}_D2"][k])
        assert abs(fs.cosmology.f_growth_rate(a) - z[f"{tag}_f"][k]) < 1e-11
    a_x = a_v = 0.1
    for k, (a_vel, a_pos) in enumerate(z["schedule"]):
        kick = np.array(fs._fs._cola_kick_coefficients(a_v, a_x, a_vel))
        assert np.abs(kick - z[f"{tag}_kick"][k]).max() <= 1e-12 * np.abs(z[f"{tag}_kick"][k]).max()
        a_v = a_vel
        drift = np.array(fs._fs._cola_drift_coefficients(a_x, a_v, a_pos))
        assert np.all(np.abs(drift - z[f"{tag}_drift"][k]) <= 1e-11 * np.abs(z[f"{tag}_drift"][k]))
        a_x = a_pos
