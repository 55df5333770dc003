"""GPU parity against the committed golden fixtures (tests/golden/*.npz,
written from the FP64 oracle by tests/golden/make_golden.py): the HIP path is
compared with stored vectors only, nothing is recomputed on the host.

Tolerances are those of the live-oracle tests: CIC per cell 1e-5 relative
for the FP32-atomic paints and one FP32 rounding for the fixed-point one;
forces 1e-5 relative RMS; 2LPT 2e-5 of the field scale; COLA positions
1e-3 mesh cells."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _announce(msg):
    print(f"[case] {msg}", flush=True)


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


def _eh(fs, z, prefix="eh_"):
    return fs.PowerSpectrum.eisenstein_hu(*(float(z[prefix + k]) for k in
                                            ("omega_m", "omega_b", "h", "ns", "sigma8")))


@pytest.mark.parametrize("case", ["uniform16", "clustered32"])
def test_paint_matches_golden(fs, case):
    """py/pm.cpp:57-139 after the py/util.h:14-24 wrap, all three paints."""
    _announce(f"golden paint {case}")
    z = _load("cic")
    n, L, pmf = int(z[f"{case}_n"]), float(z[f"{case}_boxsize"]), float(z[f"{case}_pm_factor"])
    x, ref = z[f"{case}_x"], z[f"{case}_delta"]
    fs.pm.init(n, pmf, L)
    p = fs.Particles(len(x), L)
    p.append(x.astype(np.float64))   # append takes doubles; exact for float32
    rho = np.abs(ref + 1.0)
    out = {}
    try:
        for variant in ("global", "atomic", "sorted"):
            fs.pm.set_paint_variant(variant)
            fs.pm.send_positions(p)
            out[variant] = fs.pm.compute_density(p).asarray().astype(np.float64)
            err = np.abs(out[variant] - ref)
            assert np.all(err <= 1e-5 * rho + 1e-6), (variant, float((err / (rho + 1e-6)).max()))
    finally:
        fs.pm.set_paint_variant("sorted")
    assert np.all(np.abs(out["sorted"] - ref) <= 1.2e-7 * (rho + 1.0))
    xw = p.x
    assert xw.min() >= 0 and xw.max() < L
    d = xw.astype(np.float64) - x
    assert np.all((np.abs(d) < 1e-4) | (np.abs(np.abs(d) - L) < 1e-4))   # moved by 0 or L


@pytest.mark.parametrize("case", ["lattice8", "planewave8"])
def test_force_matches_golden(fs, case):
    """Forces at the particles (py/pm.cpp:57-220, 414-508): relative RMS
    <= 1e-5; the plane wave's transverse forces vanish."""
    _announce(f"golden force {case}")
    z = _load("pm_force")
    nc, L, pmf = int(z[f"{case}_nc"]), float(z[f"{case}_boxsize"]), float(z[f"{case}_pm_factor"])
    x, ref = z[f"{case}_x"], z[f"{case}_force"]
    fs.pm.init(nc, pmf, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.Particles(len(x), L)
    p.append(x.astype(np.float64))   # append takes doubles; exact for float32
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)
    got = p.force.astype(np.float64)
    rel = float(np.sqrt(np.mean((got - ref) ** 2) / np.mean(ref**2)))
    assert rel < 1e-5, rel
    assert np.array_equal(p.x, x)      # in the box already: the wrap keeps them
    if case == "planewave8":
        rms = np.sqrt(np.mean(got[:, 0] ** 2))
        assert np.abs(got[:, 1:]).max() < 1e-5 * rms


def test_lpt_matches_golden(fs):
    """Philox white noise of seed 7 and the 2LPT state at a = 0.1
    (py/lpt.cpp:100-225, 439-591)."""
    _announce("golden 2lpt 16^3")
    z = _load("lpt")
    n, L, a, seed = int(z["n"]), float(z["boxsize"]), float(z["a"]), int(z["seed"])
    fs.cosmology.init(float(z["eh_omega_m"]))
    eh = _eh(fs, z)
    noise = fs.lpt.white_noise(n, seed).astype(np.float64)
    assert np.abs(noise - z["noise"]).max() < 1e-6 * np.abs(z["noise"]).max()
    p = fs.lpt.init(n, L, a, eh, seed, "2lpt")
    assert np.array_equal(p.id, z["id"])
    scale1 = np.sqrt(np.mean(z["dx1"].astype(np.float64) ** 2))
    for name in ("dx1", "dx2", "x", "v"):
        r = z[name].astype(np.float64)
        got = getattr(p, name).astype(np.float64)
        scale = scale1 if name == "x" else np.sqrt(np.mean(r**2))
        err = float(np.abs(got - r).max() / scale)
        assert err < 2e-5, (name, err)


@pytest.mark.parametrize("om", [0.308, 0.31])
@pytest.mark.parametrize("fused", [False, True])
def test_device_kick_drift_use_golden_coefficients(fs, om, fused):
    """Each step of the 10-step schedule on three probe particles: f = 1,
    dx1 = 1 and dx2 = 1 from v = 0, x = 0, so v and x show the kick and drift
    coefficients (py/cola.cpp:48-127) directly.  With `fused` the kick stays
    deferred into the drift; otherwise reading v applies it first."""
    _announce(f"golden kick/drift coefficients om={om} fused={fused}")
    z = _load("cosmology")
    tag = f"om{int(round(om * 1000))}"
    fs.cosmology.init(om)
    L = 64.0
    unit = np.eye(3, dtype=np.float32)[:, :1].repeat(3, axis=1)   # rows: (1,1,1), 0, 0
    a_x = a_v = 0.1
    for k, (a_vel, a_pos) in enumerate(z["schedule"]):
        kf, q1, q2 = z[f"{tag}_kick"][k]
        dt, da1, da2 = z[f"{tag}_drift"][k]
        p = fs.Particles(3, L)
        p.append(np.zeros((3, 3)))
        p.set_field("force", unit)
        p.set_field("dx1", np.roll(unit, 1, axis=0))
        p.set_field("dx2", np.roll(unit, 2, axis=0))
        p.set_times(a_x, a_v, a_x)
        fs.cola.kick(p, float(a_vel))
        v_expect = -1.5 * om * kf * np.array([1.0, q1, q2])
        if not fused:
            v = p.v.astype(np.float64)
            assert np.abs(v - v_expect[:, None]).max() <= 4e-7 * np.abs(v_expect).max(), k
        fs.cola.drift(p, float(a_pos))
        x_expect = v_expect * dt + np.array([0.0, da1, da2])
        x = p.x.astype(np.float64)
        # FP32 coefficients and products: a few roundings of the largest term
        bound = 1e-6 * (np.abs(v_expect * dt).max() + max(abs(da1), abs(da2)))
        assert np.abs(x - x_expect[:, None]).max() <= bound, k
        if fused:
            v = p.v.astype(np.float64)
            assert np.abs(v - v_expect[:, None]).max() <= 4e-7 * np.abs(v_expect).max(), k
        a_v, a_x = float(a_vel), float(a_pos)


def test_cola_two_steps_match_golden(fs):
    """16^3 particles, 32^3 mesh, a = 0.1 -> 0.3 in 2 COLA steps from the
    2LPT state of seed 11: positions within 1e-3 mesh cells of the stored
    FP64 trajectory, velocities within 1e-4 of their largest value."""
    _announce("golden cola 2 steps 16^3/32^3")
    z = _load("cola_2step")
    n_p, nc, L = int(z["n_p"]), int(z["nc"]), float(z["boxsize"])
    fs.cosmology.init(float(z["omega_m"]))
    lz = _load("lpt")
    eh = _eh(fs, lz)
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.lpt.init(n_p, L, float(z["a0"]), eh, int(z["seed"]), "cola")
    for a_vel, a_pos in z["schedule"]:
        fs.pm.force(p)
        fs.cola.kick(p, float(a_vel))
        fs.cola.drift(p, float(a_pos))
    d = p.x.astype(np.float64) - z["x"]
    d -= L * np.round(d / L)
    err = float(np.abs(d).max() / (L / nc))
    assert err < 1e-3, err
    vref = z["v"].astype(np.float64)
    verr = float(np.abs(p.v - vref).max() / np.abs(vref).max())
    assert verr < 1e-4, verr
    assert p.times[:2] == (float(z["schedule"][-1][1]), float(z["schedule"][-1][0]))
    assert abs(p.times[0] - float(z["a1"])) < 1e-12
