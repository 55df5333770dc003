"""GPU parity: COLA kick/drift, 2LPT initial conditions, a full COLA run and
the P(k) estimator against the FP64 oracle."""
import os

import numpy as np
import pytest

from oracle import cfast as ocfast
from oracle import cic as ocic
from oracle import cola as ocola
from oracle import cosmology as ocosmo
from oracle import lpt as olpt
from oracle import pk as opk
from oracle import pm as opm
from oracle import power as opower
from oracle import rng as orng

pytestmark = pytest.mark.gpu

OM = 0.31


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def cosmo(fs):
    fs.cosmology.init(OM)
    return ocosmo.Cosmology(OM)


@pytest.fixture(scope="module")
def eh(fs):
    return fs.PowerSpectrum.eisenstein_hu(OM, 0.049, 0.67, 0.96, 0.81)


@pytest.fixture(scope="module")
def eh_oracle():
    return opower.EH98NoWiggle(OM, 0.049, 0.67, 0.96, 0.81)


def test_white_noise_matches_philox_oracle(fs):
    _announce("white noise")
    n = 16
    got = fs.lpt.white_noise(n, 12345).astype(np.float64)
    ref = orng.white_noise(n, 12345)
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()
    assert abs(got.mean()) < 0.1 and abs(got.std() - 1) < 0.05


@pytest.mark.parametrize("kind", ["cola", "zeldovich", "2lpt"])
def test_lpt_matches_oracle(fs, cosmo, eh, eh_oracle, kind):
    _announce(f"lpt kind={kind}")
    n, L, a, seed = 16, 64.0, 0.1, 7
    p = fs.lpt.init(n, L, a, eh, seed, kind)
    noise = orng.white_noise(n, seed)
    ref = olpt.particles(noise, L, a, eh_oracle.P, cosmo, kind=kind)
    assert np.array_equal(p.id, ref["id"])
    for name in ("dx1", "dx2", "x"):
        got = getattr(p, name).astype(np.float64)
        r = ref[name]
        scale = np.sqrt(np.mean(ref["dx1"] ** 2)) if name == "x" else np.sqrt(np.mean(r**2))
        err = np.abs(got - r).max() / scale
        assert err < 2e-5, (name, err)
    v = p.v.astype(np.float64)
    if kind == "cola":
        assert np.all(v == 0)
    else:
        assert np.abs(v - ref["v"]).max() < 2e-5 * np.abs(ref["v"]).max()
    assert p.times == (a, a, 0.0)


def test_lpt_lattice_and_ids_at_a0(fs, cosmo, eh):
    """test/test_particles_h5.py:57-77: a = 0 gives the exact lattice."""
    _announce("lpt lattice a=0 offset")
    n, L = 8, 64.0
    p = fs.lpt.init(n, L, 0.0, eh, 1, "cola")
    i = np.arange(n, dtype=np.float32)
    q = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    expect = (q + np.float32(0.5)) * np.float32(L / n)
    assert np.array_equal(p.x, expect)
    assert np.array_equal(p.id, np.arange(1, n**3 + 1, dtype=np.uint64))


def test_kick_drift_match_oracle(fs, cosmo):
    _announce("kick drift")
    rng = np.random.default_rng(3)
    N, L = 1000, 64.0
    p = fs.Particles(N, L)
    p.append(rng.uniform(0, L, (N, 3)))
    for name in ("v", "dx1", "dx2", "force"):
        p.set_field(name, rng.standard_normal((N, 3)).astype(np.float32))
    x0, v0 = p.x.astype(np.float64), p.v.astype(np.float64)
    dx1, dx2, f = p.dx1.astype(np.float64), p.dx2.astype(np.float64), p.force.astype(np.float64)
    p.set_times(0.2, 0.15, 0.2)
    fs.cola.kick(p, 0.25)
    v1 = ocola.kick(v0, f, dx1, dx2, cosmo, 0.15, 0.2, 0.25)
    assert np.abs(p.v - v1).max() < 1e-6 * (1 + np.abs(v1).max())
    assert p.times[1] == 0.25
    fs.cola.drift(p, 0.3)
    x1 = ocola.drift(x0, p.v.astype(np.float64), dx1, dx2, cosmo, 0.2, 0.25, 0.3)
    assert np.abs(p.x - x1).max() < 4e-6 * L
    assert p.times[0] == 0.3


def test_deferred_kick_and_drift_are_bitwise_kick_then_drift(fs, cosmo, eh):
    """fs.cola.kick and fs.cola.drift defer their updates: the next
    pm.force runs the kick and the drift inside its binning pass.  Whatever
    touches the particles in between (reading v or x, another pm.force, a
    copy, an explicit sync) applies them first, as one kick + drift pass or
    separately; every order ends bitwise equal."""
    _announce("deferred kick and drift: fused into the binning vs separate passes")
    n_p, nc, L = 16, 32, 64.0
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    out = {}
    for mode in ("fused", "read_v", "read_x", "sync", "force_between", "copy", "fields"):
        p = fs.lpt.init(n_p, L, 0.1, eh, 5, "cola")
        if mode == "fields":
            fs._fs._particles_device_fields(p._particles)   # updates run at once from here
        for a_vel, a_pos in ocola.schedule(0.1, 1.0, 4):
            fs.pm.force(p)
            fs.cola.kick(p, a_vel)
            if mode == "read_v":
                _ = p.v
            elif mode == "force_between":
                fs.pm.force(p)      # same positions: same forces
            elif mode == "copy":
                q = fs.Particles(len(p), L)
                q.copy_from(p)      # q owes p's pending kick
                fs.cola.drift(q, a_pos)
                p.copy_from(q)      # and p owes q's pending kick and drift
                continue
            fs.cola.drift(p, a_pos)
            if mode == "read_x":
                _ = p.x
            elif mode == "sync":
                fs._fs._particles_sync(p._particles)
        out[mode] = (p.x, p.v)
    for mode in out:
        for k in range(2):
            assert np.array_equal(out[mode][k], out["fused"][k]), (mode, k)


def test_leapfrog_kernels_match_oracle(fs, cosmo):
    """py/leapfrog.cpp:18-94: initial velocity, kick, drift."""
    from oracle import leapfrog as olf
    _announce("leapfrog kernels")
    rng = np.random.default_rng(4)
    N, L = 1000, 64.0
    p = fs.Particles(N, L)
    p.append(rng.uniform(0, L, (N, 3)))
    for name in ("dx1", "dx2", "force"):
        p.set_field(name, rng.standard_normal((N, 3)).astype(np.float32))
    dx1, dx2, f = p.dx1.astype(np.float64), p.dx2.astype(np.float64), p.force.astype(np.float64)
    fs.leapfrog.set_initial_velocity(p, 0.1)
    v0 = olf.initial_velocity(dx1, dx2, cosmo, 0.1)
    assert np.abs(p.v - v0).max() < 1e-6 * np.abs(v0).max()
    assert p.times[1] == 0.1
    p.set_times(0.1, 0.1, 0.1)
    x0, v0 = p.x.astype(np.float64), p.v.astype(np.float64)
    fs.leapfrog.kick(p, 0.15)
    v1 = olf.kick(v0, f, cosmo, 0.1, 0.15)
    assert np.abs(p.v - v1).max() < 1e-6 * (1 + np.abs(v1).max())
    fs.leapfrog.drift(p, 0.2)
    x1 = olf.drift(x0, p.v.astype(np.float64), cosmo, 0.1, 0.2)
    assert np.abs(p.x - x1).max() < 4e-6 * L
    assert p.times[:2] == (0.2, 0.15)


def test_leapfrog_pm_run_matches_oracle(fs, cosmo, eh):
    """5 standard leapfrog steps with PM forces from 2LPT ICs (v from
    set_initial_velocity): positions within 1e-3 mesh cells of FP64."""
    from oracle import leapfrog as olf
    _announce("leapfrog run 16^3/32^3")
    n_p, nc, L, a0, a1, nstep = 16, 32, 64.0, 0.1, 0.5, 5
    p = fs.lpt.init(n_p, L, a0, eh, 13, "2lpt")
    fs.leapfrog.set_initial_velocity(p, a0)
    dx1, dx2 = p.dx1.astype(np.float64), p.dx2.astype(np.float64)
    x = p.x.astype(np.float64)
    v = olf.initial_velocity(dx1, dx2, cosmo, a0)
    fs.pm.init(nc, nc / n_p, L)
    a_x = a_v = a0
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        fs.pm.force(p)
        fs.leapfrog.kick(p, a_vel)
        fs.leapfrog.drift(p, a_pos)
        xw = np.mod(x, L)
        f = opm.force(xw, nc, L, nc / n_p)
        v = olf.kick(v, f, cosmo, a_v, a_vel)
        a_v = a_vel
        x = olf.drift(xw, v, cosmo, a_x, a_pos)
        a_x = a_pos
    d = p.x.astype(np.float64) - x
    d -= L * np.round(d / L)
    assert np.abs(d).max() / (L / nc) < 1e-3, np.abs(d).max() / (L / nc)


def test_cola_run_matches_oracle(fs, cosmo, eh):
    """10 COLA steps from 2LPT ICs: final positions within 1e-3 mesh cells."""
    _announce("cola run 16^3/32^3")
    n_p, nc, L, a0, a1, nstep = 16, 32, 64.0, 0.1, 1.0, 10
    p = fs.lpt.init(n_p, L, a0, eh, 11, "cola")
    x = p.x.astype(np.float64)
    v = p.v.astype(np.float64)
    dx1 = p.dx1.astype(np.float64)
    dx2 = p.dx2.astype(np.float64)
    fs.pm.init(nc, nc / n_p, L)
    a_x = a_v = a0
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
        # oracle step in FP64 on its own trajectory
        xw = np.mod(x, L)
        f = opm.force(xw, nc, L, nc / n_p)
        v = ocola.kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, cosmo, a_x, a_v, a_pos)
        a_x = a_pos
    got = p.x.astype(np.float64)
    d = got - x
    d -= L * np.round(d / L)
    cell = L / nc
    assert np.abs(d).max() / cell < 1e-3, np.abs(d).max() / cell


def test_config_a_run_forces_and_power_match_fp64(fs, cosmo, eh):
    """The reference's CPU case: 64^3 particles, 128^3 mesh, box 64 Mpc/h,
    2LPT at a = 0.1, 10 COLA steps to a = 1, beside an FP64 oracle
    trajectory from the same 2LPT state.  Positions: RMS below 1e-4 cells,
    99.99% within 1e-3 cells, none beyond 1e-2 (the FP32 floor of dense
    regions, as at 256^3); forces at the final positions within 1e-5
    relative RMS of FP64; final P(k) within 0.1%."""
    _announce("config A: 64^3 p / 128^3 mesh, 10 steps vs FP64 oracle")
    n_p, nc, L = 64, 128, 64.0
    p = fs.lpt.init(n_p, L, 0.1, eh, 21, "cola")
    x, v = p.x.astype(np.float64), p.v.astype(np.float64)
    dx1, dx2 = p.dx1.astype(np.float64), p.dx2.astype(np.float64)
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    a_x = a_v = 0.1
    for a_vel, a_pos in ocola.schedule(0.1, 1.0, 10):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
        xw = np.mod(x, L)
        f = opm.force(xw, nc, L, nc / n_p)
        v = ocola.kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, cosmo, a_x, a_v, a_pos)
        a_x = a_pos
    d = p.x.astype(np.float64) - x
    d -= L * np.round(d / L)
    cell = L / nc
    e = np.abs(d).max(axis=1) / cell
    rms = float(np.sqrt(np.mean(d ** 2)) / cell)
    print(f"[case] config A position error (cells): rms {rms:.2e} "
          f"p99.99 {np.quantile(e, 0.9999):.2e} max {e.max():.2e}", flush=True)
    assert rms < 1e-4 and np.quantile(e, 0.9999) < 1e-3 and e.max() < 1e-2
    # forces at the GPU's final positions
    p.set_times(1.0, 1.0, 0.0)
    fs.pm.force(p)
    xg = p.x.astype(np.float64)
    ref = opm.force(xg, nc, L, nc / n_p)
    rel = float(np.sqrt(np.mean((p.force - ref) ** 2) / np.mean(ref ** 2)))
    print(f"[case] config A final force rel RMS = {rel:.3e}", flush=True)
    assert rel <= 1e-5, rel
    # P(k) of the final density against the oracle's trajectory
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    k, P, nm = fs.pm.power_spectrum(nc=nc)
    kr, Pr, nmr = opk.power_spectrum(opm.density(np.mod(x, L), nc, L, nc / n_p), L)
    ok = nmr > 0
    assert np.array_equal(nm, nmr)
    perr = float(np.abs(P[ok] / Pr[ok] - 1).max())
    print(f"[case] config A max P(k) relative error = {perr:.3e}", flush=True)
    assert perr < 1e-3, perr


def _cola_final_state(fs, eh, n_p, nc, L, seed=1):
    """The 10-step COLA run a = 0.1 -> 1 from 2LPT ICs (script/cola.py:37-44)
    with the deterministic paint, then one more force at the final
    positions (which also wraps them into the box)."""
    p = fs.lpt.init(n_p, L, 0.1, eh, seed, "cola")
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    for a_vel, a_pos in ocola.schedule(0.1, 1.0, 10):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
    fs.pm.force(p)
    return p


def _rel_rms(f, ref):
    return float(np.sqrt(np.mean((f - ref) ** 2) / np.mean(ref ** 2)))


def _max_cell_error(d32, ref):
    """max |d - ref| / max(1, |ref|) over the mesh, slab by slab."""
    worst = 0.0
    for i in range(0, d32.shape[0], 64):
        a = d32[i:i + 64].astype(np.float64)
        b = ref[i:i + 64]
        worst = max(worst, float((np.abs(a - b) / np.maximum(1.0, np.abs(b))).max()))
    return worst


@pytest.mark.timeout(600)
def test_config_b_forces_after_cola_run_match_fp64(fs, cosmo, eh):
    """256^3 particles, 512^3 mesh, box 512 Mpc/h: after the 10-step COLA run
    (a = 0.1 -> 1), the PM forces at the clustered final positions match the
    FP64 oracle (paint, FP64 FFTs, Green, readout) to relative RMS <= 1e-5."""
    _announce("config B forces vs FP64 oracle (256^3 p / 512^3 mesh)")
    p = _cola_final_state(fs, eh, 256, 512, 512.0)
    x, f = p.x.astype(np.float64), p.force.astype(np.float64)
    ref, _ = opm.force_large(x, 512, 512.0, 2.0, progress=_announce)
    rel = _rel_rms(f, ref)
    print(f"[case] config B force rel RMS = {rel:.3e}", flush=True)
    assert rel <= 1e-5, rel


@pytest.mark.timeout(800)
def test_config_c_forces_and_mesh_after_cola_run_match_fp64(fs, cosmo, eh):
    """The headline configuration: 512^3 particles, 1024^3 mesh, box 1024,
    10 COLA steps to the clustered state at a = 1.  At the final positions:
    - each paint variant (sorted, LDS-atomic, global) equals the FP64
      density per cell to 1e-5 of max(1, |delta|), and its FP64 total
      passes pm.cpp:376's bound 10 FLT_EPSILON nc^3 (np 8 - nc^3 = 0);
    - the forces match the FP64 oracle to relative RMS <= 1e-5;
    - the forces sum to zero (momentum conservation of symmetric CIC paint
      and readout with an odd Green function) as the oracle's do."""
    n_p, nc, L = 512, 1024, 1024.0
    _announce("config C: 10-step COLA run (512^3 p / 1024^3 mesh)")
    p = _cola_final_state(fs, eh, n_p, nc, L)
    x, f = p.x.astype(np.float64), p.force.astype(np.float64)
    _announce("config C: FP64 oracle density")
    delta = ocfast.paint(x, nc, L, 2.0)
    tol = 10 * np.finfo(np.float32).eps * nc**3
    ref_total = float(delta.sum())
    assert abs(ref_total) < 1e-6 * tol
    for variant in ("sorted", "atomic", "global"):
        fs.pm.set_paint_variant(variant)
        fs.pm.send_positions(p)
        g = fs.pm.compute_density(p)
        total = fs.pm.total_density()
        err = _max_cell_error(g.asarray(), delta)
        print(f"[case] config C paint {variant}: max cell error {err:.2e}, "
              f"FP64 sum of delta {total:.3e} (bound {tol:.3e})", flush=True)
        assert err < 1e-5, (variant, err)
        assert abs(total) < tol, (variant, total)
    fs.pm.set_paint_variant("sorted")
    _announce("config C: FP64 oracle forces")
    ref, _ = opm.force_large(x, nc, L, 2.0, progress=_announce, delta=delta)
    del delta
    rel = _rel_rms(f, ref)
    rms = np.sqrt(np.mean(ref ** 2))
    mom = np.abs(f.sum(axis=0)).max() / (len(f) * rms)
    mom_ref = np.abs(ref.sum(axis=0)).max() / (len(f) * rms)
    print(f"[case] config C force rel RMS = {rel:.3e}; |sum f| / (N rms) = {mom:.2e} "
          f"(oracle {mom_ref:.2e})", flush=True)
    assert rel <= 1e-5, rel
    assert mom < 1e-6 and mom_ref < 1e-9, (mom, mom_ref)


@pytest.mark.timeout(600)
def test_config_b_trajectory_and_power_match_fp64(fs, cosmo, eh):
    """256^3 particles, 512^3 mesh, box 512: the 10-step COLA run beside an
    FP64 oracle trajectory from the same 2LPT state, and the final P(k)
    within 0.1%.  Positions: RMS below 1e-4 cells, 99.99% of particles
    within 1e-3 cells, fewer than 1e-5 of them beyond, none beyond 1e-2.  A
    handful of particles in dense regions amplify FP32 rounding: two FP32
    runs of this state that differ only in the paint's summation rounding
    (int64 fixed point vs FP32 atomics) also end up to 1.5e-3 cells apart,
    so a per-particle bound of 1e-3 cells is below the FP32 floor here."""
    _announce("config B trajectory + P(k) vs FP64 oracle")
    n_p, nc, L = 256, 512, 512.0
    p = fs.lpt.init(n_p, L, 0.1, eh, 11, "cola")
    x, v = p.x.astype(np.float64), p.v.astype(np.float64)
    dx1, dx2 = p.dx1.astype(np.float64), p.dx2.astype(np.float64)
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    a_x = a_v = 0.1
    for step, (a_vel, a_pos) in enumerate(ocola.schedule(0.1, 1.0, 10)):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
        xw = np.mod(x, L)
        f, _ = opm.force_large(xw, nc, L, nc / n_p)
        v = ocola.kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, cosmo, a_x, a_v, a_pos)
        a_x = a_pos
        _announce(f"step {step} done")
    d = p.x.astype(np.float64) - x
    d -= L * np.round(d / L)
    cell = L / nc
    e = np.abs(d).max(axis=1) / cell
    rms = float(np.sqrt(np.mean(d ** 2)) / cell)
    p9999 = float(np.quantile(e, 0.9999))
    frac = float(np.mean(e > 1e-3))
    print(f"[case] config B position error (cells): rms {rms:.2e} p99.99 {p9999:.2e} "
          f"max {e.max():.2e} fraction > 1e-3 {frac:.1e}", flush=True)
    assert rms < 1e-4 and p9999 < 1e-3 and frac < 1e-5 and e.max() < 1e-2
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    k, P, nm = fs.pm.power_spectrum(nc=nc)
    kr, Pr, nmr = opk.power_spectrum(ocfast.paint(np.mod(x, L), nc, L, nc / n_p), L)
    ok = nmr > 0
    assert np.array_equal(nm, nmr)
    perr = float(np.abs(P[ok] / Pr[ok] - 1).max())
    print(f"[case] config B max P(k) relative error = {perr:.3e}", flush=True)
    assert perr < 1e-3, perr


def test_power_spectrum_matches_oracle(fs, cosmo, eh):
    _announce("power spectrum")
    n_p, nc, L = 16, 32, 64.0
    p = fs.lpt.init(n_p, L, 0.5, eh, 5, "cola")
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.send_positions(p)
    g = fs.pm.compute_density(p)
    delta = g.asarray().astype(np.float64)
    k, P, nm = fs.pm.power_spectrum(nc=nc)
    kr, Pr, nmr = opk.power_spectrum(delta, L)
    assert np.array_equal(nm, nmr)
    ok = nmr > 0
    assert np.allclose(k[ok], kr[ok], rtol=1e-12)
    assert np.allclose(P[ok], Pr[ok], rtol=1e-3)   # P(k) within 0.1%
