"""GPU parity at the sizes of the batched (D) and differentiable (E)
configurations, and the reference's own driver cases.

D  96 boxes in one graph-replayed launch chain: 32 x (32^3 particles, 64^3
   mesh, box 32), 32 x (48^3, 96^3, 48), 32 x (64^3, 128^3, 64), 10 steps.
   Every box equals the single-box fs loop bit for bit; one box per class
   follows the FP64 oracle trajectory, and its final forces match FP64.
E  128^3 particles, 256^3 mesh, box 256, 10 steps, loss = mean squared
   difference of the final density and a second seed's.  The GPU gradient
   with respect to the white noise against the float64 autograd oracle, and
   against central finite differences on 16 random directions.
script/cola.py:18-44 (Omega_m 0.308, nc = nc_pm = 64, box 64, the Planck
   CAMB table, 9 steps) and test/pm_setup.py:13-27 (2LPT at a = 1, density
   and forces on a 64^3 mesh) beside the FP64 oracle.
"""
import os

import numpy as np
import pytest

from oracle import cfast as ocfast
from oracle import cola as ocola
from oracle import cosmology as ocosmo
from oracle import pk as opk
from oracle import pm as opm
from oracle import power as opower

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OM = 0.31
EH = (OM, 0.049, 0.67, 0.96, 0.81)


def _announce(msg):
    print(f"[case] {msg}", flush=True)


def _run_single(fs, ps, n_p, nc, L, seed, a0, a1, nstep):
    p = fs.lpt.init(n_p, L, a0, ps, seed, "cola")
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
    return p


def _oracle_trajectory(cosmo, x, v, dx1, dx2, nc, L, pm_factor, a0, a1, nstep):
    a_x = a_v = a0
    for a_vel, a_pos in ocola.schedule(a0, a1, nstep):
        xw = np.mod(x, L)
        f, _ = opm.force_large(xw, nc, L, pm_factor)
        v = ocola.kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, cosmo, a_x, a_v, a_pos)
        a_x = a_pos
    return x


def _position_errors(x_gpu, x_ref, L, nc):
    d = x_gpu.astype(np.float64) - x_ref
    d -= L * np.round(d / L)
    cell = L / nc
    e = np.abs(d).max(axis=1) / cell
    return float(np.sqrt(np.mean(d ** 2)) / cell), float(np.quantile(e, 0.9999)), float(e.max())


def _rel_rms(f, ref):
    return float(np.sqrt(np.mean((f - ref) ** 2) / np.mean(ref ** 2)))


@pytest.fixture(scope="module")
def eh(fs):
    fs.cosmology.init(OM)
    return fs.PowerSpectrum.eisenstein_hu(*EH)


@pytest.mark.timeout(600)
def test_config_d_ensemble_full_size(fs, eh):
    from fs2_amd.batch import Ensemble
    classes = [(32, 64, 32, 32.0), (48, 96, 32, 48.0), (64, 128, 32, 64.0)]
    a0, a1, nstep = 0.1, 1.0, 10
    _announce("config D: 96-box ensemble, autotune + graph replay")
    fs.cosmology.init(OM)
    cosmo = ocosmo.Cosmology(OM)
    ens = Ensemble(classes)
    seeds = list(range(1, 97))
    ens.init_lpt(eh, a0, seeds)
    ens.autotune(reps=1)
    ens.set_schedule(a0, a1, nstep)
    ens.run(graph=True)
    ens.synchronize()
    k = 0
    for ci, (nl, nm, nb, L) in enumerate(classes):
        for box in range(nb):
            p = _run_single(fs, eh, nl, nm, L, seeds[k], a0, a1, nstep)
            for name in ("x", "v"):
                assert np.array_equal(ens.field(ci, box, name), getattr(p, name)), (ci, box, name)
            k += 1
        _announce(f"config D class {ci}: {nb} boxes bitwise equal to single-box runs")
    # one box per class against the FP64 oracle
    k = 0
    for ci, (nl, nm, nb, L) in enumerate(classes):
        box = nb // 2
        seed = seeds[k + box]
        k += nb
        p0 = fs.lpt.init(nl, L, a0, eh, seed, "cola")
        x_ref = _oracle_trajectory(cosmo, *(getattr(p0, n).astype(np.float64)
                                            for n in ("x", "v", "dx1", "dx2")),
                                   nm, L, nm / nl, a0, a1, nstep)
        x_gpu = ens.field(ci, box, "x")
        rms, p9999, emax = _position_errors(x_gpu, x_ref, L, nm)
        p = _run_single(fs, eh, nl, nm, L, seed, a0, a1, nstep)
        fs.pm.force(p)
        ref, _ = opm.force_large(p.x.astype(np.float64), nm, L, nm / nl)
        rel = _rel_rms(p.force.astype(np.float64), ref)
        print(f"[case] config D class {ci} box {box}: positions rms {rms:.2e} p99.99 "
              f"{p9999:.2e} max {emax:.2e} cells; final force rel RMS {rel:.2e}", flush=True)
        assert rms < 1e-4 and p9999 < 1e-3 and emax < 1e-2
        assert rel <= 1e-5


@pytest.mark.timeout(600)
def test_config_e_gradient_full_size(fs, eh):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from fs2_amd.autograd import ColaModel
    from oracle import diff
    n, nc, L, nstep = 128, 256, 256.0, 10
    _announce("config E: GPU forward/backward 128^3 / 256^3, 10 steps")
    fs.cosmology.init(OM)
    cosmo = ocosmo.Cosmology(OM)
    model = ColaModel(n, nc, L, 0.1, 1.0, nstep, eh)
    w1 = np.random.default_rng(1).standard_normal((n,) * 3).astype(np.float32)
    w2 = np.random.default_rng(2).standard_normal((n,) * 3).astype(np.float32)
    target = model.forward(torch.tensor(w2, device="cuda")).clone()
    wd = torch.tensor(w1, device="cuda")
    delta = model.forward(wd)
    dbar = 2.0 * (delta - target) / delta.numel()
    g = model.backward(dbar).cpu().numpy().astype(np.float64)

    def loss(w):
        with torch.no_grad():
            return float(((model.forward(w) - target).double() ** 2).mean())

    # central differences of the FP32 forward on 16 random unit directions
    rng = np.random.default_rng(3)
    dirs = [rng.standard_normal((n,) * 3) for _ in range(16)]
    dirs = [d / np.linalg.norm(d) for d in dirs]
    eps = 0.5
    fd_err = []
    for d in dirs:
        dd = torch.tensor(d.astype(np.float32), device="cuda")
        fd = (loss(wd + eps * dd) - loss(wd - eps * dd)) / (2 * eps)
        an = float(np.sum(g * d))
        fd_err.append(abs(fd - an) / np.linalg.norm(g))
    _announce("config E: float64 autograd oracle")
    _, g_ref, d_ref = diff.loss_and_grad(w1.astype(np.float64),
                                         target.cpu().numpy().astype(np.float64), L, nc,
                                         0.1, 1.0, nstep, opower.EH98NoWiggle(*EH).P, cosmo)
    drel = float(np.abs(delta.cpu().numpy() - d_ref).max() / np.abs(d_ref).max())
    rel = float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
    dir_err = max(abs(np.sum(g * d) - np.sum(g_ref * d)) / np.linalg.norm(g_ref) for d in dirs)
    print(f"[case] config E: density max dev {drel:.2e} of max; gradient rel L2 vs float64 "
          f"{rel:.2e}; directional vs float64 {dir_err:.2e} |g|; FD vs adjoint "
          f"{max(fd_err):.2e} |g|", flush=True)
    assert drel < 1e-3
    assert rel < 2e-2, rel
    assert dir_err < 5e-3, dir_err
    assert max(fd_err) < 5e-3, fd_err


def test_reference_driver_case_script_cola(fs):
    """script/cola.py:18-44: Omega_m 0.308, nc = nc_pm = 64 (pm_factor 1),
    box 64, data/planck_matterpower.dat, 2LPT 'cola' ICs at a = 0.1, 9 steps
    to a = 1.  The IC realisation is this build's seeded white noise (the
    reference draws GSL ranlxd1); from that state the run follows the FP64
    oracle.  Also checks the table's P(k) reaches the 2LPT amplitude
    (power.cpp:50-54 -> lpt.cpp:364-366) through the oracle's 2LPT."""
    from oracle import lpt as olpt
    from oracle import rng as orng
    _announce("script/cola.py case: Om 0.308, 64^3 / 64^3, Planck table, 9 steps")
    om, n, L, nstep = 0.308, 64, 64.0, 9
    fs.cosmology.init(om)
    cosmo = ocosmo.Cosmology(om)
    path = os.path.join(GOLDEN, "planck_matterpower.dat")
    ps = fs.PowerSpectrum(path)
    table = opower.TablePower(path)
    p = fs.lpt.init(n, L, 0.1, ps, 1, "cola")
    ref0 = olpt.particles(orng.white_noise(n, 1), L, 0.1, table.P, cosmo, "cola")
    for name in ("dx1", "dx2", "x"):
        got = getattr(p, name).astype(np.float64)
        want = ref0[name]
        scale = np.sqrt(np.mean(ref0["dx1" if name == "x" else name] ** 2))
        err = np.abs(got - want).max() / scale
        assert err < 2e-5, (name, err)
    x0, v0, dx1, dx2 = (getattr(p, k).astype(np.float64) for k in ("x", "v", "dx1", "dx2"))
    fs.pm.init(n, 1.0, L)
    fs.pm.set_paint_variant("sorted")
    for a_vel, a_pos in ocola.schedule(0.1, 1.0, nstep):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
    x_ref = _oracle_trajectory(cosmo, x0, v0, dx1, dx2, n, L, 1.0, 0.1, 1.0, nstep)
    rms, p9999, emax = _position_errors(p.x, x_ref, L, n)
    fs.pm.force(p)
    ref, _ = opm.force_large(p.x.astype(np.float64), n, L, 1.0)
    rel = _rel_rms(p.force.astype(np.float64), ref)
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    k, P, nm = fs.pm.power_spectrum(nc=n)
    kr, Pr, nmr = opk.power_spectrum(ocfast.paint(np.mod(x_ref, L), n, L, 1.0), L)
    ok = nmr > 0
    perr = float(np.abs(P[ok] / Pr[ok] - 1).max())
    print(f"[case] script/cola.py case: positions rms {rms:.2e} p99.99 {p9999:.2e} max "
          f"{emax:.2e} cells; final force rel RMS {rel:.2e}; P(k) max rel {perr:.2e}",
          flush=True)
    assert rms < 1e-4 and p9999 < 1e-3 and emax < 1e-2
    assert rel <= 1e-5
    assert np.array_equal(nm, nmr) and perr < 1e-3
    fs.cosmology.init(OM)


def test_reference_pm_setup_case(fs):
    """test/pm_setup.py:13-27 with test_pm_density.py / test_pm_force.py:
    Omega_m 0.308, nc = 64, pm_factor 1, box 64, 2LPT at a = 1.  The FP64
    sum of delta is within eps Ng (test_pm_density.py:26), the density
    matches FP64 per cell, and the forces match FP64 to 1e-5 rel. RMS."""
    _announce("test/pm_setup.py case: 2LPT at a = 1, 64^3 mesh, pm_factor 1")
    fs.cosmology.init(0.308)
    ps = fs.PowerSpectrum(os.path.join(GOLDEN, "planck_matterpower.dat"))
    n, L = 64, 64.0
    p = fs.lpt.init(n, L, 1.0, ps, 1, "2lpt")
    fs.pm.init(n, 1.0, L)
    fs.pm.send_positions(p)
    g = fs.pm.compute_density(p)
    d = g.asarray().astype(np.float64)
    assert abs(fs.pm.total_density()) < np.finfo(np.float32).eps * n**3
    x = p.x.astype(np.float64)
    ref = ocfast.paint(x, n, L, 1.0)
    assert np.abs(d - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
    fs.pm.force(p)
    fref, _ = opm.force_large(p.x.astype(np.float64), n, L, 1.0)
    rel = _rel_rms(p.force.astype(np.float64), fref)
    print(f"[case] pm_setup case: force rel RMS {rel:.2e}", flush=True)
    assert rel <= 1e-5
    fs.cosmology.init(OM)
