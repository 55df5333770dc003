"""GPU parity of the differentiable COLA run (fs2_diff_* through ctypes and
the torch.autograd.Function) against the float64 torch autograd oracle
(oracle/diff.py): forward density, gradient with respect to the white
noise, and a directional finite-difference check of the GPU gradient."""
import numpy as np
import pytest

from oracle import cosmology as ocosmo
from oracle import power as opower

pytestmark = pytest.mark.gpu

OM = 0.31
EH = (OM, 0.049, 0.67, 0.96, 0.81)


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.fixture(scope="module")
def setup(fs, torch):
    fs.cosmology.init(OM)
    return fs.PowerSpectrum.eisenstein_hu(*EH), ocosmo.Cosmology(OM), opower.EH98NoWiggle(*EH)


def _noise(n, seed):
    return np.random.default_rng(seed).standard_normal((n, n, n)).astype(np.float32)


CASES = [
    # nc, nc_pm, boxsize, nstep: mildly nonlinear (16 Mpc/h spacing) and
    # strongly clustered (4 Mpc/h spacing) final states
    (8, 16, 128.0, 3),
    (16, 32, 256.0, 4),
    (8, 8, 128.0, 2),
    (8, 16, 128.0, 0),
    (8, 16, 32.0, 3),
    # generic mixed-radix transforms (12 = 4 x 3, 24 = 8 x 3 outside the tuned sizes)
    (12, 24, 192.0, 3),
]


@pytest.mark.parametrize("variant", ["sorted", "global"])
@pytest.mark.parametrize("nc,nc_pm,boxsize,nstep", CASES)
def test_forward_and_gradient_match_float64_autograd(setup, torch, nc, nc_pm, boxsize, nstep,
                                                     variant):
    from fs2_amd.autograd import ColaModel
    from oracle import diff
    ps, cosmo, eh = setup
    _announce(f"adjoint nc={nc} nc_pm={nc_pm} L={boxsize} nstep={nstep} {variant}")
    a0, a1 = 0.1, 1.0
    w = _noise(nc, 11)
    target = (_noise(nc_pm, 12) * 0.5).astype(np.float64)
    model = ColaModel(nc, nc_pm, boxsize, a0, a1, nstep, ps)
    model.set_paint_variant(variant)
    wd = torch.tensor(w, device="cuda")
    delta = model.forward(wd)
    loss_ref, grad_ref, delta_ref = diff.loss_and_grad(w.astype(np.float64), target, boxsize,
                                                       nc_pm, a0, a1, nstep, eh.P, cosmo)
    d = delta.cpu().numpy().astype(np.float64)
    scale = max(1.0, np.abs(delta_ref).max())
    # float32 positions after nstep PM steps: density within 1e-3 of the scale
    assert np.abs(d - delta_ref).max() < 1e-3 * scale

    dbar = 2.0 * (delta - torch.tensor(target, device="cuda", dtype=torch.float32)) / d.size
    g = model.backward(dbar).cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g - grad_ref) / np.linalg.norm(grad_ref)
    assert rel < 2e-3, rel
    # and component-wise on the large entries
    big = np.abs(grad_ref) > 0.1 * np.abs(grad_ref).max()
    assert np.abs(g - grad_ref)[big].max() < 1e-2 * np.abs(grad_ref).max()


def test_autograd_function_and_finite_differences(setup, torch):
    from fs2_amd.autograd import ColaModel, cola_density
    ps, _, _ = setup
    _announce("adjoint autograd.Function + FD")
    nc, nc_pm, L, nstep = 16, 32, 256.0, 3
    model = ColaModel(nc, nc_pm, L, 0.1, 1.0, nstep, ps)
    w = torch.tensor(_noise(nc, 21), device="cuda", requires_grad=True)
    target = torch.tensor(_noise(nc_pm, 22) * 0.5, device="cuda")
    loss = ((cola_density(w, model) - target) ** 2).mean()
    loss.backward()
    g = w.grad.detach().clone()
    # same as calling the adjoint directly
    with torch.no_grad():
        delta = model.forward(w)
        g2 = model.backward(2.0 * (delta - target) / delta.numel())
    # fixed-point tiled paints: the whole sweep is bitwise reproducible
    assert torch.equal(g, g2)

    # directional derivative by central differences of the float32 forward
    dirn = torch.tensor(np.random.default_rng(23).standard_normal((nc,) * 3).astype(np.float32),
                        device="cuda")
    dirn /= dirn.norm()

    def f(x):
        with torch.no_grad():
            return float(((model.forward(x) - target).double() ** 2).mean())
    eps = 3e-2
    fd = (f(w.detach() + eps * dirn) - f(w.detach() - eps * dirn)) / (2 * eps)
    an = float((g.double() * dirn.double()).sum())
    assert abs(fd - an) < 2e-2 * abs(an), (fd, an)


def test_backward_requires_matching_forward(setup, torch):
    from fs2_amd.autograd import ColaModel, cola_density
    ps, _, _ = setup
    _announce("adjoint error paths")
    model = ColaModel(8, 16, 32.0, 0.1, 1.0, 2, ps)
    with pytest.raises(RuntimeError, match="forward"):
        model.backward(torch.zeros((16,) * 3, device="cuda"))
    w = torch.tensor(_noise(8, 1), device="cuda", requires_grad=True)
    out = cola_density(w, model).sum()
    model.forward(w.detach())   # replaces the checkpoints of the graph above
    with pytest.raises(RuntimeError, match="checkpoints"):
        out.backward()
    with pytest.raises(ValueError):
        model.forward(torch.zeros((4,) * 3, device="cuda"))


def test_store_and_recompute_checkpoints_agree_bitwise(setup, torch):
    """Keeping every step's force meshes and bins, or rebuilding them from
    the positions, gives the same gradient bit for bit (deterministic
    paint), and a second backward of the same forward repeats it."""
    from fs2_amd.autograd import ColaModel
    ps, _, _ = setup
    _announce("adjoint store vs recompute")
    nc, nc_pm, L, nstep = 16, 32, 128.0, 5
    w = torch.tensor(_noise(nc, 31), device="cuda")
    dbar = torch.tensor(_noise(nc_pm, 32) * 1e-3, device="cuda")
    out = {}
    for policy in ("store", "recompute"):
        model = ColaModel(nc, nc_pm, L, 0.1, 1.0, nstep, ps, checkpoint=policy)
        assert model.stores_forces == (policy == "store")
        delta = model.forward(w)
        g1 = model.backward(dbar)
        g2 = model.backward(dbar)
        assert torch.equal(g1, g2)
        out[policy] = (delta, g1)
    assert torch.equal(out["store"][0], out["recompute"][0])
    assert torch.equal(out["store"][1], out["recompute"][1])


def test_gradient_directional_derivatives_float64(setup, torch):
    """<grad, d> of the GPU adjoint against the float64 oracle's directional
    derivative (exact autograd) on random directions."""
    from fs2_amd.autograd import ColaModel
    from oracle import diff
    ps, cosmo, eh = setup
    _announce("adjoint directional derivatives")
    nc, nc_pm, L, nstep = 16, 32, 256.0, 5
    w = _noise(nc, 41)
    target = (_noise(nc_pm, 42) * 0.5).astype(np.float64)
    _, grad_ref, _ = diff.loss_and_grad(w.astype(np.float64), target, L, nc_pm, 0.1, 1.0, nstep,
                                        eh.P, cosmo)
    model = ColaModel(nc, nc_pm, L, 0.1, 1.0, nstep, ps)
    delta = model.forward(torch.tensor(w, device="cuda"))
    dbar = 2.0 * (delta - torch.tensor(target, device="cuda", dtype=torch.float32)) / delta.numel()
    g = model.backward(dbar).cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(43)
    for _ in range(16):
        d = rng.standard_normal(w.shape)
        d /= np.linalg.norm(d)
        ref = float(np.sum(grad_ref * d))
        got = float(np.sum(g * d))
        assert abs(got - ref) < 5e-3 * np.linalg.norm(grad_ref), (got, ref)
