"""Mesh sizes beyond the tuned transforms (FFTW, py/fft.cpp:19-99, takes any
n; py/pm.cpp:245-255 imposes no size rule): every even n whose prime
factors are 2, 3, 5, 7 runs through runtime-planned mixed-radix passes.
Transforms against numpy at generic sizes, a separable closed form and the
round trip at 768^3 and 2048^3, PM forces, 2LPT and a COLA run against the
FP64 oracle at generic sizes, the 2048^3 PM mesh (64-bit offsets), and the
error for sizes outside the rule."""
import numpy as np
import pytest

from oracle import cola as ocola
from oracle import cosmology as ocosmo
from oracle import lpt as olpt
from oracle import pm as opm
from oracle import power as opower
from oracle import rng as orng

pytestmark = pytest.mark.gpu

GENERIC = [6, 10, 12, 20, 24, 30, 40, 56, 60, 80, 100, 120, 144, 160, 192, 210, 384]
OM = 0.31
EH = (OM, 0.049, 0.67, 0.96, 0.81)


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def eh(fs):
    fs.cosmology.init(OM)
    return fs.PowerSpectrum.eisenstein_hu(*EH)


@pytest.mark.parametrize("n", GENERIC)
def test_generic_r2c_and_c2r_match_numpy(fs, n):
    _announce(f"generic r2c/c2r n={n}")
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n)).astype(np.float32)
    g = fs.FFT(n)
    g.set_array(a)
    g.execute_forward()
    got = g.asarray_k()
    ref = np.fft.rfftn(a.astype(np.float64))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 4e-6, ("r2c", err)
    ck = (rng.standard_normal((n, n, n // 2 + 1)) +
          1j * rng.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
    g.set_array_k(ck)
    g.execute_inverse()
    got = g.asarray()
    ref = np.fft.irfftn(ck.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2)) * n**3
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 4e-6, ("c2r", err)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [768, 2048])
def test_generic_large_separable_and_round_trip(fs, n):
    """Separable input a(x) b(y) c(z) (two terms): its r2c is the sum of the
    products of 1D transforms, checked on 12 x-slabs; then c2r(r2c(a)) =
    n^3 a everywhere.  At 2048^3 a mesh holds 8.7e9 floats."""
    _announce(f"generic separable + round trip n={n}")
    rng = np.random.default_rng(11 * n)
    terms = [[rng.standard_normal(n) for _ in range(3)] for _ in range(2)]
    g = fs.FFT(n)
    a = np.empty((n, n, n), dtype=np.float32)
    for i in range(n):
        a[i] = sum((ta[i] * tb[:, None] * tc[None, :]).astype(np.float32) for ta, tb, tc in terms)
    g.set_array(a)
    g.execute_forward()
    got = g.asarray_k()
    spec = [(np.fft.fft(ta), np.fft.fft(tb), np.fft.rfft(tc)) for ta, tb, tc in terms]
    scale = max(np.abs(A).max() * np.abs(B).max() * np.abs(C).max() for A, B, C in spec)
    err = 0.0
    for kx in list(rng.integers(0, n, 10)) + [0, n // 2]:
        ref = sum(A[kx] * B[:, None] * C[None, :] for A, B, C in spec)
        err = max(err, float(np.abs(got[kx] - ref).max()))
    del got
    assert err / scale < 6e-6, err / scale
    g.execute_inverse()
    back = g.asarray()
    amax = float(np.abs(a).max())
    worst = 0.0
    inv = 1.0 / float(n) ** 3
    for i in range(0, n, 64):
        worst = max(worst, float(np.abs(back[i:i + 64] * inv - a[i:i + 64]).max()))
    _announce(f"round trip n={n}: max error {worst / amax:.2e} of max")
    assert worst / amax < 4e-6


@pytest.mark.parametrize("n_p,nc,L", [(30, 60, 60.0), (50, 100, 100.0), (96, 192, 192.0),
                                      (192, 384, 384.0)])
def test_generic_pm_force_matches_fp64(fs, eh, n_p, nc, L):
    """2LPT state at a = 0.6 (n_p^3 particles; n_p = 30 and 50 also run the
    generic 2LPT transforms), PM force on the generic nc^3 mesh against the
    FP64 oracle: relative RMS <= 1e-5."""
    _announce(f"generic PM force {n_p}^3 / {nc}^3")
    p = fs.lpt.init(n_p, L, 0.6, eh, 3, "cola")
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.force(p)
    ref, total = opm.force_large(p.x.astype(np.float64), nc, L, nc / n_p)
    f = p.force.astype(np.float64)
    rel = float(np.sqrt(np.mean((f - ref) ** 2) / np.mean(ref ** 2)))
    print(f"[case] generic PM {n_p}^3/{nc}^3 force rel RMS = {rel:.2e}", flush=True)
    assert rel <= 1e-5, rel


@pytest.mark.parametrize("n", [24, 30, 50])
def test_generic_lpt_matches_oracle(fs, eh, n):
    _announce(f"generic 2LPT n={n}")
    L, a = 2.0 * n, 0.3
    c = ocosmo.Cosmology(OM)
    p = fs.lpt.init(n, L, a, eh, 9, "2lpt")
    ref = olpt.particles(orng.white_noise(n, 9), L, a, opower.EH98NoWiggle(*EH).P, c, "2lpt")
    assert np.array_equal(p.id, ref["id"])
    for name in ("dx1", "dx2", "x", "v"):
        got = getattr(p, name).astype(np.float64)
        scale = np.sqrt(np.mean(ref["dx1" if name == "x" else name] ** 2))
        err = np.abs(got - ref[name]).max() / scale
        assert err < 2e-5, (name, err)


def test_generic_cola_run_matches_fp64_trajectory(fs, eh):
    """10 COLA steps a = 0.1 -> 1 at 30^3 particles on a 60^3 mesh (box 60)
    beside the FP64 oracle trajectory from the same 2LPT state."""
    _announce("generic COLA run 30^3 / 60^3")
    n_p, nc, L = 30, 60, 60.0
    c = ocosmo.Cosmology(OM)
    p = fs.lpt.init(n_p, L, 0.1, eh, 5, "cola")
    x, v, dx1, dx2 = (getattr(p, k).astype(np.float64) for k in ("x", "v", "dx1", "dx2"))
    fs.pm.init(nc, nc / n_p, L)
    a_x = a_v = 0.1
    for a_vel, a_pos in ocola.schedule(0.1, 1.0, 10):
        fs.pm.force(p)
        fs.cola.kick(p, a_vel)
        fs.cola.drift(p, a_pos)
        xw = np.mod(x, L)
        f, _ = opm.force_large(xw, nc, L, nc / n_p)
        v = ocola.kick(v, f, dx1, dx2, c, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, c, a_x, a_v, a_pos)
        a_x = a_pos
    d = p.x.astype(np.float64) - x
    d -= L * np.round(d / L)
    e = np.abs(d).max(axis=1) / (L / nc)
    rms = float(np.sqrt(np.mean(d ** 2)) / (L / nc))
    print(f"[case] generic run: rms {rms:.2e} max {e.max():.2e} cells", flush=True)
    assert rms < 1e-4 and e.max() < 1e-2


@pytest.mark.parametrize("n", [22, 15, 26, 2])
def test_unsupported_sizes_raise_value_error(fs, n):
    with pytest.raises(ValueError, match="not supported"):
        fs.FFT(n)
    with pytest.raises(ValueError):
        fs.pm.init(n, 1.0, 10.0)


def _sparse_cic(x, nc, L, fac):
    """CIC of few particles on a huge mesh: (flat cell index, FP64 value)."""
    i0, i1, w0, w1 = _cic_cells(x, nc, L)
    idx, val = [], []
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                ix = (i0, i1)[cx][:, 0]
                iy = (i0, i1)[cy][:, 1]
                iz = (i0, i1)[cz][:, 2]
                idx.append((ix * nc + iy) * nc + iz)
                val.append(fac * (w0, w1)[cx][:, 0] * (w0, w1)[cy][:, 1] * (w0, w1)[cz][:, 2])
    idx, val = np.concatenate(idx), np.concatenate(val)
    u, inv = np.unique(idx, return_inverse=True)
    return u, np.bincount(inv, weights=val)


def _cic_cells(x, nc, L):
    X = x * (nc / L)
    i0 = np.trunc(X).astype(np.int64)
    w1 = X - i0
    w0 = 1.0 - w1
    i0 = np.where(i0 >= nc, 0, i0)
    i1 = np.where(i0 + 1 >= nc, i0 + 1 - nc, i0 + 1)
    return i0, i1, w0, w1


@pytest.mark.timeout(900)
def test_2048_mesh_paint_and_readout(fs, eh):
    """A 2048^3 PM mesh (8.7e9 floats per mesh; offsets past 2^31): the
    deterministic and global paints of 64^3 particles equal the sparse FP64
    CIC at every touched cell and -1 elsewhere, and the tiled readout of a
    force equals the per-particle gather readout bit for bit."""
    n_p, nc, L = 64, 2048, 2048.0
    _announce("2048^3 mesh: paint vs sparse FP64, tiled vs gather readout")
    p = fs.lpt.init(n_p, L, 0.5, eh, 13, "cola")
    fs.pm.init(nc, nc / n_p, L)
    fac = (nc / n_p) ** 3
    fs.pm.send_positions(p)
    x = p.x.astype(np.float64)
    u, val = _sparse_cic(x, nc, L, fac)
    for variant in ("sorted", "global"):
        fs.pm.set_paint_variant(variant)
        g = fs.pm.compute_density(p)
        d = g.asarray().reshape(-1)
        got = d[u].astype(np.float64) + 1.0
        err = float(np.abs(got - val).max() / val.max())
        # all of the mass sits in the touched cells: every other cell is -1
        total = float(np.sum(d, dtype=np.float64)) + float(nc) ** 3
        print(f"[case] 2048^3 paint {variant}: max rel error {err:.2e}; mass {total:.6e} "
              f"(touched cells {got.sum():.6e}, expected {n_p**3 * fac:.6e})", flush=True)
        assert err < 1e-5
        assert abs(total - n_p**3 * fac) < 1e-6 * n_p**3 * fac
        assert abs(got.sum() - n_p**3 * fac) < 1e-6 * n_p**3 * fac
        del d, g
    fs.pm.set_paint_variant("sorted")
    fs.pm.force(p)   # tiled readout from the paint's bins
    f_tiled = p.force.copy()
    p.device_fields()   # device planes exposed: the next readout gathers per particle
    p.set_times(*p.times[:2], 0.0)
    fs.pm.force(p)
    assert np.array_equal(p.force, f_tiled)
    assert np.isfinite(f_tiled).all() and np.abs(f_tiled).max() > 0
