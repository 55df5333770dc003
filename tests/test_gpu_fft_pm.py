"""GPU parity: FFT passes, CIC paint/readout and the PM force against the
FP64 oracle (oracle/), through the fs._fs module and the C ABI."""
import numpy as np
import pytest

from oracle import cic as ocic
from oracle import pm as opm

pytestmark = pytest.mark.gpu

FFT_SIZES = [4, 8, 16, 32, 48, 64, 96, 128]


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("n", FFT_SIZES)
def test_r2c_matches_numpy(fs, n):
    _announce(f"r2c n={n}")
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n)).astype(np.float32)
    g = fs.FFT(n)
    g.set_array(a)
    g.execute_forward()
    got = g.asarray_k()
    ref = np.fft.rfftn(a.astype(np.float64))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    # FP32 Stockham: error ~ eps log2(N^3)
    assert err < 2e-6, err


@pytest.mark.parametrize("n", FFT_SIZES)
def test_c2r_non_hermitian_matches_irfftn(fs, n):
    _announce(f"c2r n={n}")
    rng = np.random.default_rng(100 + n)
    ck = (rng.standard_normal((n, n, n // 2 + 1)) +
          1j * rng.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
    g = fs.FFT(n)
    g.set_array_k(ck)
    g.execute_inverse()
    got = g.asarray()
    # FFTW c2r semantics == numpy irfftn (x/y complex inverse, then c2r on z
    # dropping Im of kz = 0 and kz = n/2), unnormalised
    ref = np.fft.irfftn(ck.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2)) * n**3
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_r2c_separable_closed_form(fs, n):
    """Large meshes (the persistent z passes run from n = 512): the input is
    a sum of two separable terms a(x) b(y) c(z), whose transform is the sum
    of the products of 1D transforms A(kx) B(ky) C(kz), compared slab by slab."""
    _announce(f"r2c separable n={n}")
    rng = np.random.default_rng(7 * n)
    terms = [[rng.standard_normal(n) for _ in range(3)] for _ in range(2)]
    a = np.empty((n, n, n), dtype=np.float32)
    for i in range(n):   # slab by slab: no n^3 FP64 temporaries
        a[i] = sum((ta[i] * tb[:, None] * tc[None, :]).astype(np.float32) for ta, tb, tc in terms)
    # what the device sees is the FP32 sum; its transform is that of the
    # FP64 terms up to the input rounding, bounded below with the tolerance
    g = fs.FFT(n)
    g.set_array(a)
    del a
    g.execute_forward()
    got = g.asarray_k()
    spec = [(np.fft.fft(ta), np.fft.fft(tb), np.fft.rfft(tc)) for ta, tb, tc in terms]
    scale = max(np.abs(A).max() * np.abs(B).max() * np.abs(C).max() for A, B, C in spec)
    err = 0.0
    for kx in range(n):
        ref = sum(A[kx] * B[:, None] * C[None, :] for A, B, C in spec)
        err = max(err, float(np.abs(got[kx] - ref).max()))
    assert err / scale < 4e-6, err / scale


@pytest.mark.parametrize("n", [512, 1024])
def test_round_trip_large(fs, n):
    """c2r(r2c(x)) = n^3 x at the sizes of the persistent z passes."""
    _announce(f"round trip n={n}")
    rng = np.random.default_rng(n + 3)
    a = rng.standard_normal((n, n, n), dtype=np.float32)
    g = fs.FFT(n)
    g.set_array(a)
    g.execute_forward()
    g.execute_inverse()
    back = g.asarray()
    back /= np.float32(n) ** 3
    err = float(np.abs(back - a).max()) / float(np.abs(a).max())
    assert err < 2e-6, err


def test_fft_test_data_layout(fs):
    """test/test_fft.py:9-28: set_test_data then gather -> index + 1."""
    _announce("fft test data")
    n = 8
    g = fs.FFT(n)
    g.set_test_data()
    a = g.asarray()
    assert np.array_equal(a.reshape(-1), np.arange(1, n**3 + 1, dtype=np.float32))


@pytest.mark.parametrize("x", [0, 1, 1.5, 3, 3.5, 4])
def test_pm_cic_one_particle(fs, x):
    """test/test_pm_cic.py:9-51: one particle on a 4^3 mesh, L = 64."""
    _announce(f"cic one particle x={x}")
    nc, boxsize = 4, 64.0
    dx = boxsize / nc
    fs.pm.init(nc, 1, boxsize)
    p = fs.Particles(nc, boxsize)
    p.append(np.array([[x * dx, 2 * dx, 2 * dx]]))
    fs.pm.send_positions(p)
    a = fs.pm.compute_density(p).asarray()
    total = np.sum(a.astype(np.float64) + 1.0)
    assert abs(total - 1.0) < 1.0e-15
    ref = ocic.paint(ocic.wrap(p.x, boxsize), nc, boxsize, 1.0, wdtype=np.float32)
    assert np.array_equal(a.astype(np.float64), ref)


@pytest.mark.parametrize("n,npart", [(16, 3000), (32, 20000)])
def test_paint_random_matches_oracle(fs, n, npart):
    _announce(f"paint random n={n}")
    rng = np.random.default_rng(n)
    L = 100.0
    # includes positions outside the box to exercise the fused wrap
    x = rng.uniform(-0.3 * L, 1.3 * L, size=(npart, 3))
    fs.pm.init(n, 2.0, L)
    p = fs.Particles(npart, L)
    p.append(x)
    fs.pm.send_positions(p)
    got = fs.pm.compute_density(p).asarray().astype(np.float64)
    xw = p.x
    assert xw.min() >= 0 and xw.max() < L
    ref = ocic.paint(ocic.wrap(x.astype(np.float32), L), n, L, 2.0, wdtype=np.float32)
    assert np.abs(got - ref).max() < 8 * 4e-7 * (1 + np.abs(ref).max())
    # mass: sum(delta) = Np pm^3 - Ng  (pm.cpp:357-390 tolerance)
    assert abs(got.sum() - (npart * 8.0 - n**3)) < 10 * 1.2e-7 * n**3


def _lattice_positions(n_p, L, amp, seed):
    rng = np.random.default_rng(seed)
    i = np.arange(n_p)
    q = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    q = (q + 0.5) * (L / n_p)
    return q + amp * (L / n_p) * rng.standard_normal(q.shape)


@pytest.mark.parametrize("n_p,nc", [(8, 16), (16, 32), (32, 64)])
def test_force_matches_fp64_oracle(fs, n_p, nc):
    """force relative RMS <= 1e-5 against the FP64 oracle."""
    _announce(f"force n_p={n_p} nc={nc}")
    L = 64.0
    x = _lattice_positions(n_p, L, 0.6, seed=nc)
    fs.pm.init(nc, nc / n_p, L)
    p = fs.Particles(n_p**3, L)
    p.append(x)
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)
    got = p.force.astype(np.float64)
    xw = p.x.astype(np.float64)
    ref = opm.force(xw, nc, L, nc / n_p)
    rel = np.sqrt(np.mean((got - ref) ** 2) / np.mean(ref**2))
    assert rel < 1e-5, rel
    # momentum conservation of symmetric paint/readout
    assert np.abs(got.sum(axis=0)).max() < 1e-3 * np.sqrt(np.mean(ref**2)) * len(x) ** 0.5


def test_force_plane_wave_closed_form(fs):
    """f = grad laplacian^-1 delta (doc/theory/fs.tex): delta = A cos(k.x)
    gives f = +A k/|k|^2 sin(k.x) on the mesh (pm.cpp:472-501 with
    fk = i f2 delta_k, f2 = -(L/2pi) k/|k|^2)."""
    _announce("plane wave force mesh")
    nc, L = 32, 32.0
    # particles at cell centres weighted to make delta = A cos(2 pi m x / L):
    # build the mesh through the oracle's Green function on a given delta
    i = np.arange(nc)
    X = (i * L / nc)[:, None, None] * np.ones((1, nc, nc))
    m = 3
    A = 0.1
    delta = A * np.cos(2 * np.pi * m * X / L)
    f = opm.force_meshes(delta, L)
    kphys = 2 * np.pi * m / L
    expect = A * kphys / kphys**2 * np.sin(kphys * X)
    assert np.abs(f[0] - expect).max() < 1e-12
    assert np.abs(f[1]).max() < 1e-12 and np.abs(f[2]).max() < 1e-12
    # GPU: same mesh pipeline fed the delta through the density grid
    fs.pm.init(nc, 1.0, L)
    p = fs.Particles(1, L)
    p.append(np.array([[0.5, 0.5, 0.5]]))
    p.set_times(0.5, 0.5, 0.0)
    g = fs.pm.compute_density(p)
    g.set_array(delta.astype(np.float32))
    fs.pm.compute_force(p)
    got = fs.pm.force_mesh(0).astype(np.float64)
    assert np.abs(got - expect).max() < 1e-6 * A * L
    assert np.abs(fs.pm.force_mesh(1)).max() < 1e-6 * A * L


def test_compute_force_requires_density(fs):
    _announce("force without density")
    fs.pm.init(16, 1.0, 16.0)
    p = fs.Particles(4, 16.0)
    p.append(np.array([[1.0, 2.0, 3.0]]))
    p.set_times(0.3, 0.3, 0.0)
    fs.pm.compute_density(p)
    fs.pm.compute_force(p)
    p.set_times(0.4, 0.4, 0.3)
    # pm.cpp:314-317: status is force_done, not density_done
    with pytest.raises(RuntimeError):
        fs.pm.compute_force(p)


@pytest.mark.parametrize("variant", ["sorted", "atomic", "global"])
def test_empty_particle_set(fs, variant):
    """No particles (an empty domain in the reference's loops): the density
    is -1 in every cell, the force call and a COLA kick + drift run and
    leave an empty set, and particles appended afterwards paint as usual."""
    _announce(f"empty particle set, paint {variant}")
    L, nc = 64.0, 16
    fs.cosmology.init(0.31)
    fs.pm.init(nc, 2.0, L)
    fs.pm.set_paint_variant(variant)
    try:
        p = fs.Particles(8, L)
        fs.pm.send_positions(p)
        assert np.all(fs.pm.compute_density(p).asarray() == -1.0)
        p.set_times(0.5, 0.5, 0.0)
        fs.pm.force(p)
        assert p.force.shape == (0, 3) and p.x.shape == (0, 3)
        fs.cola.kick(p, 0.55)
        fs.cola.drift(p, 0.6)
        assert len(p) == 0 and p.v.shape == (0, 3)
        assert p.times[:2] == (0.6, 0.55)
        p.append(np.array([[1.0, 2.0, 3.0]]))
        fs.pm.send_positions(p)
        d = fs.pm.compute_density(p).asarray().astype(np.float64)
        assert abs(d.sum() - (8.0 - nc**3)) < 1e-3     # one particle of mass pm_factor^3
    finally:
        fs.pm.set_paint_variant("sorted")


def test_partially_filled_particle_set(fs):
    """np_local < np_alloc: only the appended particles are painted, read out,
    kicked and drifted (the reference loops to np_local)."""
    _announce("partially filled particle set")
    L, nc, n = 64.0, 32, 3000
    rng = np.random.default_rng(5)
    fs.cosmology.init(0.31)
    fs.pm.init(nc, 2.0, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.Particles(5000, L)
    p.append(rng.uniform(0, L, (n, 3)))
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)
    got = p.force.astype(np.float64)
    assert got.shape == (n, 3)
    ref = opm.force(p.x.astype(np.float64), nc, L, 2.0)
    rel = float(np.sqrt(np.mean((got - ref) ** 2) / np.mean(ref ** 2)))
    assert rel < 1e-5, rel
    fs.cola.kick(p, 0.55)
    fs.cola.drift(p, 0.6)
    assert len(p) == n and np.all(np.isfinite(p.x)) and np.all(np.isfinite(p.v))
    fs.pm.send_positions(p)
    d = fs.pm.compute_density(p).asarray().astype(np.float64)
    assert abs(d.sum() - (n * 8.0 - nc**3)) < 10 * 1.2e-7 * nc**3   # exactly n particles


def _clustered(npart, L, rng):
    """Half uniform, half in a few tight clumps, plus particles on tile and
    box faces (lower cells on the last layer of a tile spill into the next)."""
    u = rng.uniform(0, L, (npart // 2, 3))
    centres = rng.uniform(0, L, (5, 3))
    c = centres[rng.integers(0, 5, npart - npart // 2)]
    clump = np.mod(c + 0.01 * L * rng.standard_normal(c.shape), L)
    x = np.concatenate([u, clump])
    x[:64, 0] = L * (1 - 1e-7)          # rounds to cell n -> index 0
    x[64:128, 1] = L * 15.5 / 16.0       # last layer of a 16-cell tile
    return x


@pytest.mark.parametrize("n", [8, 16, 48, 64, 96])
def test_paint_variants_match_oracle(fs, n):
    """Tiled (LDS atomics), deterministic (fixed point) and per-particle
    global-atomic paints all reproduce the oracle; the deterministic one is
    bitwise reproducible."""
    _announce(f"paint variants n={n}")
    rng = np.random.default_rng(7 * n)
    L = 50.0
    npart = 8 * n**3 // 8
    x = _clustered(npart, L, rng)
    fs.pm.init(n, 2.0, L)
    p = fs.Particles(npart, L)
    p.append(x)
    ref = ocic.paint(ocic.wrap(x.astype(np.float32), L), n, L, 2.0, wdtype=np.float32)
    # FP32 sums in any order over up to thousands of contributions per cell
    # (clumps): per-cell bound 1e-5 relative.  Fixed-point sums are exact to
    # 2^-32 per term, so the deterministic variant carries only the final
    # FP32 rounding.
    rho = np.abs(ref + 1.0)
    out = {}
    for variant in ("global", "atomic", "sorted"):
        fs.pm.set_paint_variant(variant)
        fs.pm.send_positions(p)
        out[variant] = fs.pm.compute_density(p).asarray().astype(np.float64)
        err = np.abs(out[variant] - ref)
        assert np.all(err <= 1e-5 * rho + 1e-6), (variant, (err / (rho + 1e-6)).max())
    assert np.all(np.abs(out["sorted"] - ref) <= 1.2e-7 * (rho + 1.0)), "sorted"
    fs.pm.send_positions(p)
    again = fs.pm.compute_density(p).asarray()
    assert np.array_equal(again, out["sorted"].astype(np.float32))


@pytest.mark.parametrize("n_p,nc", [(16, 32), (24, 48)])
def test_tiled_readout_matches_gather_readout(fs, n_p, nc):
    """compute_force reads the force meshes through the tile kernel when the
    paint's bins are current and through the per-particle gather otherwise;
    with the deterministic paint both see the same meshes and must agree
    bitwise."""
    _announce(f"tiled vs gather readout n_p={n_p} nc={nc}")
    L = 40.0
    x = np.mod(_lattice_positions(n_p, L, 0.8, seed=3 * nc), L)
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.Particles(n_p**3, L)
    p.append(x)
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)                 # bins current -> tile readout
    f_tile = p.force
    p.set_times(0.5, 0.5, -1.0)
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    p.set_field("x", p.x)          # same values, new version -> gather readout
    fs.pm.compute_force(p)
    f_gather = p.force
    assert np.array_equal(f_tile, f_gather)
    ref = opm.force(p.x.astype(np.float64), nc, L, nc / n_p)
    rel = np.sqrt(np.mean((f_tile - ref) ** 2) / np.mean(ref**2))
    assert rel < 1e-5, rel


@pytest.mark.timeout(600)
def test_column_readout_matches_unfused_path(fs):
    """At 1024^3 pm.force fuses the density's tile kernel with the z r2c
    (compute_density leaves the cells pending; tiles of 4 x 2 full z rows)
    and the z c2r of the three force meshes with the readout (rows
    transformed in LDS, no real force mesh written).  Forces must equal, bit
    for bit, those of the path with the density written first (read back
    between compute_density and compute_force) and those of the unfused path
    (z c2r passes writing the real meshes, per-particle gather readout), and
    force_mesh must return the same real meshes each way (finished on demand
    after the fused path)."""
    _announce("column readout vs unfused path, 128^3 particles, 1024^3 mesh")
    n_p, nc, L = 128, 1024, 1024.0
    x = np.mod(_lattice_positions(n_p, L, 2.0, seed=17), L)
    # a dense clump too: many particles per column, all three class bits
    x[:50000] = np.mod(512.0 + 3.0 * np.random.default_rng(18).standard_normal((50000, 3)), L)
    fs.pm.init(nc, nc / n_p, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.Particles(n_p**3, L)
    p.append(x)
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)                 # bins current -> fused column readout
    f_col = p.force
    assert np.isfinite(f_col).all() and np.abs(f_col).max() > 0
    mesh_col = [fs.pm.force_mesh(a)[::7, 3::11, :].copy() for a in range(3)]
    # density written by the plain tile kernel (read back), then the force
    p.set_times(0.5, 0.5, -1.0)
    fs.pm.send_positions(p)
    d = fs.pm.compute_density(p)
    rho = d.asarray()[::5, 2::9, :].copy()
    fs.pm.compute_force(p)
    assert np.array_equal(p.force, f_col), np.abs(p.force - f_col).max()
    # the density of the clump region: painted, not pending
    assert rho.max() > 0.0 and np.isclose(rho.min(), -1.0)
    p.set_times(0.5, 0.5, -1.0)
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    p.set_field("x", p.x)          # same values, new version -> unfused path
    fs.pm.compute_force(p)
    assert np.array_equal(p.force, f_col), np.abs(p.force - f_col).max()
    for a in range(3):
        assert np.array_equal(fs.pm.force_mesh(a)[::7, 3::11, :], mesh_col[a]), a


@pytest.mark.timeout(300)
def test_fused_path_with_empty_and_sparse_sets(fs):
    """The 1024^3 fused kernels (paint + z r2c, column readout) with almost
    every tile and column empty: no particles at all (the reference's empty
    domain), then six, two of them on the box's upper faces.  The sparse
    forces equal, bit for bit, those of the unfused path."""
    _announce("fused 1024^3 path: empty and sparse particle sets")
    nc, L = 1024, 1024.0
    fs.pm.init(nc, 2.0, L)
    fs.pm.set_paint_variant("sorted")
    p = fs.Particles(16, L)
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)
    assert p.force.shape == (0, 3)
    x = np.random.default_rng(23).uniform(0, L, (6, 3))
    x[0] = [L * (1 - 1e-7), 3.5, 1000.25]     # rounds to cell n -> index 0
    x[1] = [511.9, L - 0.01, L - 0.5]
    p.append(x)
    p.set_times(0.5, 0.5, 0.0)
    fs.pm.force(p)                            # fused
    f_fused = p.force.copy()
    assert np.isfinite(f_fused).all() and np.abs(f_fused).max() > 0
    p.set_times(0.5, 0.5, -1.0)
    fs.pm.send_positions(p)
    fs.pm.compute_density(p)
    p.set_field("x", p.x)                     # new version -> unfused path
    fs.pm.compute_force(p)
    assert np.array_equal(p.force, f_fused), np.abs(p.force - f_fused).max()


@pytest.mark.timeout(300)
def test_paint_variants_agree_at_1024(fs):
    """At 1024^3 the three paints take different paths: sorted (the
    default) fuses its tile kernel with the z r2c and is read out by the
    column kernel; the LDS-atomic tile paint writes the density first and
    is still read out by the column kernel; the global-atomic paint runs
    the unfused path (z r2c, z c2r passes, gather readout).  On one
    clustered set their forces agree to float-atomic rounding."""
    _announce("paint variants at 1024^3, 64^3 clustered particles")
    n_p, nc, L = 64, 1024, 1024.0
    x = np.mod(_lattice_positions(n_p, L, 3.0, seed=29), L)
    x[:20000] = np.mod(300.0 + 2.0 * np.random.default_rng(30).standard_normal((20000, 3)), L)
    fs.pm.init(nc, nc / n_p, L)
    forces = {}
    try:
        for variant in ("sorted", "atomic", "global"):
            fs.pm.set_paint_variant(variant)
            p = fs.Particles(n_p**3, L)
            p.append(x)
            p.set_times(0.5, 0.5, 0.0)
            fs.pm.force(p)
            forces[variant] = p.force.astype(np.float64)
    finally:
        fs.pm.set_paint_variant("sorted")
    ref = forces["sorted"]
    scale = np.sqrt(np.mean(ref**2))
    assert np.isfinite(ref).all() and scale > 0
    for variant in ("atomic", "global"):
        rel = float(np.sqrt(np.mean((forces[variant] - ref) ** 2)) / scale)
        print(f"[case] paint {variant} vs sorted at 1024^3: force rel RMS {rel:.2e}")
        assert rel < 1e-5, (variant, rel)


def test_mem_named_allocation(fs):
    """fs.mem.Mem (py/fs/mem.py): sizes rounded up to 16 bytes (mem.cpp:48-50),
    nothing allocated for size 0, MemoryError when the device cannot hold it."""
    _announce("mem")
    from fs2_amd.fs.mem import Mem
    m = Mem("test", 100)
    assert (m.mem_allocated, m.mem_using) == (112, 0)
    assert "test" in repr(m)
    assert (Mem("empty").mem_allocated, Mem("empty").mem_using) == (0, 0)
    with pytest.raises(MemoryError):
        Mem("too big", 1 << 50)


def test_particles_resize(fs):
    _announce("particles resize")
    p = fs.Particles(10, 1.0)
    p.append(np.arange(15, dtype=np.float64).reshape(5, 3) / 20.0)
    p.resize(3)
    assert len(p) == 3
    assert np.allclose(p.x, np.arange(9).reshape(3, 3) / 20.0)
    p.resize(8)
    assert len(p) == 8
    with pytest.raises(MemoryError):
        p.resize(11)
    with pytest.raises(AttributeError):
        fs.pm.write_packet_info("packets.h5")
