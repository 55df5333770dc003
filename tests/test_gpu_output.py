"""GPU-side output formats: the Gadget-1 binary snapshot
(py/gadget_file.cpp:15-114) and the HDF5 snapshot (py/hdf5_write.cpp:28-246),
read back and compared with the particle arrays."""
import numpy as np
import pytest

import h5read

pytestmark = pytest.mark.gpu

OM = 0.31
GADGET_HEADER = np.dtype([("np", "<i4", 6), ("mass", "<f8", 6), ("time", "<f8"),
                          ("redshift", "<f8"), ("flag_sfr", "<i4"), ("flag_feedback", "<i4"),
                          ("np_total", "<u4", 6), ("flag_cooling", "<i4"),
                          ("num_files", "<i4"), ("boxsize", "<f8"), ("omega0", "<f8"),
                          ("omega_lambda", "<f8"), ("hubble_param", "<f8"),
                          ("flag_stellarage", "<i4"), ("flag_metals", "<i4"),
                          ("np_total_highword", "<u4", 6), ("flag_entropy_instead_u", "<i4"),
                          ("fill", "V60")])


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def particles(fs):
    fs.cosmology.init(OM)
    ps = fs.PowerSpectrum.eisenstein_hu(OM, 0.049, 0.67, 0.96, 0.81)
    p = fs.lpt.init(8, 32.0, 0.25, ps, 3, "2lpt")
    return p


def _blocks(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        out.append(raw[pos + 4:pos + 4 + n])
        assert int(np.frombuffer(raw, "<i4", 1, pos + 4 + n)[0]) == n
        pos += n + 8
    return out


@pytest.mark.parametrize("longid", [False, True])
def test_gadget_binary(fs, particles, tmp_path, longid):
    _announce(f"gadget longid={longid}")
    p = particles
    path = str(tmp_path / "snap.gadget")
    p.save_gadget_binary(path, longid)
    blocks = _blocks(path)
    assert len(blocks) == 4 and len(blocks[0]) == 256
    h = np.frombuffer(blocks[0], GADGET_HEADER)[0]
    n = len(p)
    assert h["np"][1] == n and h["np_total"][1] == n and h["num_files"] == 1
    assert h["time"] == 0.25 and h["redshift"] == pytest.approx(3.0)
    assert h["boxsize"] == 32.0 and h["omega0"] == OM and h["omega_lambda"] == 1 - OM
    # particle mass = Om rho_crit L^3 / N with rho_crit = 2.775e11 h^2 M_sun / Mpc^3
    assert h["mass"][1] == pytest.approx(OM * 2.775e11 * 32.0**3 / n, rel=1e-3)
    assert np.array_equal(np.frombuffer(blocks[1], "<f4").reshape(n, 3), p.x)
    v = (p.v * np.float32(1.0 / np.sqrt(0.25))).astype(np.float32)
    assert np.allclose(np.frombuffer(blocks[2], "<f4").reshape(n, 3), v, rtol=1e-6, atol=0)
    ids = np.frombuffer(blocks[3], "<u8" if longid else "<u4")
    assert np.array_equal(ids, p.id)


@pytest.mark.skipif(not h5read.available(), reason="no HDF5 C library")
def test_hdf5_snapshot(fs, particles, tmp_path):
    _announce("hdf5 ixvf12")
    p = particles
    path = str(tmp_path / "snap.h5")
    p.save_hdf5(path, "ixvf12")
    n = len(p)
    assert np.array_equal(h5read.read(path, "id"), p.id)
    for name, ref in (("x", p.x), ("v", p.v), ("f", p.force), ("dx1", p.dx1), ("dx2", p.dx2)):
        got = h5read.read(path, name)
        assert got.shape == (n, 3) and np.array_equal(got, ref), name
    assert h5read.read(path, "parameters/boxsize") == 32.0
    assert h5read.read(path, "parameters/omega_m") == OM
    assert h5read.read(path, "parameters/ax") == 0.25
    assert h5read.read(path, "parameters/nc") == 8


@pytest.mark.skipif(not h5read.available(), reason="no HDF5 C library")
def test_hdf5_cola_velocity_and_errors(fs, particles, tmp_path):
    _announce("hdf5 c / errors")
    p = particles
    path = str(tmp_path / "snap_c.h5")
    p.save_hdf5(path, "c")
    assert np.array_equal(h5read.read(path, "v"), fs._fs._cola_velocity(p._particles))
    with pytest.raises(ValueError):
        p.save_hdf5(str(tmp_path / "bad.h5"), "xi")       # out of order
    with pytest.raises(IOError):
        p.save_hdf5(path, "x")                            # /parameters exists already
    with pytest.raises(IOError):
        p.save_gadget_binary(str(tmp_path / "no_such_dir" / "snap"), False)
