"""CPU tests of the C ABI library and its host-side numerics (no GPU calls).

* libfs2hip.so loads and exports every function include/fs2hip.h declares;
* FP64 host cosmology, COLA coefficients and power spectra match the oracle;
* error codes map onto the reference's Python exception types.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import cola as ocola
from oracle import cosmology as ocosmo
from oracle import power as opower

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fs2hip.h")
LIB = os.path.join(ROOT, "fs2_amd", "libfs2hip.so")
TABLE = os.path.join(ROOT, "tests", "golden", "planck_matterpower.dat")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs2_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_io_library_exports_every_declared_symbol():
    """libfs2io.so (HDF5 output, built where an HDF5 C library exists)."""
    iolib = os.path.join(ROOT, "fs2_amd", "libfs2io.so")
    if not os.path.exists(iolib):
        pytest.skip("libfs2io.so not built (no HDF5 C library)")
    import torch  # noqa: F401  (one HIP runtime: the one libfs2hip binds in the package)
    lib = ctypes.CDLL(iolib)
    names = declared_functions(os.path.join(ROOT, "include", "fs2io.h"))
    assert names == ["fs2_write_hdf5"]
    assert hasattr(lib, "fs2_write_hdf5")


def test_extension_method_table_matches_reference_names(fs):
    """The method names of py/py_package.cpp:28-157 that are on the hot path."""
    expected = [
        "set_loglevel", "comm_mpi_init", "comm_mpi_finalise", "comm_hello", "comm_this_node",
        "comm_n_nodes", "_cosmology_init", "_cosmology_D_growth", "_cosmology_D2_growth",
        "_power_alloc", "_power_n", "_power_i", "_particles_alloc", "_particles_slice",
        "_particles_len", "_particles_np_total", "_particles_update_np_total",
        "_particles_getitem", "_particles_id_asarray", "_particles_x_asarray",
        "_particles_force_asarray", "_particles_periodic_wrapup", "_particles_clear",
        "_particles_append", "_lpt", "_lpt_set_offset", "_lpt_set_zeldovich_force", "_pm_init",
        "_pm_compute_force", "_pm_compute_density", "_pm_domain_init", "_pm_send_positions",
        "_pm_check_total_density", "_pm_get_forces", "_pm_set_packet_size", "_cola_kick",
        "_cola_drift", "_fft_alloc", "_fft_set_test_data", "_fft_fx_global_as_array",
        "_leapfrog_initial_velocity", "_leapfrog_kick", "_leapfrog_drift",
        "_write_gadget_binary", "_hdf5_write_particles",
        "config_precision"]
    missing = [n for n in expected if not hasattr(fs._fs, n)]
    assert not missing, missing
    assert fs.config_precision() == "single"
    assert fs.comm.this_node() == 0 and fs.comm.n_nodes() == 1


@pytest.mark.parametrize("om", [0.308, 0.31])
def test_host_cosmology_matches_oracle(fs, om):
    fs.cosmology.init(om)
    c = ocosmo.Cosmology(om)
    for a in (0.05, 0.1, 0.19, 0.55, 1.0):
        D = c.D(a)
        assert abs(fs.cosmology.D_growth(a) - D) < 1e-12 * D
        assert abs(fs.cosmology.D2_growth(a) - c.D2(a, D)) < 1e-12 * abs(c.D2(a, D))
        assert abs(fs.cosmology.f_growth_rate(a) - c.f(a)) < 1e-11
        assert abs(fs._fs._cosmology_Dv_growth(a) - c.Dv(a, D)) < 1e-11 * abs(c.Dv(a, D))


def test_host_cola_coefficients_match_oracle(fs):
    om = 0.31
    fs.cosmology.init(om)
    c = ocosmo.Cosmology(om)
    for (av, ax, avel) in [(0.1, 0.1, 0.145), (0.145, 0.19, 0.235), (0.91, 0.955, 1.0)]:
        kf, q1, q2 = fs._fs._cola_kick_coefficients(av, ax, avel)
        rkf, rq1, rq2 = ocola.kick_coefficients(c, av, ax, avel)
        assert abs(kf - rkf) < 1e-13 * abs(rkf)
        assert abs(q1 - rq1) < 1e-12 * rq1 and abs(q2 - rq2) < 1e-12 * abs(rq2)
        dt, da1, da2 = fs._fs._cola_drift_coefficients(ax, avel, ax + 0.045)
        rdt, rda1, rda2 = ocola.drift_coefficients(c, ax, avel, ax + 0.045)
        assert abs(dt - rdt) < 1e-12 * abs(rdt)
        assert abs(da1 - rda1) < 1e-11 * abs(rda1)
        assert abs(da2 - rda2) < 1e-11 * abs(rda2)


def test_host_leapfrog_factors_match_oracle(fs):
    from oracle import leapfrog as olf
    om = 0.31
    fs.cosmology.init(om)
    c = ocosmo.Cosmology(om)
    for ai, af in [(0.1, 0.145), (0.5, 0.6), (0.955, 1.0)]:
        kf, dt = fs.leapfrog.factors(ai, af)
        assert abs(kf - olf.kick_factor(c, ai, af)) < 1e-12 * kf
        assert abs(dt - olf.drift_factor(c, ai, af)) < 1e-12 * dt


def test_table_power_spectrum_matches_oracle(fs):
    ps = fs.PowerSpectrum(TABLE)
    ref = opower.TablePower(TABLE)
    assert len(ps) == len(ref.k)
    k0, P0 = ps[0]
    assert k0 == pytest.approx(ref.k[0], rel=1e-15) and P0 == pytest.approx(ref.Ptab[0], rel=1e-15)
    for k in np.geomspace(2e-4, 5e3, 40):
        assert ps.P(k) == pytest.approx(float(ref.P(k)), rel=1e-11)
    assert ps.sigma(8.0) == pytest.approx(ref.sigma(8.0), rel=1e-12)
    with pytest.raises(IndexError):
        ps[len(ps)]


def test_eh98_power_matches_oracle(fs):
    ps = fs.PowerSpectrum.eisenstein_hu(0.31, 0.049, 0.67, 0.96, 0.81)
    ref = opower.EH98NoWiggle(0.31, 0.049, 0.67, 0.96, 0.81)
    for k in np.geomspace(1e-4, 50, 30):
        assert ps.P(k) == pytest.approx(float(ref.P(k)), rel=1e-9)
    assert ps.sigma(8.0) == pytest.approx(0.81, rel=1e-10)


def test_power_file_errors_raise_ioerror(fs, tmp_path):
    with pytest.raises(IOError):
        fs.PowerSpectrum(str(tmp_path / "missing.dat"))
    bad = tmp_path / "bad.dat"
    bad.write_text("0.1 1.0\n0.05 2.0\n0.2 3.0\n")
    with pytest.raises(IOError):   # k must increase (power.cpp:108-116)
        fs.PowerSpectrum(str(bad))


def test_pm_init_argument_errors(fs):
    # pm.cpp:245-255 (checked before any device allocation)
    with pytest.raises(RuntimeError):
        fs.pm.init(1, 1.0, 64.0)
    with pytest.raises(RuntimeError):
        fs.pm.init(16, 1.0, -1.0)


def _have_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_have_gpu(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly(fs):
    """No CPU fallback: device entry points raise without a GPU."""
    with pytest.raises(RuntimeError, match="no HIP device"):
        fs.Particles(10, 1.0)


def test_error_paths_without_a_device():
    """Argument and state errors come back as status codes with a message,
    before any device work (no GPU is touched here)."""
    import torch  # noqa: F401  (one HIP runtime, as the package loads it)
    lib = ctypes.CDLL(LIB)
    lib.fs2_last_error.restype = ctypes.c_char_p
    FS2_ERR_VALUE, FS2_ERR_RUNTIME = 5, 1
    u64 = ctypes.c_uint64
    # handle checks
    for name in ("fs2_cola_kick", "fs2_cola_drift"):
        f = getattr(lib, name)
        f.argtypes = [ctypes.c_void_p, ctypes.c_double]
        assert f(None, 0.5) == FS2_ERR_VALUE
        assert b"null particles handle" in lib.fs2_last_error()
    lib.fs2_particles_resize.argtypes = [ctypes.c_void_p, u64]
    assert lib.fs2_particles_resize(None, 3) == FS2_ERR_VALUE
    lib.fs2_mem_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    assert lib.fs2_mem_info(None, None, None) == FS2_ERR_VALUE
    lib.fs2_copy_bandwidth.argtypes = [u64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    assert lib.fs2_copy_bandwidth(1 << 20, 1024, 1, None) == FS2_ERR_VALUE
    # FoF state queries before any run: no groups, no tree
    n = u64(123)
    lib.fs2_fof_grp.argtypes = [ctypes.c_void_p, ctypes.POINTER(u64)]
    assert lib.fs2_fof_grp(None, ctypes.byref(n)) == 0 and n.value == 0
    h = u64()
    lib.fs2_kdtree_height.argtypes = [ctypes.POINTER(u64)]
    assert lib.fs2_kdtree_height(ctypes.byref(h)) == FS2_ERR_RUNTIME
    assert b"no kd-tree" in lib.fs2_last_error()
    # PM parameters (py/pm.cpp:245-255), checked before any allocation
    pm = ctypes.c_void_p()
    lib.fs2_pm_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_void_p)]
    assert lib.fs2_pm_create(1, 1.0, 64.0, ctypes.byref(pm)) == FS2_ERR_RUNTIME
    assert b"must be larger than 1" in lib.fs2_last_error()
    assert lib.fs2_pm_create(16, 1.0, 0.0, ctypes.byref(pm)) == FS2_ERR_RUNTIME
    assert b"must be positive" in lib.fs2_last_error()
    # FFTW's any-n is kept for even n with prime factors 2, 3, 5, 7
    for bad in (22, 15, 4100):
        assert lib.fs2_pm_create(bad, 1.0, 64.0, ctypes.byref(pm)) == FS2_ERR_VALUE
        assert b"not a supported FFT size" in lib.fs2_last_error()
    assert pm.value is None
