"""fs2_pm_power_spectrum_noalloc through the C ABI: the same bins as
fs2_pm_power_spectrum on two calls in a row, the second on a smaller mesh
with fewer bins (stale sums of the first call must not leak into it), and a
clear error above the fixed scratch size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from fs2_amd import LIBRARY
    import fs2_amd.fs  # noqa: F401  (loads the library with the _fs module)
    lib = ctypes.CDLL(LIBRARY)
    vp, i32, f64, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    dp = ctypes.POINTER(f64)
    sig = {
        "fs2_cosmology_init": [f64],
        "fs2_power_alloc_eh98": [f64, f64, f64, f64, f64, ctypes.POINTER(vp)],
        "fs2_lpt": [i32, f64, f64, u64, vp, ctypes.c_char_p, ctypes.POINTER(vp)],
        "fs2_pm_create": [i32, f64, f64, ctypes.POINTER(vp)],
        "fs2_pm_compute_density": [vp, vp, ctypes.POINTER(vp)],
        "fs2_pm_power_spectrum": [vp, i32, f64, dp, dp, dp],
        "fs2_pm_power_spectrum_noalloc": [vp, i32, f64, dp, dp, dp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = i32
    lib.fs2_last_error.restype = ctypes.c_char_p
    for name in ("fs2_pm_destroy", "fs2_particles_free", "fs2_power_free"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = None
    return lib


def _ok(lib, rc):
    assert rc == 0, lib.fs2_last_error().decode()


def _pk(lib, fn, pm, nbins):
    k, P, nm = (np.zeros(nbins), np.zeros(nbins), np.zeros(nbins))
    dp = ctypes.POINTER(ctypes.c_double)
    rc = fn(pm, nbins, 1.0, k.ctypes.data_as(dp), P.ctypes.data_as(dp), nm.ctypes.data_as(dp))
    return rc, k, P, nm


def test_noalloc_power_spectrum_matches_allocating_entry():
    lib = _lib()
    _ok(lib, lib.fs2_cosmology_init(0.31))
    ps = ctypes.c_void_p()
    _ok(lib, lib.fs2_power_alloc_eh98(0.31, 0.049, 0.67, 0.96, 0.81, ctypes.byref(ps)))
    for n_p, nc, L, nbins in ((32, 64, 64.0, 40), (16, 32, 64.0, 12)):
        print(f"[case] noalloc P(k): {n_p}^3 particles, {nc}^3 mesh, {nbins} bins", flush=True)
        p, pm, g = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _ok(lib, lib.fs2_lpt(n_p, L, 0.5, 7, ps, b"cola", ctypes.byref(p)))
        _ok(lib, lib.fs2_pm_create(nc, nc / n_p, L, ctypes.byref(pm)))
        _ok(lib, lib.fs2_pm_compute_density(pm, p, ctypes.byref(g)))
        # the first case runs the r2c in the allocating call, the second in
        # the allocation-free one (its twiddles were built by fs2_pm_create)
        calls = [lib.fs2_pm_power_spectrum, lib.fs2_pm_power_spectrum_noalloc]
        if nc == 32:
            calls.reverse()
        out = {}
        for fn in calls:
            rc, *res = _pk(lib, fn, pm, nbins)
            _ok(lib, rc)
            out[fn is lib.fs2_pm_power_spectrum_noalloc] = res
        (k0, P0, n0), (k1, P1, n1) = out[False], out[True]
        assert np.array_equal(n0, n1)            # integer mode counts: exact
        assert n0[1:].min() > 0
        ok = n0 > 0
        assert np.allclose(k1[ok], k0[ok], rtol=1e-13, atol=0)
        assert np.allclose(P1[ok], P0[ok], rtol=1e-12, atol=0)
        lib.fs2_pm_destroy(pm)
        lib.fs2_particles_free(p)
    # beyond the fixed scratch: a clear error, nothing written
    p, pm, g = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _ok(lib, lib.fs2_lpt(8, 32.0, 0.5, 3, ps, b"cola", ctypes.byref(p)))
    _ok(lib, lib.fs2_pm_create(16, 2.0, 32.0, ctypes.byref(pm)))
    _ok(lib, lib.fs2_pm_compute_density(pm, p, ctypes.byref(g)))
    rc, k, P, nm = _pk(lib, lib.fs2_pm_power_spectrum_noalloc, pm, 2049)
    assert rc == 5 and b"2048" in lib.fs2_last_error()   # FS2_ERR_VALUE
    assert not k.any() and not P.any() and not nm.any()
    lib.fs2_pm_destroy(pm)
    lib.fs2_particles_free(p)
    lib.fs2_power_free(ps)
