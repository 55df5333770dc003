"""Batched ensemble of small COLA boxes, replayed as one hipGraph.

No reference counterpart: the reference runs one box per process
(script/cola.py:28-48).  Boxes of one size form a class; every kernel of a
step covers a whole class, each class runs on its own HIP stream, and the
whole multi-step run of all classes is captured once into a hipGraph
(``fs2_batch_*`` in include/fs2hip.h).  Per box the results are bit-identical
to the single-box path (fs.lpt.init + fs.pm.force + fs.cola.kick/drift).

    ens = Ensemble([(32, 64, 32, 32.0), (48, 96, 32, 48.0), (64, 128, 32, 64.0)])
    ens.init_lpt(ps, 0.1, seeds=range(1, 97))
    ens.autotune()
    ens.set_schedule(0.1, 1.0, 10)
    ens.run()                      # graph replay (captured on first use)
    x = ens.field(2, 5, 'x')       # class 2, box 5: (64^3, 3) float32
"""
import ctypes

import numpy as np

from . import LIBRARY

_lib = None
_FIELDS = {"x": 0, "dx1": 1, "dx2": 2, "v": 3, "force": 4}


def _load():
    global _lib
    if _lib is not None:
        return _lib
    from . import fs  # noqa: F401  (loads the library with the _fs module)
    lib = ctypes.CDLL(LIBRARY)
    vp, i32, f64, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
    ip, dp = ctypes.POINTER(i32), ctypes.POINTER(f64)
    lib.fs2_last_error.restype = ctypes.c_char_p
    sig = {
        "fs2_batch_create": [i32, ip, ip, ip, dp, ctypes.POINTER(vp)],
        "fs2_batch_info": [vp, ip, ctypes.POINTER(u64)],
        "fs2_batch_init_lpt": [vp, vp, f64, ctypes.POINTER(u64)],
        "fs2_batch_set_tile": [vp, i32, i32, i32, i32],
        "fs2_batch_get_tile": [vp, i32, ip, ip, ip],
        "fs2_batch_autotune": [vp, i32, dp],
        "fs2_batch_set_schedule": [vp, f64, f64, i32],
        "fs2_batch_capture": [vp],
        "fs2_batch_run": [vp, i32],
        "fs2_batch_reset": [vp],
        "fs2_batch_get_field": [vp, i32, i32, i32, ctypes.POINTER(ctypes.c_float)],
        "fs2_synchronize": [],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = i32
    lib.fs2_batch_destroy.argtypes = [vp]
    lib.fs2_batch_destroy.restype = None
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib.fs2_last_error().decode())


def _power_pointer(ps):
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get(ps._ps, b"_PowerSpectrum")


class Ensemble:
    """classes: list of (nc_lpt, nc_pm, nbox, boxsize) tuples."""

    def __init__(self, classes):
        lib = _load()
        self.classes = [(int(a), int(b), int(c), float(d)) for a, b, c, d in classes]
        k = len(self.classes)
        nl = (ctypes.c_int * k)(*[c[0] for c in self.classes])
        nm = (ctypes.c_int * k)(*[c[1] for c in self.classes])
        nb = (ctypes.c_int * k)(*[c[2] for c in self.classes])
        L = (ctypes.c_double * k)(*[c[3] for c in self.classes])
        h = ctypes.c_void_p()
        _check(lib.fs2_batch_create(k, nl, nm, nb, L, ctypes.byref(h)))
        self._h = h
        npt = ctypes.c_uint64()
        _check(lib.fs2_batch_info(h, None, ctypes.byref(npt)))
        self.np_total = int(npt.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fs2_batch_destroy(h)
            self._h = None

    @property
    def nbox_total(self):
        return sum(c[2] for c in self.classes)

    def init_lpt(self, ps, a, seeds):
        """2LPT initial conditions ('cola' kind) at scale factor a; one seed
        per box, class-major."""
        seeds = [int(s) for s in seeds]
        if len(seeds) != self.nbox_total:
            raise ValueError(f"need {self.nbox_total} seeds, got {len(seeds)}")
        arr = (ctypes.c_uint64 * len(seeds))(*seeds)
        _check(_lib.fs2_batch_init_lpt(self._h, _power_pointer(ps), float(a), arr))

    def set_tile(self, cls, tile):
        _check(_lib.fs2_batch_set_tile(self._h, cls, *[int(t) for t in tile]))

    def tile(self, cls):
        t = [ctypes.c_int() for _ in range(3)]
        _check(_lib.fs2_batch_get_tile(self._h, cls, *[ctypes.byref(v) for v in t]))
        return tuple(v.value for v in t)

    def autotune(self, reps=3):
        """Chooses each class's paint tile shape by timing; returns
        [(tile, ms), ...] per class."""
        ms = (ctypes.c_double * len(self.classes))()
        _check(_lib.fs2_batch_autotune(self._h, int(reps), ms))
        return [(self.tile(k), ms[k]) for k in range(len(self.classes))]

    def set_schedule(self, a0, a1, nstep):
        _check(_lib.fs2_batch_set_schedule(self._h, float(a0), float(a1), int(nstep)))

    def capture(self):
        _check(_lib.fs2_batch_capture(self._h))

    def run(self, graph=True):
        """Runs the schedule from the current state (asynchronous)."""
        _check(_lib.fs2_batch_run(self._h, 1 if graph else 0))

    def reset(self):
        """Restores the initial conditions."""
        _check(_lib.fs2_batch_reset(self._h))

    def synchronize(self):
        _check(_lib.fs2_synchronize())

    def field(self, cls, box, name):
        nl = self.classes[cls][0]
        out = np.empty((nl**3, 3), dtype=np.float32)
        _check(_lib.fs2_batch_get_field(self._h, cls, box, _FIELDS[name],
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out
