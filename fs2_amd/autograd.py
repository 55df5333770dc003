"""Differentiable COLA run on the GPU as a torch.autograd.Function.

No reference counterpart: fs2 has no adjoint.  The forward is the
reference's 'cola' pipeline -- white noise -> 2LPT (py/lpt.cpp, v0 = 0) ->
nstep COLA kick/drift steps (py/cola.cpp:48-127, schedule of
script/cola.py:37-44) with PM forces (py/pm.cpp) -> final CIC density.
Both directions run as HIP kernels of libfs2hip.so through the C ABI
``fs2_diff_*`` (include/fs2hip.h); torch only owns the device buffers.

    model = ColaModel(nc=128, nc_pm=256, boxsize=512.0, a0=0.1, a1=1.0,
                      nstep=10, ps=fs.PowerSpectrum.eisenstein_hu())
    delta = cola_density(noise, model)       # noise: (nc,)*3 float32 on cuda
    loss = ((delta - target) ** 2).mean()
    loss.backward()                          # noise.grad via the adjoint sweep

One model keeps the per-step positions of its LAST forward; calling
backward after another forward on the same model raises RuntimeError.
"""
import ctypes

import torch

from . import LIBRARY

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    from . import fs  # noqa: F401  (loads the library with the _fs module)
    lib = ctypes.CDLL(LIBRARY)
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.fs2_last_error.restype = ctypes.c_char_p
    lib.fs2_diff_create.argtypes = [i32, i32, f64, f64, f64, i32, vp, ctypes.POINTER(vp)]
    lib.fs2_diff_destroy.argtypes = [vp]
    lib.fs2_diff_destroy.restype = None
    lib.fs2_diff_forward.argtypes = [vp, vp, vp]
    lib.fs2_diff_backward.argtypes = [vp, vp, vp]
    for f in (lib.fs2_diff_create, lib.fs2_diff_forward, lib.fs2_diff_backward):
        f.restype = i32
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


class ColaModel:
    """Device state of one differentiable COLA configuration.

    Args:
        nc: particles per dimension (white-noise grid, nc % 4 == 0).
        nc_pm: PM mesh per dimension.
        boxsize: box length [1/h Mpc].
        a0, a1: initial and final scale factor; nstep COLA steps between.
        ps: fs.PowerSpectrum of the linear field.
    fs.cosmology.init must have been called.
    """

    def __init__(self, nc, nc_pm, boxsize, a0, a1, nstep, ps):
        lib = _load()
        self.nc, self.nc_pm, self.nstep = int(nc), int(nc_pm), int(nstep)
        self.boxsize, self.a0, self.a1 = float(boxsize), float(a0), float(a1)
        self._ps = ps   # keeps the power spectrum alive
        h = ctypes.c_void_p()
        _check(lib.fs2_diff_create(self.nc, self.nc_pm, self.boxsize, self.a0, self.a1,
                                   self.nstep, _power_pointer(ps), ctypes.byref(h)))
        self._h = h
        self.forward_count = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fs2_diff_destroy(h)
            self._h = None

    def _device_tensor(self, t, n, name):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise RuntimeError(f"{name} must be a CUDA (HIP) tensor")
        if t.shape != (n, n, n):
            raise ValueError(f"{name} must have shape {(n, n, n)}, got {tuple(t.shape)}")
        return t.detach().to(torch.float32).contiguous()

    def forward(self, noise):
        """Final density contrast (nc_pm,)*3 float32 of the run from `noise`."""
        w = self._device_tensor(noise, self.nc, "noise")
        delta = torch.empty((self.nc_pm,) * 3, dtype=torch.float32, device=w.device)
        torch.cuda.current_stream(w.device).synchronize()
        _check(_lib.fs2_diff_forward(self._h, w.data_ptr(), delta.data_ptr()))
        self.forward_count += 1
        return delta

    def backward(self, delta_bar):
        """d loss / d noise for d loss / d delta = delta_bar (last forward)."""
        g = self._device_tensor(delta_bar, self.nc_pm, "delta_bar")
        out = torch.empty((self.nc,) * 3, dtype=torch.float32, device=g.device)
        torch.cuda.current_stream(g.device).synchronize()
        _check(_lib.fs2_diff_backward(self._h, g.data_ptr(), out.data_ptr()))
        return out


class ColaDensity(torch.autograd.Function):
    """noise -> final CIC density, with the hand-written adjoint as backward."""

    @staticmethod
    def forward(ctx, noise, model):
        delta = model.forward(noise)
        ctx.model = model
        ctx.count = model.forward_count
        return delta

    @staticmethod
    def backward(ctx, grad):
        if ctx.model.forward_count != ctx.count:
            raise RuntimeError("the model ran another forward since this graph was built; "
                               "its per-step checkpoints are gone")
        return ctx.model.backward(grad), None


def cola_density(noise, model):
    return ColaDensity.apply(noise, model)
