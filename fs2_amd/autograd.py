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
    lib.fs2_diff_create.argtypes = [i32, i32, f64, f64, f64, i32, vp, i32, ctypes.POINTER(vp)]
    lib.fs2_diff_params.argtypes = [vp] + [ctypes.POINTER(i32)] * 3 + [
        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(i32)]
    lib.fs2_diff_set_paint_variant.argtypes = [vp, i32]
    lib.fs2_diff_destroy.argtypes = [vp]
    lib.fs2_diff_destroy.restype = None
    lib.fs2_diff_forward.argtypes = [vp, vp, vp]
    lib.fs2_diff_backward.argtypes = [vp, vp, vp]
    for f in (lib.fs2_diff_create, lib.fs2_diff_forward, lib.fs2_diff_backward,
              lib.fs2_diff_params, lib.fs2_diff_set_paint_variant):
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
        checkpoint: 'auto' (store when it fits in half of the free device
            memory), 'recompute' (positions only; each step's density and
            force meshes are rebuilt in the backward sweep) or 'store'
            (also keep every step's force meshes and particle bins).
    fs.cosmology.init must have been called.
    """

    _POLICY = {"auto": 0, "recompute": 1, "store": 2}
    _VARIANT = {"sorted": 1, "global": 2}

    def __init__(self, nc, nc_pm, boxsize, a0, a1, nstep, ps, checkpoint="auto"):
        lib = _load()
        self.nc, self.nc_pm, self.nstep = int(nc), int(nc_pm), int(nstep)
        self.boxsize, self.a0, self.a1 = float(boxsize), float(a0), float(a1)
        self._ps = ps   # keeps the power spectrum alive
        h = ctypes.c_void_p()
        if checkpoint not in self._POLICY:
            raise ValueError(f"checkpoint must be one of {sorted(self._POLICY)}")
        _check(lib.fs2_diff_create(self.nc, self.nc_pm, self.boxsize, self.a0, self.a1,
                                   self.nstep, _power_pointer(ps), self._POLICY[checkpoint],
                                   ctypes.byref(h)))
        self._h = h
        self.forward_count = 0
        st = ctypes.c_int()
        _check(lib.fs2_diff_params(h, None, None, None, None, ctypes.byref(st)))
        self.stores_forces = bool(st.value)

    def set_paint_variant(self, variant):
        """'sorted' (tiled, deterministic; default) or 'global' (atomics and
        gather kernels).  Invalidates the checkpoints of the last forward."""
        if variant not in self._VARIANT:
            raise ValueError(f"variant must be one of {sorted(self._VARIANT)}")
        _check(_lib.fs2_diff_set_paint_variant(self._h, self._VARIANT[variant]))
        self.forward_count += 1

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
