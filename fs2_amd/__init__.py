"""MI355X-native PM + COLA hot path of fs2.

``fs2_amd.fs`` is the drop-in counterpart of the reference's ``fs`` package
(py/fs/*.py): the same public API over the ``_fs`` CPython module, which
calls the HIP kernels of ``libfs2hip.so`` through the C ABI declared in
``include/fs2hip.h``.  There is no CPU fallback: without a GPU every entry
point that touches device memory raises RuntimeError.
"""
import os

LIBDIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(LIBDIR, 'libfs2hip.so')

# One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64
# (soname libamdhip64.so.7) and its libraries ask for it by the file name
# libamdhip64.so, so if libfs2hip.so were loaded first (binding
# /opt/rocm/lib/libamdhip64.so.7) and torch second, the process would hold
# two HIP runtimes and the second to start could not open the device.
# Importing torch first makes libfs2hip.so bind to torch's runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass
