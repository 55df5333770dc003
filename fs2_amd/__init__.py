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
