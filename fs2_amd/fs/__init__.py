"""fs: fast cosmological simulation -- PM + COLA on MI355X.

Mirrors py/fs/__init__.py of the reference: importing the package
initialises communication (one rank here) and exposes the same modules.
FoF and the kd-tree run on the host over positions copied from the device.
"""
import atexit

from . import _fs
from . import comm
from . import cosmology
from . import lpt
from . import pm
from . import cola
from . import leapfrog
from . import fof
from . import msg
from .particles import Particles
from .power import PowerSpectrum
from .fft import FFT
from .kdtree import KdTree

_fs.comm_mpi_init()

atexit.register(_fs.comm_mpi_finalise)


def config_precision():
    """'single': state, meshes and forces are FP32 (py/py_config.cpp)."""
    return _fs.config_precision()
