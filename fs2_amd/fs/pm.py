"""Particle-mesh gravity, py/fs/pm.py."""
from . import _fs as c
from .fft import FFT

PAINT_ATOMIC = 0
PAINT_SORTED = 1
PAINT_GLOBAL = 2


def init(nc_pm, pm_factor, boxsize):
    """Initialise the PM module (idempotent for unchanged parameters).

    Args:
        nc_pm (int): mesh points per dimension
        pm_factor (float): nc_pm / nc
        boxsize (float): length of the periodic box
    """
    c._pm_init(nc_pm, pm_factor, boxsize)


def force(particles):
    """Compute particles.force from particles.x (py/fs/pm.py:17-31)."""
    send_positions(particles)
    compute_density(particles)
    check_total_density()
    compute_force(particles)
    get_forces(particles)


def send_positions(particles):
    """Periodic wrap of the positions (one rank: nothing to send)."""
    c._pm_send_positions(particles._particles)


def compute_density(particles):
    """Density mesh delta from particles.x; returns a borrowed FFT."""
    return FFT(c._pm_compute_density(particles._particles))


def check_total_density():
    """<delta> = 0 check; a no-op unless enabled by set_check_total_density."""
    c._pm_check_total_density()


def set_check_total_density(on=True):
    c._pm_set_check_total_density(bool(on))


def total_density():
    """FP64 sum of delta over the mesh (must follow compute_density)."""
    return c._pm_total_density()


def compute_force(particles):
    """particles.force from the density mesh."""
    c._pm_compute_force(particles._particles)


def get_forces(particles):
    c._pm_get_forces(particles._particles)


def domain_init(particles):
    c._pm_domain_init(particles._particles)


def set_packet_size(packet_size):
    c._pm_set_packet_size(packet_size)


def set_paint_variant(variant):
    """'sorted' (default: tiled, deterministic int64 fixed-point tile sums),
    'atomic' (tiled, FP32 LDS atomics) or 'global' (per-particle global
    FP32 atomics on a mesh cleared to -1)."""
    if isinstance(variant, str):
        variant = {'atomic': PAINT_ATOMIC, 'sorted': PAINT_SORTED,
                   'global': PAINT_GLOBAL}[variant]
    c._pm_set_paint_variant(variant)


def force_mesh(axis):
    """Real-space force mesh of one axis from the last compute_force."""
    return c._pm_force_mesh(axis)


def power_spectrum(nbins=None, dk_bin=1.0, nc=None):
    """Binned P(k) of the density mesh: returns (k, P, nmodes)."""
    if nbins is None:
        if nc is None:
            raise ValueError('give nbins or nc')
        nbins = int(nc // 2 / dk_bin)
    return c._pm_power_spectrum(nbins, dk_bin)


def write_packet_info(filename):
    """MPI particle-exchange information (py/fs/pm.py:108-114).  The
    reference's wrapper calls a method its module does not register, and
    there are no packets on one rank; this raises as the reference does."""
    raise AttributeError("_pm_write_packet_info is not available: particle packets exist only "
                         "in multi-rank runs")
