"""2LPT initial conditions, py/fs/lpt.py."""
from . import _fs as c
from .particles import Particles


def init(nc, boxsize, a, ps, seed, kind):
    """Generate 2LPT displacements and particle positions on the GPU.

    Args:
        nc (int): particles per dimension; np = nc**3.
        boxsize (float): periodic box length [Mpc/h].
        a (float): scale factor of the positions.
        ps (PowerSpectrum): linear P(k) extrapolated to a = 1.
        seed (int): seed of the white-noise field.
        kind (str): 'zeldovich', '2lpt' or 'cola' ('cola' sets v = 0).

    Returns:
        Particles.
    """
    return Particles(_particles=c._lpt(nc, boxsize, a, seed, ps._ps,
                                       kind.lower()))


def set_offset(offset):
    """Lattice offset: x = (ix + offset)*boxsize/nc (default 0.5)."""
    c._lpt_set_offset(offset)


def set_zeldovich_force(particles, a):
    """particles.force = D1(a) dx1."""
    c._lpt_set_zeldovich_force(particles._particles, a)


def white_noise(nc, seed):
    """The (nc, nc, nc) white-noise field that init(nc, ..., seed) uses."""
    return c._lpt_white_noise(nc, seed)
