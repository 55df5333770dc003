"""Standard (non-COLA) leapfrog time stepping, py/fs/leapfrog.py."""
from . import _fs as c


def set_initial_velocity(particles, a_vel):
    """v = Dv(a) dx1 + D2v(a) dx2 at scale factor a_vel (py/leapfrog.cpp:18-38)."""
    c._leapfrog_initial_velocity(particles._particles, a_vel)


def kick(particles, a_vel):
    """Advances the velocities from their current time a_v to a_vel with the
    forces at a_x: v += -1.5 Omega_m f SphiStd(a_v, a_vel)."""
    c._leapfrog_kick(particles._particles, a_vel)


def drift(particles, a_pos):
    """Advances the positions from a_x to a_pos at the current velocities:
    x += v SqStd(a_x, a_pos)."""
    c._leapfrog_drift(particles._particles, a_pos)


def factors(a_i, a_f):
    """(kick factor SphiStd, drift factor SqStd) between a_i and a_f, FP64."""
    return c._leapfrog_factors(a_i, a_f)
