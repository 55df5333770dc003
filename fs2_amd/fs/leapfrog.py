"""Standard (non-COLA) leapfrog time stepping, py/fs/leapfrog.py."""
from . import _fs as c


def set_initial_velocity(particles, a_vel):
    """v = Dv(a) dx1 + D2v(a) dx2 at scale factor a_vel (py/leapfrog.cpp:18-38)."""
    c._leapfrog_initial_velocity(particles._particles, a_vel)


def kick(particles, a_vel):
    """Update particle velocities to scale factor a_vel

    Args:
        particles (Particles)
        a_vel (float): Scale factor after kick.
    """
    c._leapfrog_kick(particles._particles, a_vel)


def drift(particles, a_pos):
    """Update particle positions to scale factor a_pos

    Args:
        particles (Particles)
        a_pos (float): Scale factor after drift.
    """
    c._leapfrog_drift(particles._particles, a_pos)


def factors(a_i, a_f):
    """(kick factor SphiStd, drift factor SqStd) between a_i and a_f, FP64."""
    return c._leapfrog_factors(a_i, a_f)
