"""COLA time stepping, py/fs/cola.py."""
from . import _fs as c


def kick(particles, a_vel):
    """Update particle velocities to scale factor a_vel."""
    c._cola_kick(particles._particles, a_vel)


def drift(particles, a_pos):
    """Update particle positions to scale factor a_pos."""
    c._cola_drift(particles._particles, a_pos)


def velocity(particles):
    """Physical COLA velocity v + Dv dx1 + D2v dx2 (py/cola.cpp:129-148)."""
    return c._cola_velocity(particles._particles)
