"""Friends-of-friends halo finder (py/fs/fof.py of the reference).

Runs on the host over the final positions, copied from the device through
pinned memory.  As in the reference, the kd-tree build reorders the
particles (x, v, dx1, dx2 and id; not force).
"""
from . import _fs as c


def find_groups(particles, ll, *, quota=32, boxsize3=None, compute_nfof=False):
    """Run the FoF halo finder: particles closer than ll (periodic) are
    friends, and groups are the connected sets of friends.

    Args:
        particles (Particles)
        ll (float): linking length

    Options:
        boxsize3 (sequence of 3 floats): size of the cuboid that steers the
            kd-tree cuts; the bounding box of the particles if None
        quota (int): maximum number of particles in a kd-tree leaf
        compute_nfof (bool): return the group sizes

    Returns:
        nfof (int32 array, np_local) if compute_nfof, else None.  nfof[i] is
        the number of members of the group whose lowest index is i, 0 if i
        is not the lowest index of its group.  particles.fof_group gives the
        group (its lowest index) of every particle.
    """
    return c._fof_find_groups(particles._particles, ll, boxsize3, quota, compute_nfof)
