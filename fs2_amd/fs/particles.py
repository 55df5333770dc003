"""Particles, py/fs/particles.py.

State lives in HBM as FP32 planes; the accessors return host copies in the
reference's layouts (x/dx1/dx2/force as (np, 3) float32, id as uint64,
indexing as rows of the 56-byte Particle record viewed as 14 float32).
"""
from . import _fs as c


class Particles(object):
    def __init__(self, np=10, boxsize=0.0, *, _particles=None):
        if _particles is None:
            self._particles = c._particles_alloc(np, boxsize)
        else:
            self._particles = _particles

    def __getitem__(self, index):
        """particles[i] -> (x, y, z); particles[i:j:k] -> (n, 14) float32 rows:
        a[0:3] x, a[3:6] v, a[6:9] dx1, a[9:12] dx2, a[12:14] id bytes."""
        return c._particles_getitem(self._particles, index)

    def __len__(self):
        return c._particles_len(self._particles)

    def update_np_total(self):
        c._particles_update_np_total(self._particles)

    def resize(self, np_local):
        """Set the local number of particles to np_local (<= the allocation)."""
        c._particles_resize(self._particles, np_local)

    def append(self, x):
        """Append particles with positions x ((n, >=3) float64; None: no-op)."""
        c._particles_append(self._particles, x)

    def clear(self):
        c._particles_clear(self._particles)

    def slice(self, frac):
        return c._particles_slice(self._particles, frac)

    def periodic_wrapup(self):
        c._particles_periodic_wrapup(self._particles)

    def save_gadget_binary(self, filename, use_longid=False):
        """Writes a Gadget-1 snapshot: 256-byte header, positions,
        velocities / sqrt(a), then ids as uint32, or uint64 when use_longid
        (py/fs/particles.py:83-89)."""
        c._write_gadget_binary(self._particles, filename, use_longid)

    def save_hdf5(self, filename, var):
        """Writes an HDF5 snapshot with /parameters and one dataset per letter
        of var (py/fs/particles.py:91-104): 'i' ids, 'x' positions, 'v'
        velocities ('c' writes the COLA-corrected velocity as v), 'f' forces,
        '1' and '2' the first- and second-order displacements at a = 1."""
        c._hdf5_write_particles(self._particles, filename, var)

    def copy_from(self, other):
        """Device-to-device copy of all state (same capacity)."""
        c._particles_copy(self._particles, other._particles)

    @property
    def np_local(self):
        return c._particles_len(self._particles)

    @property
    def np_total(self):
        return c._particles_np_total(self._particles)

    @property
    def id(self):
        return c._particles_id_asarray(self._particles)

    @property
    def x(self):
        return c._particles_x_asarray(self._particles, 0)

    @property
    def dx1(self):
        return c._particles_x_asarray(self._particles, 1)

    @property
    def dx2(self):
        return c._particles_x_asarray(self._particles, 2)

    @property
    def v(self):
        return c._particles_v_asarray(self._particles)

    @property
    def force(self):
        return c._particles_force_asarray(self._particles)

    @property
    def fof_group(self):
        """fof_group[i] is the group (its lowest particle index) of particle i
        after the last fs.fof.find_groups; None before any run."""
        return c._fof_grp()

    @property
    def times(self):
        """(a_x, a_v, a_f)"""
        return c._particles_times(self._particles)

    def set_times(self, a_x, a_v, a_f):
        c._particles_set_times(self._particles, a_x, a_v, a_f)

    def set_field(self, name, values):
        field = {'x': 0, 'dx1': 1, 'dx2': 2, 'v': 3, 'force': 4}[name]
        c._particles_set_field(self._particles, field, values)

    def device_fields(self):
        """Device pointers of the FP32 planes [3][stride] and the id array."""
        return c._particles_device_fields(self._particles)
