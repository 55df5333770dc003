"""Named memory allocations (py/fs/mem.py of the reference); the memory is
device memory here."""
from . import _fs as c


class Mem:
    """Mem(name, size=0): `size` bytes (rounded up to 16) reserved under
    `name`; MemoryError if they are not available.

    Attributes: name, mem_allocated, mem_using (bytes).
    """

    def __init__(self, name, size=0):
        self._mem = c._mem_alloc(name, size)
        self.name = name
        self.mem_allocated, self.mem_using = c._mem(self._mem)

    def __repr__(self):
        return "Mem %s, using %d/%d bytes" % (self.name, self.mem_using, self.mem_allocated)
