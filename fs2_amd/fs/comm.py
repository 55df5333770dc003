"""One rank per GPU; the slab decomposition of the reference collapses
(py/fs/comm.py)."""
from . import _fs as c


def this_node():
    """Returns: this_node (int): rank of this process (always 0)."""
    return c.comm_this_node()


def n_nodes():
    """Returns: n_nodes (int): number of ranks (always 1)."""
    return c.comm_n_nodes()
