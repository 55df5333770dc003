"""Background cosmology (FP64, host), py/fs/cosmology.py."""
from . import _fs as c


def init(omega_m):
    """Initialise the cosmology module with Omega_m (flat LCDM, H0 = 1)."""
    c._cosmology_init(omega_m)


def D_growth(a):
    """Linear growth factor D(a), normalised to D(1) = 1."""
    return c._cosmology_D_growth(a)


def D2_growth(a):
    """2nd-order growth factor D2(a) = -3/7 D^2 Omega(a)^(-1/143)."""
    return c._cosmology_D2_growth(a)


def f_growth_rate(a):
    """Growth rate f = dlnD/dlna (no fs.cosmology counterpart)."""
    return c._cosmology_f_growth_rate(a)
