"""Linear power spectrum, py/fs/power.py."""
from . import _fs as c


class PowerSpectrum:
    """ps = PowerSpectrum(filename): tabulated "k P" text file (CAMB format).

    Raises:
        IOError: if the file cannot be read.
    """

    def __init__(self, filename=None, *, _ps=None):
        if _ps is not None:
            self._ps = _ps
        else:
            self._ps = c._power_alloc(filename)

    @classmethod
    def eisenstein_hu(cls, omega_m=0.31, omega_b=0.049, h=0.67, ns=0.96,
                      sigma8=0.81):
        """Eisenstein & Hu (1998) no-wiggle P(k) normalised to sigma8."""
        return cls(_ps=c._power_eh98(omega_m, omega_b, h, ns, sigma8))

    def __len__(self):
        return c._power_n(self._ps)

    def __getitem__(self, i):
        """ith pair (k, P); IndexError out of range."""
        return c._power_i(self._ps, i)

    def P(self, k):
        return c._power_P(self._ps, k)

    def sigma(self, R):
        return c._power_sigma(self._ps, R)
