"""FFT grid handle, py/fs/fft.py."""
from . import _fs as c


class FFT:
    """3-dimensional grid in real or Fourier space.

    Args:
        nc (int): allocate a new grid with nc points per dimension, or
        _fft: wrap an existing grid capsule.
    """

    def __init__(self, arg):
        if isinstance(arg, int):
            self._fft = c._fft_alloc(arg)
        else:
            self._fft = arg

    def set_test_data(self):
        """grid(ix, iy, iz) = (ix*nc + iy)*nc + iz + 1"""
        c._fft_set_test_data(self._fft)

    def asarray(self):
        """The real grid as an (nc, nc, nc) float32 np.array."""
        return c._fft_fx_global_as_array(self._fft)

    # -- additions (no fs.FFT counterpart) --
    def set_array(self, a):
        c._fft_set_fx(self._fft, a)

    def asarray_k(self):
        """Half-complex view (nc, nc, nc//2+1) complex64, [kx][ky][kz]."""
        return c._fft_fk_as_array(self._fft)

    def set_array_k(self, a):
        c._fft_set_fk(self._fft, a)

    def execute_forward(self):
        """In-place r2c (unnormalised, sign -1), py/fft.cpp:80-89."""
        c._fft_r2c(self._fft)

    def execute_inverse(self):
        """In-place c2r (unnormalised), py/fft.cpp:91-100."""
        c._fft_c2r(self._fft)
