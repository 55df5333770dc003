"""FP64 restatement of the reference background cosmology (test oracle only).

Follows py/cosmology.cpp:
  * D1(a) = H(a) * int_0^a (a' H(a'))^-3 da', normalised to 1 at a = 1
    (cosmology.cpp:152-170, 29-36);
  * D2(a) = -3/7 D1^2 Omega(a)^(-1/143)              (cosmology.cpp:40-46);
  * Dv = a^2 D H f, D2v = 2 a^2 D2 H f, D2a = D2 - D1^2 (cosmology.cpp:48-69);
  * f = 1/(d_un a^2 H^2) - 1.5 Omega_m/(H^2 a^3)      (cosmology.cpp:71-82);
  * H = sqrt(Omega_m a^-3 + 1 - Omega_m), H0 = 1      (cosmology.cpp:104-108);
  * Omega(a) = Omega_m/(Omega_m + (1-Omega_m) a^3)    (cosmology.cpp:110-115).

The integral uses scipy.integrate.quad (QUADPACK, the Gauss-Kronrod family
GSL qag belongs to) at a tighter tolerance than the reference's 0.5e-8.
"""
import numpy as np
from scipy import integrate, special


class Cosmology:
    def __init__(self, omega_m):
        self.om = float(omega_m)
        self.norm = 1.0 / self.growth_unnormalised(1.0)

    def hubble(self, a):
        return np.sqrt(self.om / a**3 + (1.0 - self.om))

    def omega(self, a):
        return self.om / (self.om + (1.0 - self.om) * a**3)

    def growth_unnormalised(self, a):
        om = self.om
        # cosmology.cpp:142-147 integrand (a H)^-3
        f = lambda x: (om / x + (1.0 - om) * x * x) ** -1.5
        val, _ = integrate.quad(f, 0.0, a, epsabs=0.0, epsrel=1e-13, limit=200)
        return self.hubble(a) * val

    def D(self, a):
        if a == 0.0:
            return 0.0
        return self.norm * self.growth_unnormalised(a)

    def D2(self, a, D=None):
        if a == 0.0:
            return 0.0
        if D is None:
            D = self.D(a)
        return -3.0 / 7.0 * D * D * self.omega(a) ** (-1.0 / 143.0)

    def f(self, a):
        if a == 0.0:
            return 1.0
        d_un = self.growth_unnormalised(a)
        hf = self.hubble(a)
        return 1.0 / (d_un * a * a * hf * hf) - 1.5 * self.om / (hf * hf * a**3)

    def Dv(self, a, D):
        if a == 0.0:
            return 1.0
        return a * a * D * self.hubble(a) * self.f(a)

    def D2v(self, a, D2):
        return 2.0 * a * a * D2 * self.hubble(a) * self.f(a)

    @staticmethod
    def D2a(D1, D2):
        return D2 - D1 * D1

    # --- closed form, used only to pin the quadrature above -------------
    def D_closed_form(self, a):
        """D1 for flat LCDM: a 2F1(1/3, 1; 11/6; -a^3 (1-Om)/Om), normalised."""
        om = self.om
        g = lambda x: x * special.hyp2f1(1.0 / 3.0, 1.0, 11.0 / 6.0,
                                          -(x**3) * (1.0 - om) / om)
        return g(a) / g(1.0)
