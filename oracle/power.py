"""Linear power spectra (test oracle only).

TablePower follows py/power.cpp:15-144: a "k P" text table ('#' comments,
k<=0 rows skipped, k must increase, parsing stops at the first malformed
line), interpolated by a natural cubic spline of ln P in ln k (GSL
gsl_interp_cspline, power.cpp:31-38, 50-54), here scipy's CubicSpline with
bc_type='natural'.  sigma() is the trapezoid of power.cpp:146-171.

EH98NoWiggle is Eisenstein & Hu 1998 (ApJ 496, 605) eqs. 26-31, the
zero-baryon-oscillation fit with the baryon-suppressed shape parameter,
P(k) = A k^ns T0(q)^2, normalised to sigma_8 with a real-space top hat of
R = 8 Mpc/h.  T_CMB = 2.7255 K.  k in h/Mpc, P in (Mpc/h)^3.
"""
import numpy as np
from scipy import integrate, interpolate


def read_table(path):
    ks, Ps = [], []
    k_prev = 0.0
    with open(path) as fp:
        for line in fp:
            if line.startswith('#'):
                continue
            parts = line.split()
            try:
                k, P = float(parts[0]), float(parts[1])
            except (ValueError, IndexError):
                break
            if k <= 0.0:
                continue
            if k < k_prev:
                raise IOError('k must be increasing')
            ks.append(k)
            Ps.append(P)
            k_prev = k
    return np.array(ks), np.array(Ps)


class TablePower:
    def __init__(self, path=None, k=None, P=None):
        if path is not None:
            k, P = read_table(path)
        self.k = np.asarray(k, dtype=np.float64)
        self.Ptab = np.asarray(P, dtype=np.float64)
        self.spline = interpolate.CubicSpline(np.log(self.k), np.log(self.Ptab),
                                              bc_type='natural')

    def P(self, k):
        return np.exp(self.spline(np.log(k)))

    def sigma(self, R):
        """Trapezoid over the table nodes, power.cpp:146-171."""
        fac = 1.0 / (2.0 * np.pi**2)
        k0, f0, s2 = self.k[0], 0.0, 0.0
        for k, P in zip(self.k, self.Ptab):
            x = k * R
            w = 3.0 * (np.sin(x) - x * np.cos(x)) / x**3
            f1 = P * k * k * w * w
            s2 += 0.5 * (f0 + f1) * (k - k0)
            k0, f0 = k, f1
        return np.sqrt(fac * s2)


# The numerical constants of Eisenstein & Hu 1998 eqs. 26-31, by equation.
# Tests plant mistakes in a copy of this table to show that the pins fail.
EH98_CONSTANTS = {
    's': (44.5, 9.83, 10.0, 0.75),             # eq 26: s = a ln(b/Om h^2)/sqrt(1 + c (Ob h^2)^d)
    'alpha': (0.328, 431.0, 0.38, 22.3),       # eq 31
    'gamma': (0.43, 4.0),                      # eq 30
    'T0': (2.0, 1.8, 14.2, 731.0, 62.5),       # eq 29
}


def eh98_nowiggle_T(k, omega_m, omega_b, h, tcmb=2.7255, constants=None):
    c = EH98_CONSTANTS if constants is None else constants
    s1, s2, s3, s4 = c['s']
    a1, a2, a3, a4 = c['alpha']
    g1, g2 = c['gamma']
    t1, t2, t3, t4, t5 = c['T0']
    theta = tcmb / 2.7
    omh2 = omega_m * h * h
    obh2 = omega_b * h * h
    fb = omega_b / omega_m
    s = s1 * np.log(s2 / omh2) / np.sqrt(1.0 + s3 * obh2**s4)                  # eq 26
    alpha = 1.0 - a1 * np.log(a2 * omh2) * fb + a3 * np.log(a4 * omh2) * fb * fb  # eq 31
    kmpc = k * h
    gamma = omega_m * h * (alpha + (1.0 - alpha) / (1.0 + (g1 * kmpc * s) ** g2))  # eq 30
    q = k * theta * theta / gamma                                              # eq 28
    L0 = np.log(t1 * np.e + t2 * q)                                            # eq 29
    C0 = t3 + t4 / (1.0 + t5 * q)
    return L0 / (L0 + C0 * q * q)


def eh98_sound_horizon(omega_m, omega_b, h, constants=None):
    """Eq. 26, the fit to the sound horizon at the drag epoch, in Mpc."""
    s1, s2, s3, s4 = (EH98_CONSTANTS if constants is None else constants)['s']
    omh2, obh2 = omega_m * h * h, omega_b * h * h
    return s1 * np.log(s2 / omh2) / np.sqrt(1.0 + s3 * obh2**s4)


def tophat_sigma2(Pfun, R):
    """sigma^2(R) = 1/(2 pi^2) int P(k) W(kR)^2 k^2 dk, integrated in ln k."""
    def integrand(lnk):
        k = np.exp(lnk)
        x = np.asarray(k * R, dtype=np.float64)
        xs = np.maximum(x, 1e-3)
        w = np.where(x > 1e-3, 3.0 * (np.sin(xs) - xs * np.cos(xs)) / xs**3, 1.0 - x * x / 10.0)
        return Pfun(k) * w * w * k**3
    lo, hi = np.log(1e-6), np.log(1e3)
    edges = np.linspace(lo, hi, 161)
    rough = sum(abs(integrate.fixed_quad(integrand, a, b, n=32)[0])
                for a, b in zip(edges[:-1], edges[1:]))
    tot = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        v, _ = integrate.quad(integrand, a, b, epsabs=1e-15 * rough, epsrel=1e-12, limit=200)
        tot += v
    return tot / (2.0 * np.pi**2)


class EH98NoWiggle:
    def __init__(self, omega_m=0.31, omega_b=0.049, h=0.67, ns=0.96, sigma8=0.81,
                 tcmb=2.7255, constants=None):
        self.omega_m, self.omega_b, self.h, self.ns, self.tcmb = omega_m, omega_b, h, ns, tcmb
        self.constants = constants
        self.A = 1.0
        s2 = tophat_sigma2(self._unnorm, 8.0)
        self.A = sigma8**2 / s2

    def _unnorm(self, k):
        T = eh98_nowiggle_T(k, self.omega_m, self.omega_b, self.h, self.tcmb, self.constants)
        return k**self.ns * T * T

    def P(self, k):
        k = np.asarray(k, dtype=np.float64)
        return self.A * self._unnorm(k)
