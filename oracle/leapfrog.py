"""Standard (non-COLA) leapfrog restated from py/leapfrog.cpp:18-141 (test oracle only).

initial velocity (leapfrog.cpp:18-38): v = Dv(a, D1) dx1 + D2v(a, D2) dx2
kick  (leapfrog.cpp:40-68):  v += -1.5 Om f Sphi,  Sphi = int_ai^af da / (a^2 H(a)),
                             ai = a_v, af = a_vel
drift (leapfrog.cpp:70-94):  x += v Sq,  Sq = int_ai^af da / (a^3 H(a)), ai = a_x
H(a) = sqrt(Om/a^3 + 1 - Om) (H0 = 1).  The reference integrates with GSL qag
(Gauss-Kronrod 61, rtol 1e-5); here scipy.integrate.quad to 1e-13.
"""
import numpy as np
from scipy import integrate


def _H(om, a):
    return np.sqrt(om / a**3 + 1.0 - om)


def kick_factor(cosmo, ai, af):
    val, _ = integrate.quad(lambda a: 1.0 / (_H(cosmo.om, a) * a * a), ai, af,
                            epsabs=0.0, epsrel=1e-13, limit=200)
    return val


def drift_factor(cosmo, ai, af):
    val, _ = integrate.quad(lambda a: 1.0 / (_H(cosmo.om, a) * a**3), ai, af,
                            epsabs=0.0, epsrel=1e-13, limit=200)
    return val


def initial_velocity(dx1, dx2, cosmo, a):
    D1 = cosmo.D(a)
    D2 = cosmo.D2(a, D1)
    return cosmo.Dv(a, D1) * dx1 + cosmo.D2v(a, D2) * dx2


def kick(v, f, cosmo, a_v, a_vel):
    return v + (-1.5 * cosmo.om) * f * kick_factor(cosmo, a_v, a_vel)


def drift(x, v, cosmo, a_x, a_pos):
    return x + v * drift_factor(cosmo, a_x, a_pos)
