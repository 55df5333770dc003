"""COLA kick/drift restated from py/cola.cpp:20-179 (test oracle only).

kick  (cola.cpp:48-92):
    kf = (af^n - ai^n) / (n a^n sqrt(Om/a + (1-Om) a^2)),  n = nLPT = -2.5
    q1 = D1(a), q2 = D2(a) - D1(a)^2, a = a_x
    v += kf * (-1.5 Om) (f + q1 dx1 + q2 dx2)
drift (cola.cpp:94-127):
    dt = int_ai^af a^n / (a^3 H(a)) da / av^n        (Sq, cola.cpp:152-179)
    x += v dt + dx1 (D1(af)-D1(ai)) + dx2 (D2(af)-D2(ai))
velocity (cola.cpp:129-148): v + Dv dx1 + D2v(a, D1) dx2 (the reference passes
    D1 where D2v expects D2; kept).
"""
import numpy as np
from scipy import integrate

NLPT = -2.5


def kick_coefficients(cosmo, a_v, a_x, a_vel):
    om = cosmo.om
    ai, a, af = a_v, a_x, a_vel
    kf = (af**NLPT - ai**NLPT) / (NLPT * a**NLPT * np.sqrt(om / a + (1.0 - om) * a * a))
    g1 = cosmo.D(a)
    g2 = cosmo.D2(a, g1)
    return kf, g1, cosmo.D2a(g1, g2)


def drift_dt(cosmo, ai, af, av):
    om = cosmo.om
    fun = lambda a: a**NLPT / (np.sqrt(om / a**3 + 1.0 - om) * a**3)
    val, _ = integrate.quad(fun, ai, af, epsabs=0.0, epsrel=1e-13, limit=200)
    return val / av**NLPT


def drift_coefficients(cosmo, a_x, a_v, a_pos):
    ai, af = a_x, a_pos
    dt = drift_dt(cosmo, ai, af, a_v)
    gi, gf = cosmo.D(ai), cosmo.D(af)
    da1 = gf - gi
    da2 = cosmo.D2(af, gf) - cosmo.D2(ai, gi)
    return dt, da1, da2


def kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel):
    kf, q1, q2 = kick_coefficients(cosmo, a_v, a_x, a_vel)
    acc = -1.5 * cosmo.om * (f + dx1 * q1 + dx2 * q2)
    return v + acc * kf


def drift(x, v, dx1, dx2, cosmo, a_x, a_v, a_pos):
    dt, da1, da2 = drift_coefficients(cosmo, a_x, a_v, a_pos)
    return x + v * dt + (dx1 * da1 + dx2 * da2)


def velocity(v, dx1, dx2, cosmo, a):
    D1 = cosmo.D(a)
    Dv = cosmo.Dv(a, D1)
    D2v = cosmo.D2v(a, D1)   # cola.cpp:138 passes D1, kept
    return v + Dv * dx1 + D2v * dx2


def schedule(a0, a1, nstep):
    """The step schedule of script/cola.py:37-44."""
    out = []
    for i in range(nstep):
        out.append((a0 + (a1 - a0) / nstep * (i + 0.5),
                    a0 + (a1 - a0) / nstep * (i + 1.0)))
    return out
