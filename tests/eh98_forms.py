"""Eisenstein & Hu 1998 (ApJ 496, 605) eqs. 2-24, used only to pin the
oracle's no-wiggle fit (eqs. 26-31, oracle/power.py).

These equations are a different part of the same paper: the drag-epoch
sound horizon from the baryon-to-photon ratio (eq. 6), and the full transfer
function with baryon oscillations (eqs. 16-24).  The paper fits eq. 26 to
eq. 6 to within 2%, and builds eqs. 29-31 to follow the full transfer
function with its oscillations averaged out, so a mistake in a constant of
eqs. 26-31 shows as a departure from these forms.

k is in h/Mpc, the sound horizon in Mpc, T_CMB = 2.7255 K.
"""
import numpy as np


def _background(om, ob, h, tcmb):
    th = tcmb / 2.7
    omh2, obh2 = om * h * h, ob * h * h
    zeq = 2.50e4 * omh2 * th**-4                                   # eq 2
    keq = 7.46e-2 * omh2 * th**-2                                  # eq 3, 1/Mpc
    b1 = 0.313 * omh2**-0.419 * (1 + 0.607 * omh2**0.674)          # eq 4
    b2 = 0.238 * omh2**0.223
    zd = 1291 * omh2**0.251 / (1 + 0.659 * omh2**0.828) * (1 + b1 * obh2**b2)
    R = lambda z: 31.5 * obh2 * th**-4 * (z / 1e3) ** -1           # eq 5
    Req, Rd = R(zeq), R(zd)
    s = (2 / (3 * keq) * np.sqrt(6 / Req)                          # eq 6
         * np.log((np.sqrt(1 + Rd) + np.sqrt(Rd + Req)) / (1 + np.sqrt(Req))))
    return zeq, keq, zd, Rd, s


def sound_horizon(om, ob, h, tcmb=2.7255):
    """Eq. 6: the sound horizon at the drag epoch in Mpc."""
    return _background(om, ob, h, tcmb)[4]


def transfer(k, om, ob, h, tcmb=2.7255):
    """Eq. 16: T = (Ob/Om) T_b + (Oc/Om) T_c, with baryon oscillations."""
    zeq, keq, zd, Rd, s = _background(om, ob, h, tcmb)
    omh2, obh2 = om * h * h, ob * h * h
    fb, fc = ob / om, (om - ob) / om
    kk = np.asarray(k, dtype=np.float64) * h                       # 1/Mpc
    ksilk = 1.6 * obh2**0.52 * omh2**0.73 * (1 + (10.4 * omh2) ** -0.95)  # eq 7
    q = kk / (13.41 * keq)                                         # eq 10
    a1 = (46.9 * omh2) ** 0.670 * (1 + (32.1 * omh2) ** -0.532)    # eq 11
    a2 = (12.0 * omh2) ** 0.424 * (1 + (45.0 * omh2) ** -0.582)
    alpha_c = a1**-fb * a2**-(fb**3)
    bb1 = 0.944 / (1 + (458 * omh2) ** -0.708)                     # eq 12
    bb2 = (0.395 * omh2) ** -0.0266
    beta_c = 1 / (1 + bb1 * (fc**bb2 - 1))

    def T0(a, b):                                                  # eqs 19-20
        C = 14.2 / a + 386 / (1 + 69.9 * q**1.08)
        L = np.log(np.e + 1.8 * b * q)
        return L / (L + C * q * q)

    f = 1 / (1 + (kk * s / 5.4) ** 4)                              # eq 18
    Tc = f * T0(1, beta_c) + (1 - f) * T0(alpha_c, beta_c)         # eq 17
    y = (1 + zeq) / (1 + zd)                                       # eq 15
    G = y * (-6 * np.sqrt(1 + y) + (2 + 3 * y) * np.log((np.sqrt(1 + y) + 1) / (np.sqrt(1 + y) - 1)))
    alpha_b = 2.07 * keq * s * (1 + Rd) ** -0.75 * G               # eq 14
    beta_node = 8.41 * omh2**0.435                                 # eq 23
    s_tilde = s / (1 + (beta_node / (kk * s)) ** 3) ** (1 / 3)     # eq 22
    beta_b = 0.5 + fb + (3 - 2 * fb) * np.sqrt((17.2 * omh2) ** 2 + 1)  # eq 24
    x = kk * s_tilde
    j0 = np.sin(x) / x
    Tb = (T0(1, 1) / (1 + (kk * s / 5.2) ** 2)                     # eq 21
          + alpha_b / (1 + (beta_b / (kk * s)) ** 3) * np.exp(-(kk / ksilk) ** 1.4)) * j0
    return fb * Tb + fc * Tc
