"""Counter-based white noise (test oracle only).

Philox4x32-10 as published by Salmon, Moraes, Dror & Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11 (the Random123 library): a 10-round
Feistel-like network with multipliers 0xD2511F53, 0xCD9E8D57 and Weyl key
increments 0x9E3779B9, 0xBB67AE85.

White-noise layout used by the build (no reference counterpart: the
reference draws N-GenIC's GSL ranlxd1 stream, py/lpt.cpp:235-329, which is
replaced by seeded white noise): the real cell with linear index
i = (ix*n + iy)*n + iz takes normal number i of the stream, where call
c = i // 4 of Philox with counter (c_lo, c_hi, 0, 0) and key (seed_lo,
seed_hi) gives uniforms u_j = (r_j + 0.5) * 2^-32 and two Box-Muller pairs
  (sqrt(-2 ln u0) cos(2 pi u1), sqrt(-2 ln u0) sin(2 pi u1),
   sqrt(-2 ln u2) cos(2 pi u3), sqrt(-2 ln u2) sin(2 pi u3)).
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (4, N) uint64 array of 32-bit words; key: (2,) or (2, N)."""
    c0, c1, c2, c3 = [np.asarray(c, dtype=np.uint64) & MASK for c in ctr]
    k0 = np.asarray(key[0], dtype=np.uint64) & MASK
    k1 = np.asarray(key[1], dtype=np.uint64) & MASK
    for r in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
        if r < 9:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def white_noise(n, seed):
    """(n,n,n) float64 N(0,1) field laid out as described above."""
    total = n * n * n
    ncall = (total + 3) // 4
    c = np.arange(ncall, dtype=np.uint64)
    seed = int(seed)
    key = (np.uint64(seed & 0xFFFFFFFF), np.uint64((seed >> 32) & 0xFFFFFFFF))
    r0, r1, r2, r3 = philox4x32_10((c & MASK, c >> np.uint64(32),
                                    np.zeros_like(c), np.zeros_like(c)), key)
    s = 2.0 ** -32
    u0 = (r0.astype(np.float64) + 0.5) * s
    u1 = (r1.astype(np.float64) + 0.5) * s
    u2 = (r2.astype(np.float64) + 0.5) * s
    u3 = (r3.astype(np.float64) + 0.5) * s
    ra = np.sqrt(-2.0 * np.log(u0))
    rb = np.sqrt(-2.0 * np.log(u2))
    out = np.empty((ncall, 4), dtype=np.float64)
    out[:, 0] = ra * np.cos(2 * np.pi * u1)
    out[:, 1] = ra * np.sin(2 * np.pi * u1)
    out[:, 2] = rb * np.cos(2 * np.pi * u3)
    out[:, 3] = rb * np.sin(2 * np.pi * u3)
    return out.reshape(-1)[:total].reshape(n, n, n)
