"""Writes the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

Each .npz holds the inputs of one case and the oracle's outputs; the HIP
path is checked against them by tests/test_gpu_golden.py, and the oracle
itself by tests/test_golden.py (it must keep reproducing them).  Arrays
that the GPU side compares at FP32 precision (positions, displacements,
velocities) are stored as float32; meshes, forces and coefficients as
float64.

  cic.npz         CIC density, uniform 16^3 (with out-of-box and face
                  positions, py/util.h:14-24 wrap) and clustered 32^3 cases,
                  pm_factor 2 (py/pm.cpp:57-139)
  pm_force.npz    PM forces at the particles (py/pm.cpp:57-220, 414-508):
                  a perturbed 8^3 lattice on a 16^3 mesh and a plane-wave
                  displacement of the same lattice
  lpt.npz         Philox white noise (16^3, seed 7) and the 2LPT state it
                  gives at a = 0.1 in a 64 Mpc/h box (py/lpt.cpp:100-225,
                  439-591), EH98 no-wiggle P(k)
  cosmology.npz   D1, D2, f at seven scale factors and the COLA kick / drift
                  coefficients of the 10-step a = 0.1 -> 1 schedule
                  (py/cosmology.cpp:24-171, py/cola.cpp:48-179,
                  script/cola.py:37-44), Omega_m in {0.308, 0.31}
  cola_2step.npz  a 2-step COLA run (a = 0.1 -> 0.3) of 16^3 particles on a
                  32^3 mesh from the 2LPT state of seed 11
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import cic as ocic          # noqa: E402
from oracle import cola as ocola        # noqa: E402
from oracle import cosmology as ocosmo  # noqa: E402
from oracle import lpt as olpt          # noqa: E402
from oracle import pm as opm            # noqa: E402
from oracle import power as opower      # noqa: E402
from oracle import rng as orng          # noqa: E402

EH = dict(omega_m=0.31, omega_b=0.049, h=0.67, ns=0.96, sigma8=0.81)
A_GRID = np.array([0.05, 0.1, 0.2, 0.3, 0.5, 0.8, 1.0])
OMEGAS = (0.308, 0.31)


def cic_cases():
    """Positions exactly as handed to Particles.append (float32) and the
    oracle's density of their wrapped copies (FP32 CIC weights, as the
    reference computes them)."""
    rng = np.random.default_rng(2024)
    out = {}
    # uniform 16^3: 512 particles (Np pm^3 = Ng), a quarter outside the box
    n, L = 16, 100.0
    x = rng.uniform(-0.3 * L, 1.3 * L, (512, 3)).astype(np.float32)
    x[0] = 0.0                                   # lower box corner
    x[1] = L                                     # wraps to 0
    x[2] = np.nextafter(np.float32(L), np.float32(0))   # last cell, weight ~1 on it
    x[3] = [L / n * 3, L / n * 7.5, L / n * 15]  # on cell faces and a cell centre
    x[4] = -L / n * 0.5                          # wraps into the last cell's centre
    out["uniform16"] = (n, L, x)
    # clustered 32^3: 4096 particles, half in five tight clumps
    n, L = 32, 50.0
    u = rng.uniform(0, L, (2048, 3))
    centres = rng.uniform(0, L, (5, 3))
    c = centres[rng.integers(0, 5, 2048)]
    clump = np.mod(c + 0.01 * L * rng.standard_normal(c.shape), L)
    x = np.concatenate([u, clump]).astype(np.float32)
    out["clustered32"] = (n, L, x)
    arrays = {}
    for name, (n, L, x) in out.items():
        xw = ocic.wrap(x, L)
        arrays[f"{name}_n"] = np.int64(n)
        arrays[f"{name}_boxsize"] = np.float64(L)
        arrays[f"{name}_pm_factor"] = np.float64(2.0)
        arrays[f"{name}_x"] = x
        arrays[f"{name}_delta"] = ocic.paint(xw, n, L, 2.0, wdtype=np.float32)
    return arrays


def force_cases():
    """Wrapped float32 positions and FP64 forces (opm.force) at them."""
    arrays = {}
    n_p, nc, L = 8, 16, 64.0
    i = np.arange(n_p)
    q = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    q = (q + 0.5) * (L / n_p)
    rng = np.random.default_rng(99)
    cases = {
        "lattice8": q + 0.6 * (L / n_p) * rng.standard_normal(q.shape),
        # x_0 -> x_0 + eps sin(2 pi x_0 / L): a single density mode along x
        "planewave8": q + np.stack([0.3 * (L / n_p) * np.sin(2 * np.pi * q[:, 0] / L),
                                    np.zeros(len(q)), np.zeros(len(q))], 1),
    }
    for name, x in cases.items():
        xw = ocic.wrap(x.astype(np.float32), L)
        arrays[f"{name}_nc"] = np.int64(nc)
        arrays[f"{name}_boxsize"] = np.float64(L)
        arrays[f"{name}_pm_factor"] = np.float64(nc / n_p)
        arrays[f"{name}_x"] = xw
        arrays[f"{name}_force"] = opm.force(xw.astype(np.float64), nc, L, nc / n_p)
    return arrays


def lpt_case(cosmo, eh):
    n, L, a, seed = 16, 64.0, 0.1, 7
    noise = orng.white_noise(n, seed)
    ref = olpt.particles(noise, L, a, eh.P, cosmo, kind="2lpt")
    return dict(n=np.int64(n), boxsize=np.float64(L), a=np.float64(a), seed=np.int64(seed),
                noise=noise.astype(np.float32),
                dx1=ref["dx1"].astype(np.float32), dx2=ref["dx2"].astype(np.float32),
                x=ref["x"].astype(np.float32), v=ref["v"].astype(np.float32),
                id=ref["id"], **{f"eh_{k}": np.float64(v) for k, v in EH.items()})


def cosmology_table():
    arrays = {"a": A_GRID, "omega_m": np.array(OMEGAS)}
    sched = ocola.schedule(0.1, 1.0, 10)
    arrays["schedule"] = np.array(sched)
    for om in OMEGAS:
        c = ocosmo.Cosmology(om)
        tag = f"om{int(round(om * 1000))}"
        D = np.array([c.D(a) for a in A_GRID])
        arrays[f"{tag}_D1"] = D
        arrays[f"{tag}_D2"] = np.array([c.D2(a, d) for a, d in zip(A_GRID, D)])
        arrays[f"{tag}_f"] = np.array([c.f(a) for a in A_GRID])
        kick, drift = [], []
        a_x = a_v = 0.1
        for a_vel, a_pos in sched:
            kick.append(ocola.kick_coefficients(c, a_v, a_x, a_vel))
            a_v = a_vel
            drift.append(ocola.drift_coefficients(c, a_x, a_v, a_pos))
            a_x = a_pos
        arrays[f"{tag}_kick"] = np.array(kick)     # (kf, q1, q2) per step
        arrays[f"{tag}_drift"] = np.array(drift)   # (dt, da1, da2) per step
    return arrays


def cola_2step(cosmo, eh):
    n_p, nc, L, a0, a1, nstep, seed = 16, 32, 64.0, 0.1, 0.3, 2, 11
    ref = olpt.particles(orng.white_noise(n_p, seed), L, a0, eh.P, cosmo, kind="cola")
    x, v, dx1, dx2 = ref["x"], ref["v"], ref["dx1"], ref["dx2"]
    a_x = a_v = a0
    sched = ocola.schedule(a0, a1, nstep)
    for a_vel, a_pos in sched:
        xw = np.mod(x, L)
        f = opm.force(xw, nc, L, nc / n_p)
        v = ocola.kick(v, f, dx1, dx2, cosmo, a_v, a_x, a_vel)
        a_v = a_vel
        x = ocola.drift(xw, v, dx1, dx2, cosmo, a_x, a_v, a_pos)
        a_x = a_pos
    return dict(n_p=np.int64(n_p), nc=np.int64(nc), boxsize=np.float64(L), a0=np.float64(a0),
                a1=np.float64(a1), nstep=np.int64(nstep), seed=np.int64(seed),
                omega_m=np.float64(cosmo.om), schedule=np.array(sched),
                x=x.astype(np.float32), v=v.astype(np.float32))


def build_all():
    cosmo = ocosmo.Cosmology(EH["omega_m"])
    eh = opower.EH98NoWiggle(**EH)
    return {
        "cic": cic_cases(),
        "pm_force": force_cases(),
        "lpt": lpt_case(cosmo, eh),
        "cosmology": cosmology_table(),
        "cola_2step": cola_2step(cosmo, eh),
    }


def main():
    for name, arrays in build_all().items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
