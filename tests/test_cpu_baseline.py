"""The CPU baseline (oracle/cpu_step.cpp, the reference's step loops on
56-byte AoS records) computes the reference's step: its FP32 forces match
the FP64 oracle, its short trajectory follows the oracle's, and thread
counts change no position by more than FP32 atomic-order rounding."""
import numpy as np

from oracle import cola as ocola
from oracle import cosmology as ocosmo
from oracle import cpu_baseline as cb
from oracle import lpt as olpt
from oracle import pm as opm
from oracle import power as opower
from oracle import rng as orng


def _state(n_p, L):
    c = ocosmo.Cosmology(0.31)
    st = olpt.particles(orng.white_noise(n_p, 21), L, 0.1, opower.EH98NoWiggle().P, c, "cola")
    f32 = lambda a: a.astype(np.float32)
    return c, st, cb.records(f32(st["x"]), f32(st["v"]), f32(st["dx1"]), f32(st["dx2"]), st["id"])


def test_record_layout_is_the_reference_particle():
    r = np.zeros(1, dtype=cb.RECORD)
    assert cb.RECORD.itemsize == 56
    assert [cb.RECORD.fields[k][1] for k in ("x", "v", "dx1", "dx2", "id")] == [0, 12, 24, 36, 48]


def test_cpu_step_follows_fp64_oracle():
    n_p, nc, L = 16, 32, 32.0
    c, st, rec = _state(n_p, L)
    x, v, dx1, dx2 = st["x"], st["v"], st["dx1"], st["dx2"]
    outs = []
    for threads in (1, 4):
        sim = cb.CpuCola(rec.copy(), nc, L, nc / n_p, 0.31, 0.1, threads)
        a_x = a_v = 0.1
        xo, vo = x, v
        for i, (a_vel, a_pos) in enumerate(ocola.schedule(0.1, 1.0, 10)[:3]):
            sim.step(a_vel, a_pos)
            xw = np.mod(xo, L)
            f = opm.force(xw, nc, L, nc / n_p)
            if i == 0:
                rel = np.sqrt(np.mean((sim.force - f) ** 2) / np.mean(f ** 2))
                assert rel < 1e-5, rel
            vo = ocola.kick(vo, f, dx1, dx2, c, a_v, a_x, a_vel)
            a_v = a_vel
            xo = ocola.drift(xw, vo, dx1, dx2, c, a_x, a_v, a_pos)
            a_x = a_pos
        d = sim.positions() - xo
        d -= L * np.round(d / L)
        assert np.abs(d).max() / (L / nc) < 1e-3
        outs.append(sim.positions())
    d = outs[0] - outs[1]
    d -= L * np.round(d / L)
    assert np.abs(d).max() / (L / nc) < 1e-4
