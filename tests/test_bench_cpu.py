"""bench.py on the CPU: the CPU-baseline leg at a small size returns the
fields the bench line carries, and --gpus N > 1 exits with a message before
touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_cpu_baseline_small():
    """One step of the reference's CPU loops on a 2LPT state (16^3 / 32^3)."""
    import numpy as np
    from oracle import cosmology as ocosmo
    from oracle import cpu_baseline as cb
    from oracle import lpt as olpt
    from oracle import power as opower
    from oracle import rng as orng
    st = olpt.particles(orng.white_noise(16, 3), 64.0, 0.5, opower.EH98NoWiggle().P,
                        ocosmo.Cosmology(0.31), "cola")
    f32 = lambda a: a.astype(np.float32)
    recs = cb.records(f32(st["x"]), f32(st["v"]), f32(st["dx1"]), f32(st["dx2"]), st["id"])
    r = bench.cpu_baseline((recs, 0.5, 0.455), 16, 32, 64.0)
    assert {"value", "unit", "cores", "kind", "sample", "host", "phases_s"} <= set(r)
    assert r["value"] > 0 and r["unit"] == "particle-steps/s"
    assert r["cores"] >= 1 and r["kind"] == "port"
    assert "16^3 particles / 32^3 mesh" in r["sample"]
    assert set(r["phases_s"]) == {"paint", "fft", "green", "readout", "kick_drift"}
    assert r["host"]["nproc"] >= 1


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""
