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


def test_stage_bytes_with_fused_kernels():
    """With the fused paint + z r2c and z c2r + readout kernels in the
    profile, the paint, the FFT passes and the readout form one stage with
    their compulsory bytes together; the stages still sum to 156 Np + 32 Ng
    + 32 Nk."""
    Np, n = 512**3, 1024
    Ng, Nk = n**3, n * n * (n // 2 + 1)
    steps = 2
    prof = [("paint_bin", steps, 1.0, 0.0), ("paint_scatter", steps, 2.0, 0.0),
            ("paint_tile_r2c", steps, 6.0, 0.0), ("fft_y_fwd", steps, 4.0, 0.0),
            ("fft_x_fwd_green3", steps, 8.0, 0.0), ("readout_columns", steps, 14.0, 0.0),
            ("cola_kick_drift", steps, 4.0, 0.0)]
    st = bench.stage_roofline(prof, steps, Np, n)
    assert set(st) == {"paint+fft+readout", "kick_drift"}
    assert sum(v["algorithmic_bytes"] for v in st.values()) == 156 * Np + 32 * Ng + 32 * Nk
    assert st["paint+fft+readout"]["algorithmic_bytes"] == 36 * Np + 32 * Ng + 32 * Nk
    assert st["paint+fft+readout"]["ms"] == 17.5 and st["kick_drift"]["ms"] == 2.0
    frac = (36 * Np + 32 * Ng + 32 * Nk) / 17.5e-3 / 8e12
    assert abs(st["paint+fft+readout"]["frac"] - frac) < 1e-4
    # unfused: the four stages of the reference's structure
    st = bench.stage_roofline([r for r in prof if r[0] not in ("paint_tile_r2c", "readout_columns")]
                              + [("paint_tile_det", steps, 4.0, 0.0), ("fft_z_r2c", steps, 4.0, 0.0),
                                 ("fft_z_c2r", 3 * steps, 12.0, 0.0),
                                 ("readout_tile", steps, 8.0, 0.0)], steps, Np, n)
    assert set(st) == {"paint", "fft", "readout", "kick_drift"}
    assert st["fft"]["algorithmic_bytes"] == 16 * Ng + 32 * Nk
