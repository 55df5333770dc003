"""bench.py end to end on the GPU at a small size: one JSON line with the
fields of the bench contract, a roofline for the dominant kernel, the step
roofline on the compulsory 156 Np + 32 Ng + 32 Nk bytes, per-stage rows that
add up to it, and per-kernel rows whose declared bytes add up to the
issued-bytes figure."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_line_at_small_size():
    print("[case] bench.py 32^3 / 64^3, 2 steps", flush=True)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--np", "32", "--nc", "64",
                        "--box", "64", "--steps", "2", "--warmup", "1", "--legs", "main",
                        "--cpu-baseline", "off"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [ln for ln in p.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["value"] > 0 and d["higher_is_better"] is True and d["vs_baseline"] is None
    assert d["config"]["workload"].startswith("COLA 32^3 particles, 64^3 mesh")
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    assert 0 < r["frac"] < 1 and r["traffic"] is None    # counters exist for the headline only
    ks = d["kernels"]
    assert r["kernel"] in ks
    # per-kernel GB are rounded to 3 decimals: half a MB per launch at most
    step = sum(k["algorithmic_GB"] * k["launches"] for k in ks.values()) / d["steps"] * 1e9
    slack = sum(k["launches"] for k in ks.values()) / d["steps"] * 0.5e6
    sr = d["step_roofline"]
    assert abs(step - sr["issued_bytes"]) <= slack + 1.0
    Np, n = 32**3, 64
    Ng, Nk = n**3, n * n * (n // 2 + 1)
    assert sr["algorithmic_bytes"] == 156 * Np + 32 * Ng + 32 * Nk
    st = d["stage_roofline"]
    assert set(st) == {"paint", "fft", "readout", "kick_drift"}
    assert sum(v["algorithmic_bytes"] for v in st.values()) == sr["algorithmic_bytes"]
    assert all(v["ms"] > 0 and 0 < v["frac"] for v in st.values())
    # the stages hold every kernel of the step but the bookkeeping ones
    stage_ms = sum(v["ms"] for v in st.values())
    kern_ms = sum(k["total_ms"] for k in ks.values()) / d["steps"]
    assert stage_ms <= kern_ms + 1e-3 and stage_ms > 0.8 * kern_ms
