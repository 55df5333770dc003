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
    r = bench.cpu_baseline(16, 32, 64.0, 2, 1)
    assert set(r) == {"value", "unit", "cores", "kind", "sample"}
    assert r["value"] > 0 and r["unit"] == "particle-steps/s"
    assert r["cores"] == 1 and r["kind"] == "port"
    assert "32^3 mesh" in r["sample"] and "1/2 of the 16^3" in r["sample"]


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""
