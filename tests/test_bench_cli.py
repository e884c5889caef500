"""bench.py's command line on the CPU: one GPU only, and the CPU baseline leg
(the reference driver when it was built, else the C oracle) on a small grid
reporting a positive solve time scaled to the requested iteration count."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                           *args], capture_output=True, text=True,
                          timeout=timeout, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    out = _run("--gpus", "2")
    assert out.returncode == 2
    assert "one GPU" in out.stderr


def test_cpu_baseline_small_grid():
    out = _run("--cpu-only", "7", "--shape", "24,16,12")
    assert out.returncode == 0, out.stderr[-2000:]
    line = json.loads(out.stdout.strip().splitlines()[-1])
    cpu = line["cpu_baseline"]
    assert line["workload"] == "flame_step 24x16x12"
    assert cpu["unit"] == "ms" and cpu["cores"] == 1
    assert cpu["kind"] in ("reference", "port")
    assert cpu["value"] > 0.0
    if cpu["kind"] == "reference":
        # start + 7 x per-iteration of the reference, the port timed beside it
        assert "7 FP64 iterations" in cpu["sample"]
        assert cpu["port_seconds_per_iteration"] > 0.0
