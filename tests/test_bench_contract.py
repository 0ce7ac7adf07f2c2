"""Host-side parts of bench.py that need no GPU: the roofline record of an
engine kernel (fields, units, traffic scaled from the committed per-proposal
PMC figure) and the one-GPU refusal of --gpus N > 1."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_kernel_roofline_record():
    r = bench.kernel_roofline("k_eval_a", 20e-6, 13500, 33.0)
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel", "avg_launch_us"):
        assert key in r
    assert r["unit"] == "TFLOP/s" and r["peak"] == bench.PEAK_FP32_TFLOPS
    flops = bench.eval_flops_per_proposal() * 33.0
    assert abs(r["achieved"] - flops / 20e-6 / 1e12) < 1e-3
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-5
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_engine_b2_13500.json")) as f:
        per = json.load(f)["kernels"]["k_eval_a"]["traffic_bytes_per_proposal"]
    assert r["traffic"] == int(round(per * 33.0))
    d = bench.kernel_roofline("k_delta", 10e-6, 13500, 20.0)
    assert abs(d["flops_per_launch"] - bench.OPS_PER_PAIR * 2.0 * 13500 * 20.0) < 1e-6
    assert bench.kernel_roofline("k_eval_a", 0.0, 13500, 33.0) is None


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr) or "1 GPU" in (p.stdout + p.stderr)
