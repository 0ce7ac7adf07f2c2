"""Host-side parts of bench.py that need no GPU: the roofline record of an
engine kernel (fields, units, algorithmic work, the attainable non-FMA
ceiling beside the published peak, traffic passed through) and the one-GPU
refusal of --gpus N > 1."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_kernel_roofline_record():
    r = bench.kernel_roofline("k_eval_a", 20e-6, 200000, 4.0, traffic=123456, traffic_source="x")
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel", "avg_launch_us",
                "attainable_peak", "frac_of_attainable"):
        assert key in r
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s" and r["peak"] == bench.PEAK_FP32_TFLOPS
    flops = bench.eval_flops_per_proposal(2000, 2000, 1000, 3) * 4.0
    # G(r) of both constraints, the S(Q) product-sum, deviations and sums
    assert flops == 4.0 * (12 * 4000 + 2.0 * 2000 * 1000 + 4.0 * 3000)
    assert abs(r["achieved"] - flops / 20e-6 / 1e12) < 1e-3
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-5
    assert abs(r["frac_of_attainable"] - r["achieved"] / 39.3) < 1e-5
    assert r["traffic"] == 123456
    b2 = bench.kernel_roofline("k_eval_a", 20e-6, 13500, 33.0, shape=(1000, 1000, 500, 3))
    assert b2["flops_per_launch"] == bench.eval_flops_per_proposal(1000, 1000, 500, 3) * 33.0
    d = bench.kernel_roofline("k_delta", 10e-6, 13500, 20.0)
    assert abs(d["flops_per_launch"] - bench.OPS_PER_PAIR * 2.0 * 13500 * 20.0) < 1e-6
    assert bench.kernel_roofline("k_eval_a", 0.0, 13500, 33.0) is None
    assert bench.kernel_roofline("k_commit", 1e-5, 13500, 1.0) is None


def test_committed_traffic_scales_the_per_proposal_bytes():
    for prof in ("hbm_traffic_engine_sio2_200k.json", "hbm_traffic_engine_b2_13500.json"):
        path = os.path.join(ROOT, "profiles", prof)
        if not os.path.exists(path):
            continue
        with open(path) as f:
            rec = json.load(f)["kernels"]["k_eval_a"]
        if "traffic_bytes_per_proposal" not in rec or "how" not in rec:
            continue
        t, src = bench.committed_traffic(prof, "k_eval_a", 12)
        assert t == int(round(float(rec["traffic_bytes_per_proposal"]) * 12)) and prof in src
    assert bench.committed_traffic("no_such_profile.json", "k_eval_a", 12) == (None, None)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr) or "1 GPU" in (p.stdout + p.stderr)
