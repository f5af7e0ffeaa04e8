"""bench.py's contract pieces that run without a GPU: the roofline
arithmetic of the JSON line and the one-GPU guard."""
import os
import subprocess
import sys

from conftest import ROOT


def test_roofline_arithmetic():
    import bench
    res = {"walk_ms": 10.0,
           "timings": {"visits": 1000, "leaf_evals": 100, "cell_evals": 800, "wave_steps": 20}}
    r = bench.roofline(res, "fp32c", 50, 12345)
    flops = 12 * 1000 + 44 * 100 + 64 * 800
    assert r["flops_per_launch"] == flops
    assert abs(r["achieved"] - round(flops / 1e-2 / 1e12, 2)) < 1e-12
    assert r["peak"] == bench.FLOP_PEAK_TFS["fp32c"]
    assert abs(r["frac"] - round(r["achieved"] / r["peak"], 4)) <= 1e-4
    assert r["traffic"] == 12345 and r["unit"] == "TFLOP/s"
    assert r["lane_utilisation"] == round(1000 / (64.0 * 20), 4)
    # 48 B per target, FP32 values of 4 B: 5 per visit, 15 per cell, 6 per leaf
    assert r["hbm_view"]["bytes_alg_per_launch"] == 50 * 48 + 4 * (5 * 1000 + 15 * 800 + 6 * 100)
    r64 = bench.roofline(res, "fp64", 50, None)
    assert r64["hbm_view"]["bytes_alg_per_launch"] == 50 * 48 + 8 * (5 * 1000 + 15 * 800 + 6 * 100)
    assert r64["peak"] == bench.FLOP_PEAK_TFS["fp64"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
