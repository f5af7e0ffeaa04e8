"""bench.py's contract pieces that run without a GPU: the roofline
arithmetic of the JSON line and the one-GPU guard."""
import os
import subprocess
import sys

from conftest import ROOT


def test_roofline_arithmetic(tmp_path, monkeypatch):
    import bench
    res = {"walk_ms": 10.0,
           "timings": {"visits": 1000, "leaf_evals": 100, "cell_evals": 800, "wave_steps": 20}}
    r = bench.roofline(res, "fp32c", 50)
    flops = 12 * 1000 + 44 * 100 + 64 * 800
    assert r["bound"] == "valu"
    assert r["flops_per_launch"] == flops
    assert abs(r["achieved"] - round(flops / 1e-2 / 1e12, 2)) < 1e-12
    assert r["peak"] == bench.FLOP_PEAK_TFS["fp32c"]
    assert abs(r["frac"] - round(r["achieved"] / r["peak"], 4)) <= 1e-4
    assert abs(r["frac_unpacked"] - round(r["achieved"] / 78.6, 4)) <= 1e-4
    assert r["unit"] == "TFLOP/s"
    assert r["lane_utilisation"] == round(1000 / (64.0 * 20), 4)
    # 48 B per target, FP32 values of 4 B: 5 per visit, 15 per cell, 6 per leaf
    assert r["per_lane_bytes"] == 50 * 48 + 4 * (5 * 1000 + 15 * 800 + 6 * 100)
    r64 = bench.roofline(res, "fp64", 50)
    assert r64["per_lane_bytes"] == 50 * 48 + 8 * (5 * 1000 + 15 * 800 + 6 * 100)
    assert r64["peak"] == bench.FLOP_PEAK_TFS["fp64"]


def test_walk_traffic_comes_with_its_source():
    """The walk's HBM bytes come from the committed counter summary, with the
    file, command and workload they were measured on."""
    import bench
    for prec in ("fp32c", "fp64"):
        traffic, source = bench.walk_traffic(prec)
        assert traffic is not None and traffic > 0
        assert os.path.exists(os.path.join(ROOT, "profiles", source["source"]))
        assert "rocprofv3" in source["command"] and source["workload"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
