"""Benchmark of the filament velocity stage on one MI355X.

Metric: point-velocities/s at 1M filament points, tree_theta = 0.4, FP32 far
field with compensated sums against the plain FP64 path.  One point-velocity
is the full velocity of one active point (FP64 LIA + the nonlocal sum over
all 27 periodic image shells); a step is one velocity stage: ghost gather +
LIA, octree build (keys, sort, nodes, moments) and the tree walk.

Workload: seeded random loops (radius U[0.005, 0.02] cm, uniform centres and
orientations, spacing 0.8 delta) filling a periodic cube L = 0.2 cm up to
1e6 points, delta = 5e-4 cm, Gamma = 9.97e-4 cm^2/s, a = 1.3e-8 cm,
tree_extra_correction on, second-order derivatives.  Inputs are uploaded to
HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--points N]

Also measured, under "ensemble": 256 independent vortex rings in open
boundaries (point counts U{200..5000}, radius U[0.02, 0.06] cm, helical
perturbation of 1% of R), FP32 pair terms with FP64 LIA, advanced 500 AB
steps in persistent launches replayed from a captured HIP graph, with the
tile shape picked by the library's autotuner.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "point-velocities/s @1M pts tree θ=0.4: FP32+comp vs FP64 GPU, rel err≤1e-5"
HBM_PEAK_GBS = 8000.0


def workload(n_points: int, seed: int = 20261202):
    from qvort_amd import geometry as G
    return G.random_loops(0.2, 5e-4, target_points=n_points, seed=seed, spacing=0.8)


def run_precision(fil, precision: str, theta: float, steps: int, warmup: int):
    from qvort_amd import gpu
    p = gpu.params(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.2,
                   delta=5e-4, corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6,
                   precision=precision)
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        for _ in range(warmup):
            c.velocity_only()
        walk_ms, build_ms, total_ms, gl_ms = [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            c.velocity_only()
            t = c.timings()
            walk_ms.append(t["walk_ms"])
            build_ms.append(t["build_ms"])
            total_ms.append(t["total_ms"])
            gl_ms.append(t["ghost_lia_ms"])
        t1 = time.perf_counter()
        tim = c.timings()
        u, _, _, _ = c.velocity()
        # the AB update is not part of the metric; timed on its own
        ta = time.perf_counter()
        c.ab_step()
        ab_ms = (time.perf_counter() - ta) * 1e3
    return dict(wall_s=t1 - t0, walk_ms=float(np.mean(walk_ms)), build_ms=float(np.mean(build_ms)),
                device_ms=float(np.mean(total_ms)), ghost_lia_ms=float(np.mean(gl_ms)),
                ab_ms=ab_ms, timings=tim, u=u)


def cpu_baseline(fil, theta: float, n_sample: int):
    """The C restatement (oracle, OpenMP) on this host: full octree build of the
    same 1M-point input plus the walk of a strided sample of targets."""
    from oracle import port
    p = port.make_params(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.2,
                         corea=1.3e-8)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    act = np.flatnonzero(fil.infront)
    stride = max(1, act.size // n_sample)
    targets = act[::stride][:n_sample]
    _, _, build_s, walk_s = port.velocity_targets(p, fil, targets, nthreads=nthreads)
    n = act.size
    t_stage = build_s + walk_s * n / targets.size
    return dict(value=n / t_stage, unit="point-velocities/s", cores=nthreads, kind="port",
                sample=(f"oracle (C, OpenMP {nthreads} threads): full octree build of the same "
                        f"{n}-point input ({build_s:.2f} s) + walk of {targets.size} strided "
                        f"targets ({walk_s:.2f} s), stage time extrapolated to all targets"))


def error_vs_direct(fil, theta: float, u32, u64, n_sample: int = 200) -> dict:
    """Part of the CPU leg: the stage's velocities against the FP64 direct sum
    on a strided sample of targets -- the theta -> 0+ limit of the tree (every
    leaf with the walk's exclusions, shells composed as calc_velocity does),
    summed by the C oracle on this host.  This is the truncation error of the
    reference's multipole acceptance (identical in the reference) plus, for
    FP32C, its own rounding."""
    from oracle import port
    p = port.make_params(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.2,
                         corea=1.3e-8)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    act = np.flatnonzero(fil.infront)
    sample = act[::max(1, act.size // n_sample)][:n_sample].astype(np.int32)
    ref, _ = port.leaf_sum_targets(p, fil, sample, nthreads=nthreads)
    nr = np.linalg.norm(ref, axis=1)
    out = {"vs": "fp64 direct sum (theta -> 0+ leaf sum, tree exclusions), C oracle",
           "sample": int(sample.size)}
    for name, u in (("fp32c", u32), ("fp64", u64)):
        e = np.linalg.norm(u[sample] - ref, axis=1) / nr
        out[name] = {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e ** 2)))}
    return out


FLOP_PEAK_TFS = {"fp32c": 157.3, "fp64": 78.6}  # MI355X published fp32 / fp64 vector
UNPACKED_PEAK_TFS = 78.6  # one FMA per lane and cycle, FP32 or FP64 (no packed FP32)


def walk_traffic(precision: str):
    """HBM bytes per launch of the walk from the committed rocprofv3 counter
    summary (FETCH_SIZE / WRITE_SIZE scaled by a plain copy's), with the
    summary's file, command and workload: the counters cannot be read inside
    this run."""
    path = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh).get(f"k_walk_{precision}")
    if not d:
        return None, None
    return d["bytes_per_launch"], {k: d[k] for k in ("source", "command", "workload")}


def roofline(res, precision: str, n: int):
    """The walk kernel against the vector-ALU roofline: the walk issues no
    MFMA (no dense contraction), its arithmetic runs on the vector ALUs.
    Algorithmic flops, counting +,-,*,/,sqrt as one: 12 per node test (per
    target and visited node), 44 per leaf segment, 64 per accepted cell,
    with the counts taken from the kernel's own 64-bit counters.  `peak` is
    MI355X's published vector peak of the precision; FP32's 157.3 TFLOP/s
    counts packed FP32 (two per lane and cycle), which this build does not
    generate, so `frac_unpacked` states the fraction of the 78.6 TFLOP/s that
    non-packed FP32 and FP64 reach.  `per_lane_bytes` counts the node data
    each target's walk reads (48 B per target; per visit 5 values, per cell
    15, per leaf 6, at 4 or 8 bytes): not HBM traffic -- one scalar load of
    a node serves the 64 targets of a wavefront -- which `traffic` gives."""
    t = res["timings"]
    visits, leaves, cells = t["visits"], t["leaf_evals"], t["cell_evals"]
    flops = 12 * visits + 44 * leaves + 64 * cells
    tflops = flops / (res["walk_ms"] * 1e-3) / 1e12
    peak = FLOP_PEAK_TFS[precision]
    b = 8 if precision == "fp64" else 4
    bytes_lane = n * 48 + b * (5 * visits + 15 * cells + 6 * leaves)
    traffic, source = walk_traffic(precision)
    return {"bound": "valu", "achieved": round(tflops, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(tflops / peak, 4),
            "frac_unpacked": round(tflops / UNPACKED_PEAK_TFS, 4),
            "traffic": traffic, "traffic_source": source,
            "kernel": f"k_walk<{precision}>", "flops_per_launch": int(flops),
            "walk_ms": round(res["walk_ms"], 3),
            "lane_utilisation": round(visits / (64.0 * max(1, t["wave_steps"])), 4),
            "counts": {"visits": visits, "leaves": leaves, "cells": cells,
                       "wave_steps": t["wave_steps"]},
            "per_lane_bytes": int(bytes_lane)}


PAIR_FLOPS = 52  # one FP32 segment term, counting FMA as 2: see run_ensemble


DIRECT_PAIR_FLOPS = 45  # one desingularised segment term (a6 form), +,-,*,/,sqrt as one


def run_direct(n_points: int = 262144, steps: int = 3) -> dict:
    """The tiled direct sum (velocity='BS', biot_savart, timestep.f90:462-482)
    of one system of random loops in an open box, FP64 and FP32C, against
    the vector FLOP peaks: 45 algorithmic flops per source-target pair."""
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=n_points, seed=7, spacing=0.8, wrap=False)
    n = int(np.count_nonzero(fil.infront))
    pairs = n * (n - 2)  # every active source but the target and its behind
    out = {"workload": f"random loops, open box L=0.1 cm, delta=5e-4 cm, velocity BS, {n} points",
           "points": n, "pairs": pairs, "pair_flops": DIRECT_PAIR_FLOPS}
    for prec in ("fp64", "fp32c"):
        p = gpu.params(velocity="BS", tree_theta=0.0, boundary="open", box_size=0.1, delta=5e-4,
                       corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6, precision=prec)
        with gpu.Context(p) as c:
            c.upload_filaments(fil)
            c.velocity_only()
            ms = []
            for _ in range(steps):
                c.velocity_only()
                ms.append(c.timings()["walk_ms"])
        t = float(np.mean(ms))
        tf = DIRECT_PAIR_FLOPS * pairs / (t * 1e-3) / 1e12
        out[prec] = {"kernel_ms": round(t, 3), "achieved_tflops": round(tf, 2),
                     "peak_tflops": FLOP_PEAK_TFS[prec], "frac": round(tf / FLOP_PEAK_TFS[prec], 4)}
    return out


def run_ensemble(steps: int, seed: int = 20261202):
    """Ensemble of 256 rings advanced together (qv_batch_*).  dt is half the
    stability bound of timestep_check (initial.f90:606-622) for the finest
    spacing in the ensemble.  One pair term of the FP32 path costs 52 flops:
    relative vector 3, R1 x e 9, |R1 x e|^2 5, R2 = R1 + e 3, |R1|^2 and
    |R2|^2 10, two rsqrt 2, |R1|, |R2|, |R1||R2| 3, R1.R2 5, g 3, f 3, FP64
    accumulation 6."""
    import math
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    systems = G.ring_ensemble(256, seed=seed)
    counts = np.array([s.pcount for s in systems])
    spacing = min(float(np.min(np.linalg.norm(s.x[s.infront - 1] - s.x, axis=1)))
                  for s in systems)
    dmin = spacing / 2.0
    dt = 0.5 * abs(dmin ** 2 / (9.97e-4 * math.log(dmin / (math.pi * 1.3e-8))))
    kw = dict(velocity="BS", boundary="open", box_size=0.2, corea=1.3e-8, quant_circ=9.97e-4,
              dt=dt)
    with gpu.Batch(gpu.params(precision="fp64", **kw), systems) as b:
        u64 = np.concatenate(b.velocity())
    with gpu.Batch(gpu.params(precision="fp32c", **kw), systems) as b:
        chosen, tune_ms = b.tune(3)
        u32 = np.concatenate(b.velocity())
        b.step(10, graph=True)  # warm-up: capture and first replays
        ms = b.step(steps, graph=True)
        info = b.info()
    nr = np.linalg.norm(u64, axis=1)
    e = np.linalg.norm(u32 - u64, axis=1) / nr
    n = int(counts.sum())
    pairs = int(np.sum(counts * (counts - 2)))
    tflops = PAIR_FLOPS * pairs / (ms / steps * 1e-3) / 1e12
    return {"workload": "256 rings, open, points U{200..5000}, R U[0.02,0.06] cm, "
                        "helix 1% of R, fp32 pairs + fp64 LIA, AB steps",
            "systems": 256, "points": n, "steps": steps, "dt": dt,
            "ms_per_step": round(ms / steps, 4),
            "point_velocities_per_s": n * steps / (ms * 1e-3),
            "pairs_per_step": pairs, "pair_flops": PAIR_FLOPS,
            "achieved_tflops": round(tflops, 2), "peak_tflops_fp32": FLOP_PEAK_TFS["fp32c"],
            "frac": round(tflops / FLOP_PEAK_TFS["fp32c"], 4),
            "graph": True, "variant": chosen, "targets_per_unit": info["tile"],
            "work_units": info["nunits"], "tune_ms": [round(t, 4) for t in tune_ms],
            "error_vs_fp64": {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e ** 2)))}}


def run_config3(fil, theta: float, steps: int) -> dict:
    """The 1M-point tangle run as the reference's main loop runs it
    (run.f90:36-169): `steps` steps of dt = 1e-6 s, FP32 far field, the host
    topology step with the neighbour search every step, the boundary step;
    wall clock from Python over the whole run.  Beside it, the same number of
    steps without the line operations (qv_step: velocity, AB, wrap on the
    device)."""
    from qvort_amd import gpu, line
    p = gpu.params(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.2,
                   delta=5e-4, corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6, precision="fp32c")
    st = line.LineState.from_arrays(capacity=2 * fil.pcount + 64, x=fil.x,
                                    infront=fil.infront, behind=fil.behind)
    cnt = line.new_counts()
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        t0 = time.perf_counter()
        for _ in range(steps):
            c.velocity_only()
            c.ab_step()
            line.topology_step(c, p, st, cnt, search=True)
            c.enforce_periodic()
        wall = time.perf_counter() - t0
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        c.step(1)
        t0 = time.perf_counter()
        c.step(steps)
        wall_dev = time.perf_counter() - t0
    return {"steps": steps, "dt": 1e-6, "wall_s": round(wall, 3),
            "ms_per_step": round(wall / steps * 1e3, 3),
            "counts": line.counts_dict(cnt),
            "points_after": int(np.count_nonzero(st.infront)),
            "without_line_ops": {"wall_s": round(wall_dev, 3),
                                 "ms_per_step": round(wall_dev / steps * 1e3, 3)}}


def run_adjoint(n_points: int = 8192, nsteps: int = 200) -> dict:
    """The adjoint configuration: a periodic Kelvin-wave line of 8192 points
    (L = 0.1 cm); one VJP of the FP64 direct-sum velocity (qv_vjp_device)
    timed on the device stream, and the gradient of the line length after
    200 AB steps checkpointed every 20 (forward, recomputation and backward
    through the hand-written VJP), wall clock."""
    import torch

    from qvort_amd import adjoint as AD
    from qvort_amd import geometry as G
    fil = G.kelvin_line(n_points, 0.1, amp_sigma=2e-4, seed=20261202)
    eng = AD.Engine(fil, box_size=0.1, boundary="periodic")
    try:
        x = torch.as_tensor(fil.x, device="cuda")
        w = torch.ones_like(x)
        out = torch.empty_like(x)
        eng._bind()
        eng.ctx.vjp_device(x.data_ptr(), w.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            eng.ctx.vjp_device(x.data_ptr(), w.data_ptr(), out.data_ptr())
        e1.record()
        torch.cuda.synchronize()
        vjp_ms = e0.elapsed_time(e1) / 5
        dt = 0.25 * AD.dt_max(0.1 / n_points)
        t0 = time.perf_counter()
        J, g = AD.gradient("length", fil.x, eng, nsteps, dt, checkpoint_every=20)
        grad_s = time.perf_counter() - t0
    finally:
        eng.close()
    n = n_points
    pairs = 27 * n * (n - 2)
    return {"workload": f"periodic Kelvin-wave line, {n} points, L=0.1 cm, 27 shells",
            "vjp_ms": round(vjp_ms, 3), "vjp_pairs": pairs,
            "gradient_steps": nsteps, "checkpoint_every": 20,
            "gradient_wall_s": round(grad_s, 3)}


def copy_calibration(gib: float = 2.0, reps: int = 5) -> None:
    """A plain device copy of known size (torch), for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": reps}))


def run_next_rows(fil, theta: float) -> dict:
    """The calls next to the velocity stage on the same workload, wall-clock
    per call from Python (host-device copies included): the reconnection
    step's neighbour search on the stage's tree (pclose_tree), the velocity
    at 1,000 and 100,000 random probe points by tree_walk_general's rules
    (mesh_velocity), one topology step (pinsert on the host, the search,
    precon and premove on the host, state down and back up), and one whole
    step with it (velocity stage, AB update, topology step, wrap)."""
    from qvort_amd import gpu, line
    p = gpu.params(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.2,
                   delta=5e-4, corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6, precision="fp32c")
    y = np.random.Generator(np.random.PCG64(7)).uniform(-0.1, 0.1, size=(100_000, 3))
    out = {}
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        c.velocity_only()
        c.closest()
        t0 = time.perf_counter()
        for _ in range(3):
            c.closest()
        out["closest_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 3)
        c.probe_velocity(y[:1000])
        t0 = time.perf_counter()
        c.probe_velocity(y[:1000])
        out["probe_tree_1k_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        t0 = time.perf_counter()
        c.probe_velocity(y)
        out["probe_tree_100k_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        st = line.LineState.from_arrays(capacity=2 * fil.pcount + 64, x=fil.x,
                                        infront=fil.infront, behind=fil.behind)
        cnt = line.new_counts()
        c.ab_step()
        t0 = time.perf_counter()
        line.topology_step(c, p, st, cnt, search=True)
        out["topology_step_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        out["topology_counts"] = line.counts_dict(cnt)
        # one whole step as run.f90 orders it: velocity stage, AB update,
        # line operations with the search, wrap
        c.enforce_periodic()
        t0 = time.perf_counter()
        c.velocity_only()
        c.ab_step()
        line.topology_step(c, p, st, cnt, search=True)
        c.enforce_periodic()
        out["step_with_topology_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--theta", type=float, default=0.4)
    ap.add_argument("--cpu-sample", type=int, default=600_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-next-rows", action="store_true")
    ap.add_argument("--config3-steps", type=int, default=100)
    ap.add_argument("--no-config3", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32c", "direct", "ensemble", "adjoint"],
                    default=None, help="run one part for a profiler (no JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain copy of known size (for rocprofv3 counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the velocity "
              "stage would shard over targets with a replicated tree, see DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    if args.only == "direct":
        print(json.dumps(run_direct(steps=args.steps)), file=sys.stderr, flush=True)
        return
    if args.only == "ensemble":
        print(json.dumps(run_ensemble(args.ensemble_steps)), file=sys.stderr, flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps(run_adjoint()), file=sys.stderr, flush=True)
        return
    fil = workload(args.points)
    n = fil.n_active
    print(f"workload: {n} points, {len(fil.loop_len)} loops", file=sys.stderr, flush=True)
    if args.only:
        r = run_precision(fil, args.only, args.theta, args.steps, args.warmup)
        print(f"{args.only}: {r['wall_s'] / args.steps * 1e3:.2f} ms/step, walk "
              f"{r['walk_ms']:.3f} ms", file=sys.stderr, flush=True)
        return
    r64 = run_precision(fil, "fp64", args.theta, args.steps, args.warmup)
    print(f"fp64: {r64['wall_s'] / args.steps * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
    r32 = run_precision(fil, "fp32c", args.theta, args.steps, args.warmup)
    print(f"fp32c: {r32['wall_s'] / args.steps * 1e3:.2f} ms/step", file=sys.stderr, flush=True)

    act = fil.infront != 0
    nr = np.linalg.norm(r64["u"][act], axis=1)
    e = np.linalg.norm(r32["u"][act] - r64["u"][act], axis=1) / nr
    err = {"vs": "fp64 GPU, same tree", "max": float(e.max()),
           "rms": float(np.sqrt(np.mean(e ** 2))),
           "max_over_rms_velocity": float(np.max(np.linalg.norm(
               r32["u"][act] - r64["u"][act], axis=1)) / np.sqrt(np.mean(nr ** 2)))}

    ms32 = r32["wall_s"] / args.steps * 1e3
    ms64 = r64["wall_s"] / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": n / (ms32 * 1e-3),
        "unit": "point-velocities/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "random loops, periodic cube L=0.2 cm, delta=5e-4 cm",
                   "n_points": n, "tree_theta": args.theta, "shells": 27,
                   "precision": "fp32 far field, compensated sums; fp64 LIA and decisions",
                   "parallelism": "1 GPU"},
        "fp64": {"value": n / (ms64 * 1e-3), "ms_per_step": round(ms64, 3),
                 "device_ms": round(r64["device_ms"], 3), "build_ms": round(r64["build_ms"], 3),
                 "walk_ms": round(r64["walk_ms"], 3), "ghost_lia_ms": round(r64["ghost_lia_ms"], 3),
                 "ab_ms": round(r64["ab_ms"], 3)},
        "fp32c": {"device_ms": round(r32["device_ms"], 3), "build_ms": round(r32["build_ms"], 3),
                  "walk_ms": round(r32["walk_ms"], 3),
                  "ghost_lia_ms": round(r32["ghost_lia_ms"], 3), "ab_ms": round(r32["ab_ms"], 3)},
        "fp32c_speedup_over_fp64": round(ms64 / ms32, 3),
        "error": err,
        "evals_per_point": r64["timings"]["evals"] / n,
        "tree_nodes": r64["timings"]["nodes"],
        "roofline": roofline(r32, "fp32c", n),
        "roofline_fp64": roofline(r64, "fp64", n),
    }
    if not args.no_direct:
        out["direct_sum"] = run_direct()
        out["direct_sum_65k"] = run_direct(65536)
        print(f"direct sum: {out['direct_sum']} {out['direct_sum_65k']}", file=sys.stderr,
              flush=True)
    if not args.no_config3:
        out["config3_run"] = run_config3(fil, args.theta, args.config3_steps)
        print(f"config 3 run: {out['config3_run']}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint()
        print(f"adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    if not args.no_ensemble:
        out["ensemble"] = run_ensemble(args.ensemble_steps)
        print(f"ensemble: {out['ensemble']['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    if not args.no_next_rows:
        out["next_rows"] = run_next_rows(fil, args.theta)
        print(f"next rows: {out['next_rows']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fil, args.theta, args.cpu_sample)
        out["error_vs_direct"] = error_vs_direct(fil, args.theta, r32["u"], r64["u"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
