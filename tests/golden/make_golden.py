"""Write the golden vectors of tests/golden/*.npz from the reference build.

Runs the flang build of the reference hot path (oracle/_ref/qvort_ref_driver,
built by ``make -C oracle ref`` from /root/reference/src) on small seeded
inputs and stores inputs and outputs as compressed numpy archives.  Each
case records the run.in parameters it used.  Run from the repository root:

    python tests/golden/make_golden.py              # velocity / tree / AB cases
    python tests/golden/make_golden.py --closest    # neighbour-search cases
    python tests/golden/make_golden.py --probe      # velocity at probe points
    python tests/golden/make_golden.py --rk3        # pmotion_RK3 trajectories
    python tests/golden/make_golden.py --formats    # var.dat and tree_mesh files
    python tests/golden/make_golden.py --line       # pinsert / precon / premove
    python tests/golden/make_golden.py --mesh       # mesh velocity files
    python tests/golden/make_golden.py --boundary   # open-face loop removal, remove_mean_pos

The neighbour-search cases (closest_*.npz, index in closest_index.json) run
one step of ghostp, construct_tree, pmotion, optionally pinsert, then
pclose_tree or pclose, and store the build state (x0, infront0, behind0),
the state the search saw (x, infront, behind), closest/closestd before the
search (closest_in, closestd_in) and after it.

The probe cases (probe_*.npz, index in probe_index.json) run ghostp,
construct_tree, optionally pmotion, then sum the velocity at probe points as
mesh_velocity does, and store the input state, the probes, the positions and
ghosts the sums saw (x, ghosti) and u.

The Runge-Kutta cases (rk3_*.npz, index in rk3_index.json) run the main loop
with runge_kutta T (pmotion_RK3, then the wrap) and store the input, the
positions and the Runge-Kutta storage u1 after nsteps, and the velocity of
the first step's last stage.

The format cases (fmt_*.npz, index in fmt_index.json) run nsteps of the main
loop with tree_print T and shots 1, then write data/var.dat with data_dump's
statements, and restart from that file for one more step (data_restore's
statements).  They store the input state, the var.dat bytes after nsteps,
the positions and AB history after nsteps, the text of every
data/tree_mesh%03d.log (mesh_<n>), and from the restarted run the positions
after the extra step and its var.dat (var_dat_restart).

The mesh cases (mesh_*.npz, index in mesh_index.json) run nsteps of the main
loop with mesh_size > 0: the reference's mesh_velocity after the motion and
data/mesh%03d.dat as print_mesh writes it.  They store the input state and
the bytes of every mesh file (mesh_<n>).

The boundary cases (bnd_*.npz, index in bnd_index.json) run nsteps of the
main loop on states with loops across an open face of openx / openxy boxes
(enforce_periodic_yz / _z remove them, boundary_loop_remove) or with
remove_mean_pos T, and store the input, the first step's velocity, the
positions, AB history and links after nsteps, and boundary_loop_remove's
count and length.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402
from qvort_amd import geometry as G  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cases():
    loops = G.random_loops(0.1, 5e-4, n_loops=4, seed=11)
    loops_open = G.random_loops(0.1, 5e-4, n_loops=4, seed=12, wrap=False)
    holes = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=3, seed=13), 40, seed=5)
    kel = G.kelvin_line(256, 0.1, amp_sigma=2e-3, seed=14)
    ring = G.single_ring(512, 0.05)
    base = dict(box_size=0.1, delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4)
    rng = np.random.Generator(np.random.PCG64(99))
    hist = [rng.normal(0, 0.05, size=loops.x.shape) for _ in range(3)]
    return [
        ("ring512_bs_open", ring, dict(base, velocity="BS", tree_theta=0.0, boundary="open",
                                       box_size=0.2, delta=6.2e-4, dt=1e-5), 20, False, None),
        ("loops_tree_periodic_t04", loops, dict(base, velocity="Tree", tree_theta=0.4,
                                                boundary="periodic"), 3, True, None),
        ("loops_bs_periodic", loops, dict(base, velocity="BS", tree_theta=0.0,
                                          boundary="periodic"), 3, False, None),
        ("loops_tree_open_t06", loops_open, dict(base, velocity="Tree", tree_theta=0.6,
                                                 boundary="open"), 2, True, None),
        ("loops_tree_openx_t03", loops, dict(base, velocity="Tree", tree_theta=0.3,
                                             boundary="openx"), 2, False, None),
        ("loops_bs_openxy", loops, dict(base, velocity="BS", tree_theta=0.0,
                                        boundary="openxy"), 2, False, None),
        ("loops_tree_fourth", loops, dict(base, velocity="Tree", tree_theta=0.4,
                                          boundary="periodic", deriv_order="fourth"), 2, False,
         None),
        ("loops_tree_noextra", loops, dict(base, velocity="Tree", tree_theta=0.5,
                                           boundary="periodic", tree_extra_correction=False), 2,
         False, None),
        ("holes_tree_periodic", holes, dict(base, velocity="Tree", tree_theta=0.4,
                                            boundary="periodic"), 2, True, None),
        ("kelvin_tree_periodic", kel, dict(base, velocity="Tree", tree_theta=0.4,
                                           boundary="periodic"), 3, False, None),
        ("loops_ab4_history", loops, dict(base, velocity="BS", tree_theta=0.0,
                                          boundary="periodic"), 2, False, hist),
    ]


def closest_cases():
    loops = G.random_loops(0.1, 5e-4, n_loops=6, seed=5)
    holes = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=8, seed=9), 40, seed=3)
    loops_open = G.random_loops(0.1, 5e-4, n_loops=5, seed=11, wrap=False)
    base = dict(box_size=0.1, delta=3.5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4,
                boundary="periodic")
    return [
        ("closest_tree_pinsert", loops, dict(base, velocity="Tree", tree_theta=0.4), True),
        ("closest_tree_holes", holes, dict(base, velocity="Tree", tree_theta=0.3), True),
        ("closest_tree_open", loops_open, dict(base, velocity="Tree", tree_theta=0.6,
                                               boundary="open", box_size=0.2), False),
        ("closest_bs_pinsert", holes, dict(base, velocity="BS", tree_theta=0.0), True),
        ("closest_bs_with_tree", loops, dict(base, velocity="BS", tree_theta=0.4), False),
    ]


def main_closest():
    index = {}
    for name, fil, params, pinsert in closest_cases():
        r = refrun.closest_run(fil, pinsert=pinsert, **params)
        arrays = dict(x0=fil.x, infront0=fil.infront, behind0=fil.behind, x=r["x"],
                      infront=r["infront"], behind=r["behind"], closest_in=r["closest_in"],
                      closestd_in=r["closestd_in"], closest=r["closest"],
                      closestd=r["closestd"])
        meta = dict(params=params, pinsert=pinsert, pcount0=fil.pcount, pcount=r["pcount"])
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        act = r["infront"] != 0
        print(name, fil.pcount, "->", r["pcount"], "closest=0:",
              int(np.sum(r["closest"][act] == 0)), flush=True)
    with open(os.path.join(OUT, "closest_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def probe_cases():
    loops = G.random_loops(0.1, 5e-4, n_loops=4, seed=41)
    loops_open = G.random_loops(0.1, 5e-4, n_loops=4, seed=42, wrap=False)
    rng = np.random.Generator(np.random.PCG64(43))
    y = rng.uniform(-0.05, 0.05, size=(300, 3))
    base = dict(box_size=0.1, delta=5e-4, dt=1e-5, corea=1.3e-8, quant_circ=9.97e-4)
    return [
        ("probe_bs_periodic_motion", loops, y, dict(base, velocity="BS", tree_theta=0.0,
                                                    boundary="periodic"), True),
        ("probe_bs_openx", loops, y, dict(base, velocity="BS", tree_theta=0.0,
                                          boundary="openx"), False),
        ("probe_tree_periodic", loops, y, dict(base, velocity="Tree", tree_theta=0.4,
                                               boundary="periodic"), False),
        ("probe_tree_open_motion", loops_open, y, dict(base, velocity="Tree", tree_theta=0.6,
                                                       boundary="open"), True),
    ]


def main_probe():
    index = {}
    for name, fil, y, params, motion in probe_cases():
        r = refrun.probe_run(fil, y, motion=motion, **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind, probes=y, x=r["x"],
                      ghosti=r["ghosti"], u=r["u"])
        meta = dict(params=params, motion=motion, pcount=fil.pcount, nprobe=int(y.shape[0]))
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "|u|max", float(np.abs(r["u"]).max()), flush=True)
    with open(os.path.join(OUT, "probe_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def rk3_cases():
    ring = G.single_ring(256, 0.03)
    loops = G.random_loops(0.1, 5e-4, n_loops=4, seed=61)
    holes = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=3, seed=62), 30, seed=6)
    base = dict(delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4, runge_kutta=True)
    return [
        ("rk3_ring_bs_open", ring, dict(base, velocity="BS", tree_theta=0.0, boundary="open",
                                        box_size=0.2, delta=8e-4, dt=1e-5), 3),
        ("rk3_loops_tree_periodic", loops, dict(base, velocity="Tree", tree_theta=0.4,
                                                boundary="periodic", box_size=0.1), 3),
        ("rk3_holes_bs_openx", holes, dict(base, velocity="BS", tree_theta=0.0,
                                           boundary="openx", box_size=0.1), 2),
    ]


def main_rk3():
    index = {}
    for name, fil, params, nsteps in rk3_cases():
        r = refrun.run(fil, nsteps=nsteps, threads=1, **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind, u_first=r["u"],
                      x_final=r["x"], u1_final=r["u1"])
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount)
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, flush=True)
    with open(os.path.join(OUT, "rk3_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def line_cases():
    scene = G.reconnection_scene(seed=81)
    scene2 = G.reconnection_scene(seed=82, gap=0.2)
    tangle = G.random_loops(0.03, 5e-4, n_loops=6, seed=83)
    base = dict(delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4, box_size=0.1)
    return [
        ("line_scene_tree_periodic", scene, dict(base, velocity="Tree", tree_theta=0.4,
                                                 boundary="periodic", recon_shots=1), 3),
        ("line_scene_bs_open", scene2, dict(base, velocity="BS", tree_theta=0.0,
                                            boundary="open", recon_shots=1), 3),
        ("line_tangle_tree_periodic", tangle, dict(base, velocity="Tree", tree_theta=0.4,
                                                   boundary="periodic", box_size=0.03,
                                                   recon_shots=1), 3),
        ("line_scene_fourth_openx", scene, dict(base, velocity="Tree", tree_theta=0.5,
                                                boundary="openx", deriv_order="fourth",
                                                recon_shots=2), 4),
        # spacing 1.05 delta everywhere and no empty slot: pinsert appends a
        # slot for every segment
        ("line_sparse_ring_bs_periodic", G.single_ring(60, 60 * 1.05 * 5e-4 / (2 * np.pi)),
         dict(base, velocity="BS", tree_theta=0.0, boundary="periodic", recon_shots=1), 2),
    ]


def main_line():
    index = {}
    for name, fil, params, nsteps in line_cases():
        r = refrun.line_run(fil, nsteps=nsteps, threads=1, **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind)
        stages = [st for st in refrun.LINE_STAGES if st in r]
        for st in stages:
            for k, v in r[st].items():
                arrays[f"{st}__{k}"] = np.asarray(v)
        fin = r["final"]
        for k in ("pcount", "x", "u1", "u2", "infront", "behind"):
            arrays[f"final__{k}"] = np.asarray(fin[k])
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount, stages=stages,
                    counters1=r["counters1"],
                    counters_final={k: fin[k] for k in r["counters1"]})
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "->", fin["pcount"], r["counters1"], meta["counters_final"],
              flush=True)
    with open(os.path.join(OUT, "line_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def fmt_cases():
    holes = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=3, seed=71), 20, seed=7)
    loops = G.random_loops(0.1, 5e-4, n_loops=2, seed=72)
    base = dict(delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4, box_size=0.1)
    return [
        ("fmt_holes_tree_periodic", holes, dict(base, velocity="Tree", tree_theta=0.4,
                                                boundary="periodic"), 2),
        ("fmt_loops_bs_openx", loops, dict(base, velocity="BS", tree_theta=0.5,
                                           boundary="openx"), 2),
    ]


def main_formats():
    index = {}
    for name, fil, params, nsteps in fmt_cases():
        r = refrun.run(fil, nsteps=nsteps, threads=1, dump_var=True, tree_print=True, shots=1,
                       **params)
        rr = refrun.run(None, nsteps=1, threads=1, dump_var=True, tree_print=True, shots=1,
                        restore_var=r["var_dat"], **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind,
                      var_dat=np.frombuffer(r["var_dat"], dtype=np.uint8),
                      x_final=r["x"], u1_final=r["u1"], u2_final=r["u2"], u3_final=r["u3"],
                      x_restart=rr["x"],
                      var_dat_restart=np.frombuffer(rr["var_dat"], dtype=np.uint8))
        meshes = dict(r.get("tree_mesh", {}))
        meshes.update(rr.get("tree_mesh", {}))
        for k, text in meshes.items():
            arrays[f"mesh_{k}"] = np.frombuffer(text.encode(), dtype=np.uint8)
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount, meshes=sorted(meshes))
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "meshes", sorted(meshes), flush=True)
    with open(os.path.join(OUT, "fmt_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def mesh_cases():
    holes = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=3, seed=91), 20, seed=9)
    loops = G.random_loops(0.1, 5e-4, n_loops=3, seed=92)
    base = dict(delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4, box_size=0.1, shots=1)
    return [
        ("mesh_holes_bs_periodic", holes, dict(base, velocity="BS", tree_theta=0.0,
                                               boundary="periodic", mesh_size=6,
                                               mesh_shots=1), 2),
        # the first print waits for itime > 2; files are numbered from there
        ("mesh_loops_tree_openx", loops, dict(base, velocity="Tree", tree_theta=0.4,
                                              boundary="openx", mesh_size=8, mesh_shots=2,
                                              mesh_print_delay=2), 4),
        ("mesh_loops_rk3_tree_periodic", loops, dict(base, velocity="Tree", tree_theta=0.5,
                                                     boundary="periodic", mesh_size=4,
                                                     mesh_shots=1, runge_kutta=True), 2),
        ("mesh_holes_bs_openxy", holes, dict(base, velocity="BS", tree_theta=0.0,
                                             boundary="openxy", mesh_size=4, mesh_shots=1,
                                             mesh_print_delay=1), 2),
    ]


def main_mesh():
    index = {}
    for name, fil, params, nsteps in mesh_cases():
        r = refrun.run(fil, nsteps=nsteps, threads=1, **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind, x_final=r["x"])
        grids = r.get("mesh", {})
        for k, raw in grids.items():
            arrays[f"mesh_{k}"] = np.frombuffer(raw, dtype=np.uint8)
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount, meshes=sorted(grids))
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "meshes", sorted(grids), flush=True)
    with open(os.path.join(OUT, "mesh_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def _ring(n, radius, centre, plane="xy"):
    """A ring of n points in the given coordinate plane."""
    r = G.single_ring(n, radius)
    if plane == "xz":
        r.x = r.x[:, [0, 2, 1]].copy()
    r.x = r.x + np.asarray(centre, dtype=np.float64)
    return r


def boundary_cases():
    L = 0.1
    inner = G.random_loops(L, 5e-4, n_loops=3, seed=61)
    # openx: one ring straddles x = +L/2 and one x = -L/2 (both removed after
    # the first motion), one straddles y = +L/2 (wrapped); holes in f(:)
    openx = G.with_holes(G.combine(inner, _ring(60, 0.004, (L / 2 - 0.0015, 0.01, 0.0)),
                                   _ring(48, 0.003, (-L / 2 + 0.001, -0.02, 0.01), "xz"),
                                   _ring(52, 0.0035, (0.02, L / 2 - 0.001, 0.0))), 12, seed=62)
    # openxy: a ring straddles y = -L/2 (removed), one z = +L/2 (wrapped), one
    # x = +L/2 (removed)
    openxy = G.combine(_ring(56, 0.004, (0.0, -L / 2 + 0.0012, 0.0)),
                       _ring(50, 0.0035, (0.01, 0.02, L / 2 - 0.001), "xz"),
                       _ring(44, 0.003, (L / 2 - 0.0005, 0.0, -0.02)),
                       G.random_loops(L, 5e-4, n_loops=2, seed=63))
    kel = G.kelvin_line(256, L, amp_sigma=2e-3, seed=64)
    base = dict(box_size=L, delta=5e-4, dt=1e-6, corea=1.3e-8, quant_circ=9.97e-4)
    return [
        ("bnd_openface_bs_openx", openx, dict(base, velocity="BS", tree_theta=0.0,
                                              boundary="openx"), 3),
        ("bnd_openface_tree_openxy", openxy, dict(base, velocity="Tree", tree_theta=0.4,
                                                  boundary="openxy"), 3),
        ("bnd_openface_tree_openx", openx, dict(base, velocity="Tree", tree_theta=0.5,
                                                boundary="openx"), 2),
        ("bnd_kelvin_meanpos_bs_periodic", kel, dict(base, velocity="BS", tree_theta=0.0,
                                                     boundary="periodic",
                                                     remove_mean_pos=True), 6),
        ("bnd_kelvin_meanpos_tree_periodic", kel, dict(base, velocity="Tree", tree_theta=0.4,
                                                       boundary="periodic",
                                                       remove_mean_pos=True), 5),
    ]


def main_boundary():
    index = {}
    for name, fil, params, nsteps in boundary_cases():
        r = refrun.run(fil, nsteps=nsteps, threads=1, topology=True, **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind, u=r["u"],
                      x_final=r["x"], u1_final=r["u1"], u2_final=r["u2"], u3_final=r["u3"],
                      infront_final=r["infront_final"], behind_final=r["behind_final"],
                      boundary=np.array([r["boundary_count"]], dtype=np.int64),
                      boundary_length=np.array([r["boundary_length"]]))
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount, n_active=fil.n_active)
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "active", fil.n_active, "->",
              int(np.count_nonzero(r["infront_final"])), "loops removed",
              r["boundary_count"], "length", r["boundary_length"], flush=True)
    with open(os.path.join(OUT, "bnd_index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


def main():
    if "--boundary" in sys.argv[1:]:
        main_boundary()
        return
    if "--mesh" in sys.argv[1:]:
        main_mesh()
        return
    if "--line" in sys.argv[1:]:
        main_line()
        return
    if "--formats" in sys.argv[1:]:
        main_formats()
        return
    if "--rk3" in sys.argv[1:]:
        main_rk3()
        return
    if "--probe" in sys.argv[1:]:
        main_probe()
        return
    if "--closest" in sys.argv[1:]:
        main_closest()
        return
    index = {}
    for name, fil, params, nsteps, dump, hist in cases():
        u1, u2, u3 = hist if hist is not None else (None, None, None)
        r = refrun.run(fil, nsteps=nsteps, dump_tree=dump, threads=1, u1=u1, u2=u2, u3=u3,
                       **params)
        arrays = dict(x0=fil.x, infront=fil.infront, behind=fil.behind, u=r["u"],
                      u_li=r["u_li"], u_bs=r["u_bs"], ghosti=r["ghosti"], ghostb=r["ghostb"],
                      x_final=r["x"], u1_final=r["u1"], u2_final=r["u2"], u3_final=r["u3"],
                      evals=np.array([r["evals"]], dtype=np.int64))
        if hist is not None:
            arrays.update(u1_0=u1, u2_0=u2, u3_0=u3)
        if dump:
            arrays["tree"] = r["tree"]
        meta = dict(params=params, nsteps=nsteps, pcount=fil.pcount, n_active=fil.n_active)
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
        index[name] = meta
        print(name, fil.pcount, "evals", r["evals"], flush=True)
    with open(os.path.join(OUT, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
