"""Run the flang build of the reference hot path (oracle/_ref/qvort_ref_driver).

Test infrastructure only (see oracle/__init__.py).  The driver is built by
``make -C oracle ref`` from the reference sources; on a machine without them
the prebuilt driver is used if present.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "qvort_ref_driver")
NODE_BYTES = 3 * 4 + 4 * 8 + 3 * 8 + 3 * 8 + 9 * 8


def available() -> bool:
    return os.path.exists(DRIVER)


def run_in_text(*, velocity="Tree", tree_theta=0.4, boundary="periodic", box_size=0.1,
                delta=5e-4, dt=1e-6, quant_circ=9.97e-4, corea=8.244023e-9,
                tree_extra_correction=True, deriv_order="second", nsteps=1,
                runge_kutta=False, shots=1000000, tree_print=False,
                recon_shots=1000000, mesh_size=0, mesh_shots=100,
                mesh_print_delay=0, remove_mean_pos=False) -> str:
    lines = [f"runge_kutta {'T' if runge_kutta else 'F'}",
        f"nsteps {nsteps}", f"shots {shots}", f"recon_shots {recon_shots}", f"dt {dt!r}",
        f"tree_print {'T' if tree_print else 'F'}",
        f"delta {delta!r}", f"box_size {box_size!r}", f"velocity {velocity}",
        f"boundary {boundary}", f"quant_circ {quant_circ!r}", f"corea {corea!r}",
        f"tree_theta {tree_theta!r}",
        f"tree_extra_correction {'T' if tree_extra_correction else 'F'}",
        f"deriv_order {deriv_order}", "initf single_loop", "pcount 16",
        f"mesh_size {mesh_size}", f"mesh_shots {mesh_shots}",
        f"mesh_print_delay {mesh_print_delay}",
        f"remove_mean_pos {'T' if remove_mean_pos else 'F'}",
    ]
    return "\n".join(lines) + "\n"


def _write_state(path, fil, nsteps, flags, nsample, u1=None, u2=None, u3=None):
    n = fil.pcount
    z = np.zeros((n, 3))
    with open(path, "wb") as fh:
        np.array([1, n, nsteps, flags, nsample], dtype="<i4").tofile(fh)
        np.ascontiguousarray(fil.x, dtype="<f8").tofile(fh)
        np.ascontiguousarray(fil.infront, dtype="<i4").tofile(fh)
        np.ascontiguousarray(fil.behind, dtype="<i4").tofile(fh)
        for a in (u1, u2, u3):
            np.ascontiguousarray(z if a is None else a, dtype="<f8").tofile(fh)


def run(fil, *, nsteps=1, dump_tree=False, timing=False, nsample=0, threads=None,
        u1=None, u2=None, u3=None, timeout=3600, dump_var=False, restore_var=None,
        topology=False, **params) -> dict:
    """Run the reference driver; returns a dict of numpy arrays.

    dump_var: also return the bytes of data/var.dat written after the run with
    data_dump's statements ("var_dat").  restore_var: bytes of a var.dat to
    restart from with data_restore's statements (fil is then ignored; the run
    continues at the restored itime + 1).  With tree_print=True and shots
    dividing a step, the reference's data/tree_mesh%03d.log files come back
    as {"tree_mesh": {number: text}}; with mesh_size > 0 the data/mesh%03d.dat
    files of the mesh velocity as {"mesh": {number: bytes}}."""
    if not available():
        raise FileNotFoundError(DRIVER)
    flags = ((1 if dump_tree else 0) | (2 if timing else 0) | (64 if dump_var else 0)
             | (128 if restore_var is not None else 0) | (512 if topology else 0))
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(prefix="qvref_") as wd:
        with open(os.path.join(wd, "run.in"), "w") as fh:
            fh.write(run_in_text(nsteps=nsteps, **params))
        os.makedirs(os.path.join(wd, "data"), exist_ok=True)
        if restore_var is not None:
            with open(os.path.join(wd, "data", "var.dat"), "wb") as fh:
                fh.write(restore_var)
            from types import SimpleNamespace
            fil = SimpleNamespace(pcount=0, x=np.zeros((0, 3)), infront=np.zeros(0, np.int32),
                                  behind=np.zeros(0, np.int32))
        _write_state(os.path.join(wd, "state.bin"), fil, nsteps, flags, nsample, u1, u2, u3)
        proc = subprocess.run([DRIVER], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0 or "FATAL" in proc.stdout:
            raise RuntimeError(f"reference driver failed: {proc.stdout[-2000:]}{proc.stderr[-2000:]}")
        out = {"stdout": proc.stdout}
        meshes = {}
        for name in sorted(os.listdir(os.path.join(wd, "data"))):
            if name.startswith("tree_mesh") and name.endswith(".log"):
                with open(os.path.join(wd, "data", name)) as fh:
                    meshes[int(name[9:-4])] = fh.read()
        if meshes:
            out["tree_mesh"] = meshes
        grids = {}
        for name in sorted(os.listdir(os.path.join(wd, "data"))):
            if name.startswith("mesh") and name.endswith(".dat"):
                with open(os.path.join(wd, "data", name), "rb") as fh:
                    grids[int(name[4:-4])] = fh.read()
        if grids:
            out["mesh"] = grids
        if dump_var:
            with open(os.path.join(wd, "data", "var.dat"), "rb") as fh:
                out["var_dat"] = fh.read()
        if timing:
            with open(os.path.join(wd, "timing.txt")) as fh:
                for line in fh:
                    k, v = line.split()
                    out[k] = float(v)
            return out
        raw = open(os.path.join(wd, "out.bin"), "rb").read()
        n = int(np.frombuffer(raw, dtype="<i4", count=1)[0])
        assert restore_var is not None or n == fil.pcount
        out["evals"] = int(np.frombuffer(raw, dtype="<i8", count=1, offset=4)[0])
        arr = np.frombuffer(raw, dtype="<f8", count=27 * n, offset=12).reshape(9, n, 3)
        for k, name in enumerate(["u", "u_li", "u_bs", "ghosti", "ghostb", "x", "u1", "u2", "u3"]):
            out[name] = arr[k].copy()
        if topology:
            off = 12 + 8 * 27 * n
            links = np.frombuffer(raw, dtype="<i4", count=2 * n + 1, offset=off)
            out["infront_final"] = links[:n].copy()
            out["behind_final"] = links[n:2 * n].copy()
            out["boundary_count"] = int(links[2 * n])
            out["boundary_length"] = float(np.frombuffer(raw, dtype="<f8", count=1,
                                                         offset=off + 4 * (2 * n + 1))[0])
        if dump_tree:
            out["tree"] = parse_tree(open(os.path.join(wd, "tree.bin"), "rb").read())
        return out


def parse_tree(raw: bytes) -> np.ndarray:
    """Node records as rows of 22 doubles, the layout of the C oracle's dump:
    level, pcount, leafj, pos(3), width, cent(3), circ(3), Cij(9) column-major."""
    dt = np.dtype([("i", "<i4", 3), ("pos", "<f8", 4), ("cent", "<f8", 3), ("circ", "<f8", 3),
                   ("C", "<f8", 9)])
    rec = np.frombuffer(raw, dtype=dt)
    out = np.zeros((rec.shape[0], 22))
    out[:, 0:3] = rec["i"]
    out[:, 3:7] = rec["pos"]
    out[:, 7:10] = rec["cent"]
    out[:, 10:13] = rec["circ"]
    out[:, 13:22] = rec["C"]
    return out


def closest_run(fil, *, pinsert=False, threads=1, timeout=600, **params) -> dict:
    """One step of ghostp, construct_tree, pmotion, [pinsert], then pclose_tree
    (tree_theta > 0) or pclose, in the order of run.f90:41-66.  Returns the
    state the search saw (pcount, x, infront, behind after pinsert), the
    closest/closestd arrays before the search and after it."""
    if not available():
        raise FileNotFoundError(DRIVER)
    flags = 4 | (8 if pinsert else 0)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(prefix="qvref_") as wd:
        with open(os.path.join(wd, "run.in"), "w") as fh:
            fh.write(run_in_text(nsteps=1, **params))
        os.makedirs(os.path.join(wd, "data"), exist_ok=True)
        _write_state(os.path.join(wd, "state.bin"), fil, 1, flags, 0)
        proc = subprocess.run([DRIVER], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0 or "FATAL" in proc.stdout:
            raise RuntimeError(f"reference driver failed: {proc.stdout[-2000:]}{proc.stderr[-2000:]}")
        raw = open(os.path.join(wd, "closest.bin"), "rb").read()
    n = int(np.frombuffer(raw, dtype="<i4", count=1)[0])
    off = 4
    out = {"pcount": n}

    def take(dtype, count, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a if shape is None else a.reshape(shape)

    out["x"] = take("<f8", 3 * n, (n, 3))
    out["infront"] = take("<i4", n)
    out["behind"] = take("<i4", n)
    out["closest_in"] = take("<i4", n)
    out["closestd_in"] = take("<f8", n)
    out["closest"] = take("<i4", n)
    out["closestd"] = take("<f8", n)
    assert off == len(raw)
    return out


def probe_run(fil, probes, *, motion=False, threads=1, timeout=600, **params) -> dict:
    """Velocity at probe points the way mesh_velocity sums it
    (timestep.f90:374-452): ghostp, construct_tree, [pmotion], then per probe
    biot_savart_general(_shift) or tree_walk_general over the shells.
    Returns the positions and ghosts the sums saw and u at the probes."""
    if not available():
        raise FileNotFoundError(DRIVER)
    flags = 16 | (32 if motion else 0)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    y = np.ascontiguousarray(probes, dtype="<f8")
    with tempfile.TemporaryDirectory(prefix="qvref_") as wd:
        with open(os.path.join(wd, "run.in"), "w") as fh:
            fh.write(run_in_text(nsteps=1, **params))
        os.makedirs(os.path.join(wd, "data"), exist_ok=True)
        _write_state(os.path.join(wd, "state.bin"), fil, 1, flags, 0)
        with open(os.path.join(wd, "probes.bin"), "wb") as fh:
            np.array([y.shape[0]], dtype="<i4").tofile(fh)
            y.tofile(fh)
        proc = subprocess.run([DRIVER], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0 or "FATAL" in proc.stdout:
            raise RuntimeError(f"reference driver failed: {proc.stdout[-2000:]}{proc.stderr[-2000:]}")
        raw = open(os.path.join(wd, "probe.bin"), "rb").read()
    n = int(np.frombuffer(raw, dtype="<i4", count=1)[0])
    arr = np.frombuffer(raw, dtype="<f8", offset=4)
    out = {"pcount": n, "x": arr[:3 * n].reshape(n, 3).copy(),
           "ghosti": arr[3 * n:6 * n].reshape(n, 3).copy(),
           "u": arr[6 * n:].reshape(-1, 3).copy()}
    assert out["u"].shape == y.shape
    return out


LINE_STAGES = ("motion", "pinsert", "search", "precon", "premove")


def line_run(fil, *, nsteps=1, threads=1, timeout=600, u1=None, u2=None, u3=None,
             **params) -> dict:
    """The main loop with the line operations of run.f90:60-72 (pinsert, then
    when recon_shots divides itime the neighbour search, precon and premove).
    Returns the state after each operation of the first step
    ({stage: {pcount, x, u, u1, u2, ghosti, ghostb, ghostii, ghostbb,
    infront, behind, closest, closestd}} for the stages in LINE_STAGES that
    ran), the counters after the first step ("counters1") and the state and
    counters after nsteps ("final")."""
    if not available():
        raise FileNotFoundError(DRIVER)
    flags = 256
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(prefix="qvref_") as wd:
        with open(os.path.join(wd, "run.in"), "w") as fh:
            fh.write(run_in_text(nsteps=nsteps, **params))
        os.makedirs(os.path.join(wd, "data"), exist_ok=True)
        _write_state(os.path.join(wd, "state.bin"), fil, nsteps, flags, 0, u1, u2, u3)
        proc = subprocess.run([DRIVER], cwd=wd, env=env, capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0 or "FATAL" in proc.stdout:
            raise RuntimeError(f"reference driver failed: {proc.stdout[-2000:]}{proc.stderr[-2000:]}")
        raw = open(os.path.join(wd, "line.bin"), "rb").read()
    off = 0

    def take(dtype, count, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a if shape is None else a.reshape(shape)

    def counters():
        c = take("<i4", 3)
        r = take("<f8", 3)
        return {"recon_count": int(c[0]), "remove_count": int(c[1]),
                "small_loop_remove_count": int(c[2]), "total_length": float(r[0]),
                "avg_sep": float(r[1]), "small_loop_remove_length": float(r[2])}

    recon = params.get("recon_shots", 1000000)
    stages = LINE_STAGES if recon <= 1 or 1 % recon == 0 else LINE_STAGES[:2]
    out = {}
    for st in stages:
        n = int(take("<i4", 1)[0])
        d = {"pcount": n}
        for name in ("x", "u", "u1", "u2", "ghosti", "ghostb", "ghostii", "ghostbb"):
            d[name] = take("<f8", 3 * n, (n, 3))
        for name in ("infront", "behind", "closest"):
            d[name] = take("<i4", n)
        d["closestd"] = take("<f8", n)
        out[st] = d
    out["counters1"] = counters()
    n = int(take("<i4", 1)[0])
    fin = {"pcount": n, "x": take("<f8", 3 * n, (n, 3)), "u1": take("<f8", 3 * n, (n, 3)),
           "u2": take("<f8", 3 * n, (n, 3)), "infront": take("<i4", n), "behind": take("<i4", n)}
    fin.update(counters())
    out["final"] = fin
    assert off == len(raw)
    return out
