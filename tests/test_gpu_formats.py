"""The Fortran driver's restart file, leaf print and mesh velocity files
against the reference build's own files (tests/golden/fmt_*.npz,
tests/golden/mesh_*.npz): data/var.dat after a run, the restart from the
reference's var.dat, data/tree_mesh%03d.log and data/mesh%03d.dat."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, fmt_names, load_golden, mesh_names

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "qvort_amd", "lib", "qv_run")
NAMES = fmt_names()
MESH_NAMES = mesh_names()


def _run_in(path, p, nsteps):
    lines = [f"nsteps {nsteps}", "shots 1", f"dt {p['dt']!r}", f"delta {p['delta']!r}",
             f"box_size {p['box_size']!r}", f"velocity {p['velocity']}",
             f"boundary {p['boundary']}", f"quant_circ {p['quant_circ']!r}",
             f"corea {p['corea']!r}", f"tree_theta {p['tree_theta']!r}", "initf state_file",
             "tree_print T", "gpu_precision fp64"]
    (path / "run.in").write_text("\n".join(lines) + "\n")


def _drive(path):
    r = subprocess.run([DRIVER], cwd=path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0


def _t_after(nsteps, dt):
    t = 0.0
    for _ in range(nsteps):
        t = t + dt
    return t


@pytest.mark.parametrize("name", NAMES)
def test_driver_writes_var_dat_and_tree_mesh(name, tmp_path):
    from qvort_amd import formats
    g = load_golden(name)
    meta = g["meta"]
    n, ns = meta["pcount"], meta["nsteps"]
    print("case", name, flush=True)
    _run_in(tmp_path, meta["params"], ns)
    z = np.zeros((n, 3))
    with open(tmp_path / "state.bin", "wb") as fh:
        np.array([1, n, ns, 0, 0], dtype="<i4").tofile(fh)
        g["x0"].astype("<f8").tofile(fh)
        g["infront"].astype("<i4").tofile(fh)
        g["behind"].astype("<i4").tofile(fh)
        for _ in range(3):
            z.astype("<f8").tofile(fh)
    _drive(tmp_path)
    # the first step's tree is built from the input: the same text
    mesh1 = (tmp_path / "data" / "tree_mesh001.log").read_text()
    assert mesh1 == g["mesh_1"].tobytes().decode()
    for k in range(2, ns + 1):
        rows = formats.read_tree_mesh(tmp_path / "data" / f"tree_mesh{k:03d}.log")
        assert np.array_equal(rows, formats.read_tree_mesh(g[f"mesh_{k}"].tobytes().decode()))
    raw = (tmp_path / "data" / "var.dat").read_bytes()
    assert len(raw) == g["var_dat"].size
    d = formats.read_var_dat(tmp_path / "data" / "var.dat")
    assert d["pcount"] == n and d["itime"] == ns
    assert d["t"] == _t_after(ns, meta["params"]["dt"])
    f = d["f"]
    act = g["infront"] != 0
    assert np.array_equal(f["infront"], g["infront"]) and np.array_equal(f["behind"], g["behind"])
    assert np.max(np.abs(f["x"] - g["x_final"])[act]) <= 1e-15
    for k in (1, 2):
        ref = g[f"u{k}_final"][act]
        assert np.max(np.linalg.norm(f[f"u{k}"][act] - ref, axis=1)
                      / np.linalg.norm(ref, axis=1)) <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_driver_restarts_from_reference_var_dat(name, tmp_path):
    from qvort_amd import formats
    g = load_golden(name)
    meta = g["meta"]
    ns = meta["nsteps"]
    print("case", name, flush=True)
    _run_in(tmp_path, meta["params"], ns + 1)
    (tmp_path / "data").mkdir()
    (tmp_path / "data" / "var.dat").write_bytes(g["var_dat"].tobytes())
    _drive(tmp_path)
    # the restarted step continues the numbering and builds its tree from
    # the restored positions: the reference's text
    assert not (tmp_path / "data" / "tree_mesh001.log").exists()
    mesh = (tmp_path / "data" / f"tree_mesh{ns + 1:03d}.log").read_text()
    assert mesh == g[f"mesh_{ns + 1}"].tobytes().decode()
    d = formats.read_var_dat(tmp_path / "data" / "var.dat")
    assert d["itime"] == ns + 1
    act = g["infront"] != 0
    assert np.max(np.abs(d["f"]["x"] - g["x_restart"])[act]) <= 1e-15
    ref = formats.read_var_dat_bytes(g["var_dat_restart"].tobytes())
    assert d["t"] == ref["t"]


def _write_state(path, g, ns):
    n = g["x0"].shape[0]
    z = np.zeros((n, 3))
    with open(path / "state.bin", "wb") as fh:
        np.array([1, n, ns, 0, 0], dtype="<i4").tofile(fh)
        g["x0"].astype("<f8").tofile(fh)
        g["infront"].astype("<i4").tofile(fh)
        g["behind"].astype("<i4").tofile(fh)
        for _ in range(3):
            z.astype("<f8").tofile(fh)


@pytest.mark.parametrize("name", MESH_NAMES)
def test_driver_writes_mesh_velocity(name, tmp_path):
    """mesh_size > 0: the same files (numbers, t, grid planes, u_norm = 0)
    and u_sup within 1e-12 of the largest reference value; mesh sums in
    Tree mode with the reference's factors (mesh_reference_scale T)."""
    from qvort_amd import formats
    g = load_golden(name)
    meta = g["meta"]
    p, ns = meta["params"], meta["nsteps"]
    print("case", name, flush=True)
    lines = [f"nsteps {ns}", f"shots {p['shots']}", f"dt {p['dt']!r}", f"delta {p['delta']!r}",
             f"box_size {p['box_size']!r}", f"velocity {p['velocity']}",
             f"boundary {p['boundary']}", f"quant_circ {p['quant_circ']!r}",
             f"corea {p['corea']!r}", f"tree_theta {p['tree_theta']!r}", "initf state_file",
             f"runge_kutta {'T' if p.get('runge_kutta') else 'F'}", "gpu_precision fp64",
             f"mesh_size {p['mesh_size']}", f"mesh_shots {p['mesh_shots']}",
             f"mesh_print_delay {p.get('mesh_print_delay', 0)}", "mesh_reference_scale T"]
    (tmp_path / "run.in").write_text("\n".join(lines) + "\n")
    _write_state(tmp_path, g, ns)
    _drive(tmp_path)
    made = sorted(int(f.name[4:-4]) for f in (tmp_path / "data").iterdir()
                  if f.name.startswith("mesh") and f.name.endswith(".dat"))
    assert made == meta["meshes"]
    for k in made:
        d = formats.read_mesh_dat(tmp_path / "data" / f"mesh{k:03d}.dat")
        ref_path = tmp_path / f"ref{k:03d}.dat"
        ref_path.write_bytes(g[f"mesh_{k}"].tobytes())
        ref = formats.read_mesh_dat(ref_path)
        assert d["mesh_size"] == ref["mesh_size"] and d["t"] == ref["t"]
        assert np.array_equal(d["x"], ref["x"]) and not np.any(d["u_norm"])
        scale = np.max(np.linalg.norm(ref["u_sup"], axis=1))
        err = np.max(np.linalg.norm(d["u_sup"] - ref["u_sup"], axis=1)) / scale
        print(name, k, "max err", err, flush=True)
        assert err <= 1e-12

