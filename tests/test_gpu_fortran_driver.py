"""The Fortran driver (qvort_amd/lib/qv_run, iso_c_binding module
qvort_amd/fortran/qv_gpu.f90) against the golden trajectory of the
reference: run.in + state.bin in, data/var0001.log out."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "qvort_amd", "lib", "qv_run")


def _read_var(path, n):
    raw = open(path, "rb").read()
    t = np.frombuffer(raw, "<f8", 1, 0)[0]
    pc = np.frombuffer(raw, "<i4", 1, 8)[0]
    assert pc == n
    off = 12
    xyz = np.frombuffer(raw, "<f8", 3 * n, off).reshape(3, n).T
    off += 24 * n
    inf = np.frombuffer(raw, "<i4", n, off)
    off += 4 * n
    u = np.frombuffer(raw, "<f8", 3 * n, off).reshape(3, n).T
    return t, xyz, inf, u


@pytest.mark.parametrize("name", ["ring512_bs_open", "loops_tree_periodic_t04",
                                  "loops_ab4_history"])
def test_driver_trajectory(name, tmp_path):
    g = load_golden(name)
    meta = g["meta"]
    p = meta["params"]
    n = meta["pcount"]
    print("case", name, flush=True)
    lines = [f"nsteps {meta['nsteps']}", f"shots {meta['nsteps']}", f"dt {p['dt']!r}",
             f"delta {p['delta']!r}", f"box_size {p['box_size']!r}",
             f"velocity {p['velocity']}", f"boundary {p['boundary']}",
             f"quant_circ {p['quant_circ']!r}", f"corea {p['corea']!r}",
             f"tree_theta {p['tree_theta']!r}", "initf state_file", "gpu_precision fp64"]
    (tmp_path / "run.in").write_text("\n".join(lines) + "\n")
    z = np.zeros((n, 3))
    with open(tmp_path / "state.bin", "wb") as fh:
        np.array([1, n, meta["nsteps"], 0, 0], dtype="<i4").tofile(fh)
        g["x0"].astype("<f8").tofile(fh)
        g["infront"].astype("<i4").tofile(fh)
        g["behind"].astype("<i4").tofile(fh)
        for k in (1, 2, 3):
            g.get(f"u{k}_0", z).astype("<f8").tofile(fh)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    t, x, inf, u = _read_var(tmp_path / "data" / "var0001.log", n)
    act = g["infront"] != 0
    assert np.array_equal(inf, g["infront"])
    assert abs(t - meta["nsteps"] * p["dt"]) < 1e-18
    assert np.max(np.abs(x - g["x_final"])[act]) <= 1e-15
    assert os.path.exists(tmp_path / "data" / "ts.log")


@pytest.mark.parametrize("name", ["bnd_openface_bs_openx", "bnd_openface_tree_openxy",
                                  "bnd_kelvin_meanpos_bs_periodic"])
def test_driver_boundary_and_mean_removal(name, tmp_path):
    """The driver's boundary step removes the loops that leave an openx /
    openxy box through an open face, and remove_mean_pos T recentres the
    points, as the reference's main loop does: links and positions after
    the run."""
    g = load_golden(name)
    meta = g["meta"]
    p = meta["params"]
    n = meta["pcount"]
    print("case", name, flush=True)
    lines = [f"nsteps {meta['nsteps']}", f"shots {meta['nsteps']}", f"dt {p['dt']!r}",
             f"delta {p['delta']!r}", f"box_size {p['box_size']!r}",
             f"velocity {p['velocity']}", f"boundary {p['boundary']}",
             f"quant_circ {p['quant_circ']!r}", f"corea {p['corea']!r}",
             f"tree_theta {p['tree_theta']!r}", "initf state_file", "gpu_precision fp64",
             f"remove_mean_pos {'T' if p.get('remove_mean_pos') else 'F'}"]
    (tmp_path / "run.in").write_text("\n".join(lines) + "\n")
    z = np.zeros((n, 3))
    with open(tmp_path / "state.bin", "wb") as fh:
        np.array([1, n, meta["nsteps"], 0, 0], dtype="<i4").tofile(fh)
        g["x0"].astype("<f8").tofile(fh)
        g["infront"].astype("<i4").tofile(fh)
        g["behind"].astype("<i4").tofile(fh)
        for _ in range(3):
            z.astype("<f8").tofile(fh)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    _, x, inf, _ = _read_var(tmp_path / "data" / "var0001.log", n)
    assert np.array_equal(inf, g["infront_final"])
    act = g["infront_final"] != 0
    assert np.max(np.abs(x - g["x_final"])[act]) <= 1e-15
