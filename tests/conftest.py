import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqvort_gpu.so)")


def golden_names():
    with open(os.path.join(GOLDEN, "index.json")) as fh:
        return sorted(json.load(fh))


def closest_names():
    with open(os.path.join(GOLDEN, "closest_index.json")) as fh:
        return sorted(json.load(fh))


def line_names():
    with open(os.path.join(GOLDEN, "line_index.json")) as fh:
        return sorted(json.load(fh))


def fmt_names():
    with open(os.path.join(GOLDEN, "fmt_index.json")) as fh:
        return sorted(json.load(fh))


def mesh_names():
    with open(os.path.join(GOLDEN, "mesh_index.json")) as fh:
        return sorted(json.load(fh))


def probe_names():
    with open(os.path.join(GOLDEN, "probe_index.json")) as fh:
        return sorted(json.load(fh))


def rk3_names():
    with open(os.path.join(GOLDEN, "rk3_index.json")) as fh:
        return sorted(json.load(fh))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(bytes(d["meta"]).decode())
    return d


class Fil:
    """Minimal filament record (x, infront, behind) for the oracle binding."""

    def __init__(self, x, infront, behind):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.infront = np.ascontiguousarray(infront, dtype=np.int32)
        self.behind = np.ascontiguousarray(behind, dtype=np.int32)

    @property
    def pcount(self):
        return self.x.shape[0]


# run.in labels of the time loop and its output, not of the velocity
_LOOP_KEYS = ("runge_kutta", "shots", "mesh_size", "mesh_shots", "mesh_print_delay")


def oracle_params(p):
    from oracle import port
    kw = {k: v for k, v in p.items() if k not in _LOOP_KEYS}
    kw["deriv_order"] = 4 if kw.get("deriv_order") == "fourth" else 2
    return port.make_params(**kw)


def gpu_params(p, precision="fp64"):
    from qvort_amd import gpu
    kw = {k: v for k, v in p.items() if k not in _LOOP_KEYS}
    kw["deriv_order"] = 4 if kw.get("deriv_order") == "fourth" else 2
    return gpu.params(precision=precision, **kw)


def rel_err(u, ref, mask=None):
    """Per-point |u - ref| / |ref| over the points of mask (points with
    |ref| == 0 use the RMS norm)."""
    if mask is not None:
        u, ref = u[mask], ref[mask]
    nr = np.linalg.norm(ref, axis=1)
    scale = np.where(nr > 0, nr, np.sqrt(np.mean(nr ** 2)) + 1e-300)
    return np.linalg.norm(u - ref, axis=1) / scale


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import port
    return port.lib()


def bnd_names():
    with open(os.path.join(GOLDEN, "bnd_index.json")) as fh:
        return sorted(json.load(fh))
