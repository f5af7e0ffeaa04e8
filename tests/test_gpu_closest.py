"""Neighbour search of the reconnection step on the GPU (qv_closest).

Against the reference build's fixtures: the tree of qv_velocity at the
pre-motion positions, then the state after pmotion and pinsert uploaded,
then the search; closest and closestd bit for bit.  At larger sizes against
the oracle, also bit for bit.
"""
import numpy as np
import pytest

from conftest import Fil, closest_names, gpu_params, load_golden

pytestmark = pytest.mark.gpu

NAMES = closest_names()


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", NAMES)
def test_closest_vs_golden(name):
    from qvort_amd import gpu
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name, meta["pcount0"], "->", meta["pcount"])
    with gpu.Context(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront0"], g["behind0"])
        c.velocity()
        c.upload(g["x"], g["infront"], g["behind"])
        cl, cd = c.closest(g["closest_in"], g["closestd_in"])
    assert np.array_equal(cl, g["closest"])
    assert np.array_equal(cd, g["closestd"], equal_nan=True)


@pytest.mark.parametrize("boundary", ["periodic", "open"])
def test_tree_closest_after_motion_vs_oracle(boundary):
    """200k points: the tree at x0, the search after one AB step on the
    device (positions downloaded and handed to the oracle as they are)."""
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.2, 5e-4, target_points=200_000, seed=31,
                         wrap=boundary == "periodic")
    n = fil.pcount
    kw = dict(velocity="Tree", tree_theta=0.4, boundary=boundary, box_size=0.2,
              corea=1.3e-8, dt=1e-5)
    _say("tree closest", boundary, n)
    rng = np.random.Generator(np.random.PCG64(5))
    cl0 = rng.integers(0, n, size=n).astype(np.int32)
    cd0 = rng.uniform(0, 1, size=n)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        c.ab_step()
        x_now = c.download()[0]
        cl, cd = c.closest(cl0, cd0)
    assert np.max(np.abs(x_now - fil.x)) > 0
    p = port.make_params(**kw)
    rcl, rcd = port.closest(p, fil, Fil(x_now, fil.infront, fil.behind), cl0, cd0,
                            nthreads=16)
    assert np.array_equal(cl, rcl)
    assert np.array_equal(cd, rcd)
    _say("closest=0:", int(np.sum(cl == 0)))


def test_direct_closest_vs_oracle():
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.with_holes(G.random_loops(0.1, 5e-4, target_points=20_000, seed=32), 300, seed=4)
    n = fil.pcount
    kw = dict(velocity="BS", tree_theta=0.0, boundary="periodic", box_size=0.1, corea=1.3e-8)
    cl0 = np.full(n, 3, np.int32)
    cd0 = np.full(n, 0.5)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        cl, cd = c.closest(cl0, cd0)
    rcl, rcd = port.closest(port.make_params(**kw), fil, fil, cl0, cd0, nthreads=16)
    assert np.array_equal(cl, rcl)
    assert np.array_equal(cd, rcd)
    act = fil.infront != 0
    assert np.all(cl[~act] == 3) and np.all(cd[~act] == 0.5)


def test_closest_state_errors():
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, n_loops=3, seed=33)
    with gpu.Context(gpu.params(velocity="Tree", tree_theta=0.4)) as c:
        c.upload_filaments(fil)
        with pytest.raises(gpu.QvortError) as ei:
            c.closest()
        assert ei.value.status == gpu.QV_ERR_STATE
        c.velocity()
        c.closest()
        m = fil.pcount - 10
        c.upload(fil.x[:m], np.zeros(m, np.int32), np.zeros(m, np.int32))
        with pytest.raises(gpu.QvortError) as ei:
            c.closest()
        assert ei.value.status == gpu.QV_ERR_STATE
