"""Velocity at probe points on the GPU (qv_probe_velocity).

BS sums bit for bit against the oracle on the same inputs (positions as the
device holds them after a step, ghosts of the step's ghostp) and within
1e-12 of the reference fixtures; Tree sums within 1e-12 relative of the
oracle on the same tree (the device sums moments in another order).  The
analytic gradient against Richardson-extrapolated central differences of the
oracle (1e-6 relative) and against div u = 0, which holds for every
straight segment's field.
"""
import numpy as np
import pytest

from conftest import Fil, gpu_params, load_golden, probe_names

pytestmark = pytest.mark.gpu

NAMES = probe_names()


def _say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1))


@pytest.mark.parametrize("name", NAMES)
def test_probe_vs_golden(name):
    from qvort_amd import gpu
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name)
    with gpu.Context(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        c.velocity()
        if meta["motion"]:
            c.ab_step()
        u = c.probe_velocity(g["probes"], reference_scale=True)
    assert _rel(u, g["u"]) <= 1e-12


@pytest.mark.parametrize("boundary", ["periodic", "openx", "openxy", "open"])
def test_bs_probe_bitexact_vs_oracle(boundary):
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, n_loops=6, seed=51, wrap=boundary != "open")
    kw = dict(velocity="BS", tree_theta=0.0, boundary=boundary, box_size=0.1, corea=1.3e-8,
              dt=1e-5)
    rng = np.random.Generator(np.random.PCG64(52))
    y = rng.uniform(-0.05, 0.05, size=(1000, 3))
    p = port.make_params(**kw)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        c.ab_step()
        x_now = c.download()[0]
        u_stale = c.probe_velocity(y)
        u_fresh = c.probe_velocity(y, ghostp=True)
    gi0, _, _, _ = port.ghostp(p, fil)
    cur = Fil(x_now, fil.infront, fil.behind)
    gi1, _, _, _ = port.ghostp(p, cur)
    assert np.array_equal(u_stale, port.probe_velocity(p, fil, cur, gi0, y, nthreads=16))
    assert np.array_equal(u_fresh, port.probe_velocity(p, fil, cur, gi1, y, nthreads=16))


@pytest.mark.parametrize("theta", [0.3, 0.6])
def test_tree_probe_vs_oracle(theta):
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=20_000, seed=53)
    kw = dict(velocity="Tree", tree_theta=theta, boundary="periodic", box_size=0.1,
              corea=1.3e-8)
    rng = np.random.Generator(np.random.PCG64(54))
    y = rng.uniform(-0.05, 0.05, size=(2000, 3))
    p = port.make_params(**kw)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u = c.probe_velocity(y)
        ur = c.probe_velocity(y, reference_scale=True)
    gi, _, _, _ = port.ghostp(p, fil)
    ref = port.probe_velocity(p, fil, fil, gi, y, ref_scaling=False, nthreads=16)
    refr = port.probe_velocity(p, fil, fil, gi, y, ref_scaling=True, nthreads=16)
    _say("tree probe", theta, _rel(u, ref), _rel(ur, refr))
    assert _rel(u, ref) <= 1e-12
    assert _rel(ur, refr) <= 1e-12


def test_tree_probe_follows_theta_of_set_params():
    """The probe walk's acceptance thresholds are per tree and per theta: a
    new tree_theta through set_params applies to the next probe call on the
    same tree."""
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=5_000, seed=57)
    kw = dict(velocity="Tree", boundary="periodic", box_size=0.1, corea=1.3e-8)
    rng = np.random.Generator(np.random.PCG64(58))
    y = rng.uniform(-0.05, 0.05, size=(500, 3))
    with gpu.Context(gpu.params(tree_theta=0.4, **kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u4 = c.probe_velocity(y)
        c.set_params(gpu.params(tree_theta=0.6, **kw))
        u6 = c.probe_velocity(y)
    for th, u in ((0.4, u4), (0.6, u6)):
        p = port.make_params(tree_theta=th, **kw)
        gi, _, _, _ = port.ghostp(p, fil)
        ref = port.probe_velocity(p, fil, fil, gi, y, ref_scaling=False, nthreads=16)
        assert _rel(u, ref) <= 1e-12


@pytest.mark.parametrize("nprobe", [1, 63, 65, 130])
def test_tree_probe_odd_counts_vs_oracle(nprobe):
    """Partial groups of 64 and shell chunks of unequal length (27 shells
    over up to 27 wavefronts per group)."""
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=3_000, seed=59)
    kw = dict(velocity="Tree", tree_theta=0.5, boundary="openx", box_size=0.1, corea=1.3e-8)
    y = np.random.Generator(np.random.PCG64(60)).uniform(-0.06, 0.06, size=(nprobe, 3))
    p = port.make_params(**kw)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u = c.probe_velocity(y, reference_scale=True)
    gi, _, _, _ = port.ghostp(p, fil)
    ref = port.probe_velocity(p, fil, fil, gi, y, ref_scaling=True, nthreads=16)
    assert _rel(u, ref) <= 1e-12


def test_tree_probe_split_and_unsplit_agree():
    """Few probes split their shells over several wavefronts (partial sums
    added in shell order), many do not: the same probes agree both ways."""
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=20_000, seed=55)
    kw = dict(velocity="Tree", tree_theta=0.4, boundary="periodic", box_size=0.1,
              corea=1.3e-8)
    rng = np.random.Generator(np.random.PCG64(56))
    y = rng.uniform(-0.05, 0.05, size=(600_000, 3))
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u_all = c.probe_velocity(y)
        u_few = c.probe_velocity(y[:3000])
    _say("split vs unsplit", _rel(u_few, u_all[:3000]))
    assert _rel(u_few, u_all[:3000]) <= 1e-13


def test_gradient_vs_differences_and_divergence():
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, n_loops=5, seed=55)
    kw = dict(velocity="BS", tree_theta=0.0, boundary="periodic", box_size=0.1, corea=1.3e-8)
    rng = np.random.Generator(np.random.PCG64(56))
    y = rng.uniform(-0.05, 0.05, size=(64, 3))
    p = port.make_params(**kw)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u, G3 = c.probe_velocity(y, grad=True)
    gi, _, _, _ = port.ghostp(p, fil)

    def uo(yy):
        return port.probe_velocity(p, fil, fil, gi, yy, nthreads=16)

    assert np.array_equal(u, uo(y))
    h = 1e-6
    fd = np.zeros_like(G3)
    for m in range(3):
        d = []
        for hh in (h, 2 * h):
            e = np.zeros(3)
            e[m] = hh
            d.append((uo(y + e) - uo(y - e)) / (2 * hh))
        fd[:, :, m] = (4 * d[0] - d[1]) / 3
    scale = np.linalg.norm(G3.reshape(-1, 9), axis=1)
    err = np.linalg.norm((G3 - fd).reshape(-1, 9), axis=1) / scale
    _say("gradient rel err", err.max())
    assert err.max() <= 1e-6
    div = np.abs(np.trace(G3, axis1=1, axis2=2)) / scale
    assert div.max() <= 1e-10


def test_probe_state_and_argument_errors():
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, n_loops=2, seed=57)
    y = np.zeros((4, 3))
    with gpu.Context(gpu.params(velocity="BS")) as c:
        c.upload_filaments(fil)
        with pytest.raises(gpu.QvortError) as ei:
            c.probe_velocity(y)
        assert ei.value.status == gpu.QV_ERR_STATE
        c.probe_velocity(y, ghostp=True)
    with gpu.Context(gpu.params(velocity="Tree", tree_theta=0.4)) as c:
        c.upload_filaments(fil)
        with pytest.raises(gpu.QvortError) as ei:
            c.probe_velocity(y)
        assert ei.value.status == gpu.QV_ERR_STATE
        c.velocity()
        c.probe_velocity(y)
        with pytest.raises(gpu.QvortError) as ei:
            c.probe_velocity(y, grad=True)
        assert ei.value.status == gpu.QV_ERR_ARG


def test_empty_core_rule():
    """velocity_quasip's empty core (hamiltonian.f90:42-58): within corea of
    any slot's position -- an empty slot's too, since the reference loops
    over all of f(:) -- u and the gradient are zero; elsewhere the flag
    changes nothing.  Tree mode has no such rule in the reference and is
    refused."""
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.with_holes(G.random_loops(0.1, 5e-4, n_loops=4, seed=58), 10, seed=3)
    a = 1.3e-8
    kw = dict(velocity="BS", tree_theta=0.0, boundary="periodic", box_size=0.1, corea=a)
    act = np.flatnonzero(fil.infront)
    hole = np.flatnonzero(fil.infront == 0)[0]
    fil.x[hole] = [0.011, -0.007, 0.004]  # an empty slot's (stale) position
    rng = np.random.Generator(np.random.PCG64(59))
    dirs = rng.normal(size=(4, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    inside = np.stack([fil.x[act[10]] + 0.5 * a * dirs[0], fil.x[act[200]] + 0.9 * a * dirs[1],
                       fil.x[hole] + 0.5 * a * dirs[2]])
    outside = np.stack([fil.x[act[10]] + 2.0 * a * dirs[3],
                        rng.uniform(-0.04, 0.04, size=3)])
    y = np.concatenate([inside, outside])
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.velocity()
        u0, g0 = c.probe_velocity(y, grad=True)
        u1, g1 = c.probe_velocity(y, grad=True, empty_core=True)
    assert not np.any(u1[:3]) and not np.any(g1[:3])
    assert np.any(u0[:3])
    assert np.array_equal(u1[3:], u0[3:]) and np.array_equal(g1[3:], g0[3:])
    with gpu.Context(gpu.params(velocity="Tree", tree_theta=0.4, corea=a)) as c:
        c.upload_filaments(fil)
        c.velocity()
        with pytest.raises(gpu.QvortError) as ei:
            c.probe_velocity(y, empty_core=True)
        assert ei.value.status == gpu.QV_ERR_ARG
