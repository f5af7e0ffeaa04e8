"""GPU parity: libqvort_gpu.so against the golden vectors of the reference
and against the oracle, through the C ABI.

Bars (FP64 path): tree occupancy, cell bounds, leaf indices and the eval
counter identical; the moments of every node of at most 8 points
identical; direct (BS) and tree velocities within 1e-12 relative per
point; AB trajectories within 1e-15 absolute.  FP32C
path: within 1e-5 relative per point of the FP64 reference.
"""
import numpy as np
import pytest

from conftest import Fil, golden_names, gpu_params, load_golden, oracle_params, rel_err

pytestmark = pytest.mark.gpu

NAMES = golden_names()


def _ctx(p):
    from qvort_amd import gpu
    return gpu.Context(p)


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", NAMES)
def test_fp64_velocity_vs_golden(name):
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name, meta["pcount"])
    with _ctx(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        u, uli, ubs, ev = c.velocity()
    act = g["infront"] != 0
    assert np.max(rel_err(uli, g["u_li"], act)) <= 1e-14
    if meta["params"]["velocity"] == "BS":
        # direct sum: the reference's pair term with refined reciprocals,
        # source chunks summed in a fixed order
        scale = np.linalg.norm(g["u"][act], axis=1)
        err = np.linalg.norm(ubs[act] - g["u_bs"][act], axis=1) / scale
        assert np.max(err) <= 1e-12
    else:
        assert ev == int(g["evals"][0])
        # empty slots of the reference's arrays hold whatever was there
        scale = np.linalg.norm(g["u"][act], axis=1)
        err = np.linalg.norm(ubs[act] - g["u_bs"][act], axis=1) / scale
        assert np.max(err) <= 1e-12
    assert np.max(rel_err(u, g["u"], act)) <= 1e-12
    assert not np.any(u[~act])


@pytest.mark.parametrize("name", [n for n in NAMES if "tree" in load_golden(n)])
def test_tree_nodes_vs_golden(name):
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name)
    with _ctx(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        c.velocity()
        t = c.tree_dump()
    _check_tree(t, g["tree"], meta["params"]["box_size"])


def _check_tree(t, ref, L):
    assert t.shape == ref.shape
    # level, pcount, leaf index, corner and width: bit-identical
    assert np.array_equal(t[:, 0:7], ref[:, 0:7])
    # moments of every node of at most 8 points (summed in slot order, as
    # new_tree sums them): bit-identical
    small = (ref[:, 1] <= 8) & (ref[:, 0] > 0)
    assert np.array_equal(t[small, 7:22], ref[small, 7:22])
    _say("nodes", len(ref), "bit-identical", int(small.sum()))
    # bigger nodes (summed in chunks): equal up to summation order
    assert np.max(np.abs(t[:, 7:10] - ref[:, 7:10])) <= 1e-15 * L
    circ_scale = np.max(np.abs(ref[:, 10:13])) + 1e-300
    assert np.max(np.abs(t[:, 10:13] - ref[:, 10:13])) <= 1e-13 * circ_scale
    c_scale = np.max(np.abs(ref[:, 13:22])) + 1e-300
    assert np.max(np.abs(t[:, 13:22] - ref[:, 13:22])) <= 1e-13 * c_scale


@pytest.mark.parametrize("name", NAMES)
def test_fp64_trajectory_vs_golden(name):
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name, meta["nsteps"])
    with _ctx(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        if "u1_0" in g:
            c.upload_history(g["u1_0"], g["u2_0"], g["u3_0"])
        c.step(meta["nsteps"])
        x, u, u1, u2, u3 = c.download()
    act = g["infront"] != 0
    assert np.max(np.abs(x - g["x_final"])[act]) <= 1e-15
    for a, b in ((u1, "u1_final"), (u2, "u2_final"), (u3, "u3_final")):
        ref = g[b]
        if np.any(ref[act]):
            assert np.max(rel_err(a, ref, act)) <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_fp32c_velocity_vs_golden(name):
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name, "fp32c")
    with _ctx(gpu_params(meta["params"], precision="fp32c")) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        u, uli, ubs, ev = c.velocity()
    act = g["infront"] != 0
    assert np.max(rel_err(uli, g["u_li"], act)) <= 1e-14
    e = rel_err(u, g["u"], act)
    _say("fp32c max/rms", e.max(), np.sqrt(np.mean(e ** 2)))
    assert np.max(e) <= 1e-5
    if meta["params"]["velocity"] == "Tree":
        assert ev == int(g["evals"][0])


@pytest.mark.parametrize("theta,boundary", [(0.3, "periodic"), (0.4, "periodic"),
                                            (0.6, "periodic"), (0.4, "open"), (0.4, "openx"),
                                            (0.5, "openxy")])
def test_fp64_tree_vs_oracle_20k(theta, boundary):
    from oracle import port
    from qvort_amd import geometry as G
    fil = G.random_loops(0.1, 5e-4, target_points=20000, seed=31, wrap=boundary != "open")
    kw = dict(velocity="Tree", tree_theta=theta, boundary=boundary, box_size=0.1,
              corea=1.3e-8)
    _say("oracle", fil.pcount, theta, boundary)
    u_o, uli_o, ubs_o, ev_o = port.velocity(port.make_params(**kw), fil, nthreads=16)
    _say("gpu")
    with _ctx(gpu_params(kw)) as c:
        c.upload_filaments(fil)
        u, uli, ubs, ev = c.velocity()
        t = c.tree_dump()
    assert ev == ev_o
    e = rel_err(u, u_o)
    _say("max", e.max())
    assert np.max(e) <= 1e-12
    _check_tree(t, port.tree_dump(port.make_params(**kw), fil), 0.1)
    with _ctx(gpu_params(kw, precision="fp32c")) as c:
        c.upload_filaments(fil)
        u32, _, _, ev32 = c.velocity()
    e32 = rel_err(u32, u_o)
    _say("fp32c max/rms", e32.max(), np.sqrt(np.mean(e32 ** 2)))
    assert ev32 == ev_o
    assert np.max(e32) <= 1e-5


def test_coincident_points_report_depth_error():
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    a = G.random_loops(0.1, 5e-4, n_loops=1, seed=41)
    n = a.pcount
    x = np.concatenate([a.x, a.x])
    inf = np.concatenate([a.infront, a.infront + n]).astype(np.int32)
    beh = np.concatenate([a.behind, a.behind + n]).astype(np.int32)
    with _ctx(gpu.params(velocity="Tree", tree_theta=0.4)) as c:
        c.upload(x, inf, beh)
        with pytest.raises(gpu.QvortError) as ei:
            c.velocity()
        assert ei.value.status == gpu.QV_ERR_TREE_DEPTH


def test_argument_and_state_errors():
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    a = G.random_loops(0.1, 5e-4, n_loops=1, seed=42)
    with _ctx(gpu.params(velocity="BS")) as c:
        bad = a.infront.copy()
        bad[3] = a.pcount + 5
        with pytest.raises(gpu.QvortError) as ei:
            c.upload(a.x, bad, a.behind)
        assert ei.value.status == gpu.QV_ERR_ARG
        c.upload_filaments(a)
        with pytest.raises(gpu.QvortError) as ei:
            c.ab_step()
        assert ei.value.status == gpu.QV_ERR_STATE
