"""The C restatement (oracle) against the golden vectors of the reference.

The fixtures under tests/golden were written by tests/golden/make_golden.py
from the flang build of the reference (oracle/_ref).  The oracle must match
them bit for bit: velocities, their local and nonlocal parts, the eval
counter, every node of the tree (bounds, occupancy, moments) and the AB
trajectories.
"""
import numpy as np
import pytest

from conftest import Fil, bnd_names, golden_names, load_golden, oracle_params

NAMES = golden_names()


@pytest.mark.parametrize("name", NAMES)
def test_velocity_bitexact(name, oracle_lib):
    from oracle import port
    g = load_golden(name)
    p = oracle_params(g["meta"]["params"])
    fil = Fil(g["x0"], g["infront"], g["behind"])
    u, uli, ubs, ev = port.velocity(p, fil, nthreads=2)
    assert np.array_equal(uli, g["u_li"])
    assert np.array_equal(ubs, g["u_bs"])
    assert np.array_equal(u, g["u"])
    if g["meta"]["params"]["velocity"] == "Tree":
        assert ev == int(g["evals"][0])


@pytest.mark.parametrize("name", NAMES)
def test_ghosts_bitexact(name, oracle_lib):
    from oracle import port
    g = load_golden(name)
    p = oracle_params(g["meta"]["params"])
    fil = Fil(g["x0"], g["infront"], g["behind"])
    gi, gb, _, _ = port.ghostp(p, fil)
    act = fil.infront != 0
    assert np.array_equal(gi[act], g["ghosti"][act])
    assert np.array_equal(gb[act], g["ghostb"][act])


@pytest.mark.parametrize("name", [n for n in NAMES if "tree" in load_golden(n)])
def test_tree_nodes_bitexact(name, oracle_lib):
    from oracle import port
    g = load_golden(name)
    p = oracle_params(g["meta"]["params"])
    fil = Fil(g["x0"], g["infront"], g["behind"])
    t = port.tree_dump(p, fil)
    assert t.shape == g["tree"].shape
    assert np.array_equal(t, g["tree"])


@pytest.mark.parametrize("name", NAMES)
def test_trajectory_bitexact(name, oracle_lib):
    from oracle import port
    g = load_golden(name)
    meta = g["meta"]
    p = oracle_params(meta["params"])
    fil = Fil(g["x0"], g["infront"], g["behind"])
    h = [g.get(f"u{k}_0") for k in (1, 2, 3)]
    x, u1, u2, u3, _ = port.run(p, fil, meta["nsteps"], *h, nthreads=2)
    assert np.array_equal(x, g["x_final"])
    assert np.array_equal(u1, g["u1_final"])
    assert np.array_equal(u2, g["u2_final"])
    assert np.array_equal(u3, g["u3_final"])


@pytest.mark.parametrize("name", bnd_names())
def test_boundary_step_bitexact(name, oracle_lib):
    """Loops that leave an openx / openxy box through an open face are removed
    (boundary_loop_remove) and remove_mean_pos recentres x and y: positions,
    AB history, links and the removal counters of the reference, bit for bit."""
    from oracle import port
    g = load_golden(name)
    meta = g["meta"]
    p = oracle_params(meta["params"])
    fil = Fil(g["x0"], g["infront"], g["behind"])
    x, u1, u2, u3, uf, inf, beh, cnt, length = port.run(p, fil, meta["nsteps"], nthreads=2,
                                                        topology=True)
    assert np.array_equal(uf, g["u"])
    assert np.array_equal(inf, g["infront_final"])
    assert np.array_equal(beh, g["behind_final"])
    assert cnt == int(g["boundary"][0])
    assert length == float(g["boundary_length"][0])
    assert np.array_equal(x, g["x_final"])
    assert np.array_equal(u1, g["u1_final"])
    assert np.array_equal(u2, g["u2_final"])
    assert np.array_equal(u3, g["u3_final"])


def test_boundary_cases_remove_loops_and_recentre():
    """What the fixtures hold: loops were removed in the open-face cases (so
    the removal is exercised), none in the others, and remove_mean_pos left
    the mean x and y of the points at zero to rounding."""
    for name in bnd_names():
        g = load_golden(name)
        act = g["infront_final"] != 0
        if "openface" in name:
            assert int(g["boundary"][0]) >= 2
            assert act.sum() < np.count_nonzero(g["infront"])
            # every point left lies inside the box along the open axes
            L = g["meta"]["params"]["box_size"]
            axes = [0] if g["meta"]["params"]["boundary"] == "openx" else [0, 1]
            assert np.all(np.abs(g["x_final"][act][:, axes]) <= L / 2)
            # cleared slots hold zeros (clear_particle)
            gone = (g["infront"] != 0) & ~act
            assert not np.any(g["x_final"][gone]) and not np.any(g["u1_final"][gone])
        else:
            assert int(g["boundary"][0]) == 0
            assert np.all(np.abs(g["x_final"][act][:, :2].mean(axis=0)) < 1e-17)
