"""The C restatement (oracle) against the golden vectors of the reference.

The fixtures under tests/golden were written by tests/golden/make_golden.py
from the flang build of the reference (oracle/_ref).  The oracle must match
them bit for bit: velocities, their local and nonlocal parts, the eval
counter, every node of the tree (bounds, occupancy, moments) and the AB
trajectories.
"""
import numpy as np
import pytest

from conftest import Fil, golden_names, load_golden, oracle_params

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
