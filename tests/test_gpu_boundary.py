"""The boundary step through the C ABI against the reference: loops that
leave an openx / openxy box through an open face are removed
(enforce_periodic_yz / _z with boundary_loop_remove, periodic.f90:217-327),
and remove_mean_pos recentres x and y after each AB update
(timestep.f90:86-92).  Fixtures: tests/golden/bnd_*.npz from the reference
build.

Bars: links after the run identical, the removal count identical and its
length to 1e-13 relative; positions within 1e-15 cm; cleared slots hold
zeros; the AB history of the points left within 1e-12 relative."""
import numpy as np
import pytest

from conftest import bnd_names, gpu_params, load_golden, rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", bnd_names())
def test_boundary_step_vs_golden(name):
    from qvort_amd import gpu
    g = load_golden(name)
    meta = g["meta"]
    _say("case", name, meta["pcount"], meta["nsteps"])
    with gpu.Context(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        c.step(meta["nsteps"])
        x, u, u1, u2, u3 = c.download()
        inf, beh = c.download_links()
        cnt, length = c.boundary_removed()
    assert np.array_equal(inf, g["infront_final"])
    assert np.array_equal(beh, g["behind_final"])
    act = g["infront_final"] != 0
    assert cnt == int(g["boundary"][0])
    ref_len = float(g["boundary_length"][0])
    assert abs(length - ref_len) <= 1e-13 * max(ref_len, 1e-300)
    assert np.max(np.abs(x - g["x_final"])[act]) <= 1e-15
    gone = (g["infront"] != 0) & ~act
    assert not np.any(x[gone]) and not np.any(u1[gone]) and not np.any(u2[gone])
    for a, b in ((u1, "u1_final"), (u2, "u2_final"), (u3, "u3_final")):
        if np.any(g[b][act]):
            assert np.max(rel_err(a, g[b], act)) <= 1e-12
    _say("removed", cnt, "length", length)


def test_boundary_step_with_line_state_uploads():
    """The removal also runs when the state arrives by upload (the topology
    step's path): a loop across the open face of an openx box is removed by
    qv_enforce_periodic after qv_upload, and the oracle's boundary step on
    the same arrays gives the same links, positions and counters."""
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    L = 0.1
    ring = G.single_ring(40, 0.004)
    ring.x = ring.x + np.array([L / 2 - 0.001, 0.0, 0.0])
    fil = G.combine(G.random_loops(L, 5e-4, n_loops=2, seed=3), ring)
    kw = dict(velocity="BS", boundary="openx", box_size=L, corea=1.3e-8)
    with gpu.Context(gpu.params(**kw)) as c:
        c.upload_filaments(fil)
        c.enforce_periodic()
        x, _, _, _, _ = c.download()
        inf_d, beh_d = c.download_links()
        cnt, length = c.boundary_removed()
    z = np.zeros_like(fil.x)
    xo, inf, beh, _, _, co, lo = port.enforce_boundary(port.make_params(**kw), fil.x,
                                                      fil.infront, fil.behind, z, z)
    assert (cnt, length) == (co, lo) and cnt == 1
    assert np.array_equal(inf_d, inf) and np.array_equal(beh_d, beh)
    assert int(np.count_nonzero(inf)) == fil.n_active - 40
    act = inf != 0
    assert np.array_equal(x[act], xo[act]) and not np.any(x[~act])
