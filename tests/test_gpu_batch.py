"""Ensembles of independent systems advanced in one persistent launch per step.

Per system, the FP64 ensemble velocity (the reference's biot_savart order)
and the single-system GPU velocity are within 1e-12 relative of the
oracle's; trajectories follow the oracle's velocity + AB ladder + wrap
within 1e-15 cm.  FP32C stays within 1e-5 relative of FP64 per point.  The
per-target sums run in slot order whatever the tile shape, so every tile
shape and the graph-replayed step give identical results.
"""
import numpy as np
import pytest

from conftest import Fil, rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _open_systems(seed=71):
    """Rings of uneven sizes (tile edges 64/65/256/257 included), a ring
    with empty slots and a three-point loop, open boundaries."""
    from qvort_amd import geometry as G
    rings = G.ring_ensemble(5, nmin=60, nmax=900, seed=seed)
    for n in (64, 65, 256, 257):
        rings.append(G.ring_ensemble(1, nmin=n, nmax=n, seed=seed + n)[0])
    rings.append(G.with_holes(G.ring_ensemble(1, nmin=300, nmax=300, seed=seed + 1)[0], 17))
    rings.append(G.single_ring(3, 0.03))
    return rings


def _periodic_systems(seed=81):
    from qvort_amd import geometry as G
    return [G.random_loops(0.1, 5e-4, n_loops=k, seed=seed + k, wrap=True) for k in (1, 2, 3)]


def _kw(boundary):
    return dict(velocity="BS", boundary=boundary, box_size=0.1 if boundary != "open" else 0.2,
                corea=1.3e-8, dt=1e-5 if boundary == "open" else 1e-6)


def _batch(kw, systems, precision="fp64", order=2):
    from qvort_amd import gpu
    return gpu.Batch(gpu.params(precision=precision, deriv_order=order, **kw), systems)


def _oracle(kw, order=2):
    from oracle import port
    return port.make_params(deriv_order=order, **kw)


@pytest.mark.parametrize("boundary,order", [("open", 2), ("open", 4), ("periodic", 2)])
def test_fp64_velocity_per_system(boundary, order):
    from oracle import port
    from qvort_amd import gpu
    systems = _open_systems() if boundary == "open" else _periodic_systems()
    kw = _kw(boundary)
    _say("ensemble", boundary, order, [s.pcount for s in systems])
    with _batch(kw, systems, order=order) as b:
        us = b.velocity()
    p = _oracle(kw, order)
    for s, u in zip(systems, us):
        act = s.infront != 0
        with gpu.Context(gpu.params(deriv_order=order, **kw)) as c:
            c.upload_filaments(s)
            u1, _, _, _ = c.velocity()
        ref, _, _, _ = port.velocity(p, Fil(s.x, s.infront, s.behind), nthreads=16)
        assert np.max(rel_err(u, ref, act)) <= 1e-12
        assert np.max(rel_err(u1, ref, act)) <= 1e-12, s.pcount
        assert not np.any(u[~act])


@pytest.mark.parametrize("boundary", ["open", "periodic"])
def test_fp32c_velocity_error(boundary):
    systems = _open_systems(72) if boundary == "open" else _periodic_systems(82)
    kw = _kw(boundary)
    with _batch(kw, systems) as b:
        u64 = b.velocity()
    with _batch(kw, systems, "fp32c") as b:
        u32 = b.velocity()
    worst = 0.0
    for s, a, r in zip(systems, u32, u64):
        worst = max(worst, float(np.max(rel_err(a, r, s.infront != 0))))
    _say("fp32c max rel err", worst)
    assert worst <= 1e-5


@pytest.mark.parametrize("boundary", ["open", "periodic"])
def test_fp64_trajectory_vs_oracle(boundary):
    """6 steps: Euler, AB2, AB3, then AB4, each followed by the wrap."""
    from oracle import port
    systems = _open_systems(73)[:6] if boundary == "open" else _periodic_systems(83)
    kw = _kw(boundary)
    nsteps = 6
    with _batch(kw, systems) as b:
        b.step(nsteps, graph=False)
        xs, _ = b.download()
    p = _oracle(kw)
    for s, x in zip(systems, xs):
        xr, _, _, _, _ = port.run(p, Fil(s.x, s.infront, s.behind), nsteps, nthreads=16)
        act = s.infront != 0
        assert np.max(np.abs(x[act] - xr[act])) <= 1e-15, s.pcount


def test_graph_replays_match_eager_steps():
    """A captured 10-step graph, replayed three times with eager work and
    synchronisations between the replays, against the same steps launched
    one by one; plus a remainder of eager steps after the replays."""
    systems = _open_systems(74)
    kw = _kw("open")
    with _batch(kw, systems, "fp32c") as g, _batch(kw, systems, "fp32c") as e:
        for r in range(3):
            g.step(10, graph=True)
            e.step(10, graph=False)
            xg, ug = g.download()
            xe, ue = e.download()
            for a, c in zip(xg + ug, xe + ue):
                assert np.array_equal(a, c), r
        g.step(13, graph=True)
        e.step(13, graph=False)
        xg, _ = g.download()
        xe, _ = e.download()
        for a, c in zip(xg, xe):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_tile_shapes_identical_and_tuner(precision):
    systems = _open_systems(75)
    kw = _kw("open")
    with _batch(kw, systems, precision) as b:
        nv = b.info()["nvariants"]
        base = None
        for v in range(nv):
            b.set_variant(v)
            u = b.velocity()
            if base is None:
                base = u
            for a, c in zip(u, base):
                assert np.array_equal(a, c), v
        chosen, ms = b.tune(2)
        _say("tile times (ms)", ms, "chosen", chosen)
        assert len(ms) == nv and all(t > 0 for t in ms)
        assert ms[chosen] == min(ms)
        assert b.info()["variant"] == chosen
        for a, c in zip(b.velocity(), base):
            assert np.array_equal(a, c)


def test_ring_speeds_in_ensemble():
    """Unperturbed rings of several radii move along their axis at
    Gamma/(4 pi R)(ln(8R/a) - 1/2) (initial_loop.f90:17)."""
    from qvort_amd import geometry as G
    radii = [0.02, 0.03, 0.045, 0.06]
    systems = [G.single_ring(int(round(2 * np.pi * R / 6e-4)), R) for R in radii]
    kw = _kw("open")
    nsteps = 20
    with _batch(kw, systems, "fp32c") as b:
        x0 = [s.x.copy() for s in systems]
        b.step(nsteps, graph=True)
        xs, us = b.download()
    for R, s, a, x in zip(radii, systems, x0, xs):
        ref = G.ring_speed(R, 9.97e-4, kw["corea"])
        v = np.mean(x[:, 2] - a[:, 2]) / (nsteps * kw["dt"])
        assert abs(abs(v) - ref) <= 5e-3 * ref, (R, v, ref)


def test_rejects_links_outside_system():
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    systems = _open_systems(76)[:3]
    bad = G.Filaments(systems[1].x.copy(), systems[1].infront.copy(), systems[1].behind.copy(),
                      systems[1].loop_start, systems[1].loop_len)
    bad.infront[0] = bad.pcount + 1  # a slot of the next system
    with pytest.raises(gpu.QvortError) as ei:
        _batch(_kw("open"), [systems[0], bad, systems[2]])
    assert ei.value.status == gpu.QV_ERR_ARG
