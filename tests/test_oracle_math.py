"""Pins of the oracle that do not depend on the reference run: closed forms,
invariants and brute force.  Each is chosen so that a plausible slip (a
dropped term, a wrong sign, a transposed operand, a wrong exclusion) fails.
"""
import math

import numpy as np
import pytest

from conftest import Fil, rel_err

from qvort_amd import geometry as G


def _params(**kw):
    from oracle import port
    return port.make_params(**kw)


def _quad_segment(xi, a, b, gamma, n=4000):
    """Gamma/(4 pi) int (s - x) x ds / |s - x|^3 over a->b, Gauss-Legendre."""
    t, w = np.polynomial.legendre.leggauss(n)
    t = 0.5 * (t + 1.0)
    w = 0.5 * w
    s = a[None, :] + t[:, None] * (b - a)[None, :]
    r = s - xi[None, :]
    ds = (b - a)[None, :]
    integrand = np.cross(r, ds) / np.linalg.norm(r, axis=1)[:, None] ** 3
    return gamma / (4 * math.pi) * (w[:, None] * integrand).sum(axis=0)


@pytest.mark.parametrize("form", [0, 1])
def test_segment_matches_quadrature(form, oracle_lib):
    from oracle import port
    rng = np.random.Generator(np.random.PCG64(1))
    p = _params(quant_circ=9.97e-4)
    for _ in range(20):
        xi = rng.uniform(-1e-3, 1e-3, 3)
        xj = rng.uniform(-1e-3, 1e-3, 3)
        gj = xj + rng.normal(0, 4e-4, 3)
        if np.linalg.norm(np.cross(xj - xi, gj - xj)) < 1e-9:
            continue
        u = port.segment_velocity(p, xi, xj, gj, form=form)
        q = _quad_segment(xi, xj, gj, 9.97e-4)
        assert np.linalg.norm(u - q) <= 1e-9 * np.linalg.norm(q)


def test_segment_forms_agree_and_antisymmetric(oracle_lib):
    from oracle import port
    rng = np.random.Generator(np.random.PCG64(2))
    p = _params()
    for _ in range(50):
        xi, xj = rng.uniform(-0.05, 0.05, 3), rng.uniform(-0.05, 0.05, 3)
        gj = xj + rng.normal(0, 4e-4, 3)
        sh = rng.choice([-0.1, 0.0, 0.1], size=3)
        u0 = port.segment_velocity(p, xi, xj, gj, sh, form=0)
        u1 = port.segment_velocity(p, xi, xj, gj, sh, form=1)
        ur = port.segment_velocity(p, xi, gj, xj, sh, form=1)  # reversed segment
        assert np.linalg.norm(u0 - u1) <= 1e-11 * np.linalg.norm(u1)
        assert np.linalg.norm(u1 + ur) <= 1e-12 * np.linalg.norm(u1)


def test_ring_translation_speed(oracle_lib):
    """Ring of radius R: speed Gamma/(4 pi R)(ln(8R/a) - 1/2)
    (initial_loop.f90:17), from the LIA term plus the direct BS sum."""
    from oracle import port
    R, n, a = 0.05, 512, 1.3e-8
    fil = G.single_ring(n, R)
    p = _params(velocity="BS", boundary="open", box_size=0.2, corea=a)
    u, uli, ubs, _ = port.velocity(p, fil)
    uz = u[:, 2]
    ref = G.ring_speed(R, 9.97e-4, a)
    assert np.allclose(np.abs(uz), ref, rtol=5e-3)
    assert np.all(np.sign(uz) == np.sign(uz[0]))
    # in-plane components vanish by symmetry
    assert np.max(np.abs(u[:, :2])) < 1e-9 * ref


def test_lia_ring_curvature(oracle_lib):
    """u_LI = Gamma/(4 pi) ln((1.21/a) sqrt(l_i l_b)) kappa b: for a ring the
    3-point stencil gives kappa = 1/R up to O(l^2/R^2)."""
    from oracle import port
    R, n, a = 0.03, 2048, 1.3e-8
    fil = G.single_ring(n, R)
    p = _params(velocity="BS", boundary="open", box_size=0.2, corea=a)
    _, uli, _, _ = port.velocity(p, fil)
    ell = 2 * R * math.sin(math.pi / n)
    beta = 9.97e-4 / (4 * math.pi) * math.log(1.21 / a * ell)
    assert np.allclose(np.abs(uli[:, 2]), beta / R, rtol=1e-5)


def test_straight_line_no_self_velocity(oracle_lib):
    """A straight periodic line has zero self-induced velocity (LIA term
    zero, nonlocal sum cancels by symmetry)."""
    from oracle import port
    fil = G.kelvin_line(200, 0.1, amp_sigma=0.0)
    for vel, th in (("BS", 0.0), ("Tree", 0.4)):
        p = _params(velocity=vel, tree_theta=th, boundary="periodic", box_size=0.1)
        u, uli, ubs, _ = port.velocity(p, fil)
        assert np.max(np.abs(uli)) == 0.0
        assert np.max(np.abs(u)) < 1e-12


def _direct_tree_rule(fil, p_gamma, L, shells):
    """numpy brute force of the leaf formula over every source and shell with
    the tree's exclusions {i, behind(i), infront(i)} (tree.f90:368-371)."""
    x = fil.x
    n = x.shape[0]
    g = np.empty_like(x)
    g[:] = x[fil.infront - 1]
    for k in range(3):
        d = x[:, k] - g[:, k]
        g[:, k] = np.where(d > L / 2, g[:, k] + L, np.where(d < -L / 2, g[:, k] - L, g[:, k]))
    u = np.zeros((n, 3))
    idx = np.arange(n)
    for i in range(n):
        excl = {i, fil.behind[i] - 1, fil.infront[i] - 1}
        m = np.array([j not in excl for j in idx])
        for s in shells:
            R1 = x[m] + s - x[i]
            R2 = g[m] + s - x[i]
            a = np.linalg.norm(R1, axis=1)
            b = np.linalg.norm(R2, axis=1)
            c = a * b * (a * b + (R1 * R2).sum(axis=1))
            u[i] += (((a + b) / c)[:, None] * np.cross(R1, R2)).sum(axis=0)
    return u * p_gamma / (4 * math.pi)


def test_tree_small_theta_is_direct_sum(oracle_lib):
    """theta -> 0+ opens every cell: the walk reduces to the leaf sum over all
    segments with the tree's exclusion rule."""
    from oracle import port
    fil = G.random_loops(0.1, 5e-4, n_loops=2, seed=21)
    L = 0.1
    shells = [np.array([a, b, c]) * L for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    p = _params(velocity="Tree", tree_theta=1e-9, boundary="periodic", box_size=L)
    _, _, ubs, ev = port.velocity(p, fil)
    ref = _direct_tree_rule(fil, 9.97e-4, L, shells)
    assert np.max(rel_err(ubs, ref)) < 1e-11
    n = fil.pcount
    assert ev == 2 * n + 27 * n * (n - 3) + 26 * 3 * n - 26 * 3 * n  # every leaf, 3 excluded


def test_leaf_sum_is_small_theta_tree(oracle_lib):
    """The leaf sum (the truncation-error reference) equals the theta -> 0+
    walk up to round-off, with the same number of leaf terms, and the
    numpy brute force of the tree's exclusion rule."""
    from oracle import port
    fil = G.random_loops(0.1, 5e-4, n_loops=2, seed=22)
    L = 0.1
    p = _params(velocity="Tree", tree_theta=1e-9, boundary="periodic", box_size=L)
    u, uli, ubs, ev = port.velocity(p, fil)
    act = np.flatnonzero(fil.infront).astype(np.int32)
    ul, evl = port.leaf_sum_targets(p, fil, act)
    assert np.max(rel_err(ul, u[act])) < 1e-11
    assert evl == ev - 2 * act.size
    shells = [np.array([a, b, c]) * L for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    ref = _direct_tree_rule(fil, 9.97e-4, L, shells)
    assert np.max(rel_err(ul - uli[act], ref[act])) < 1e-11


def test_bs_brute_force_tiny(oracle_lib):
    """Direct sum (biot_savart + 26 shifted copies incl. the self image,
    timestep.f90:462-507) against an independent numpy evaluation."""
    from oracle import port
    fil = G.random_loops(0.1, 5e-4, n_loops=1, seed=23, rmin=0.002, rmax=0.003)
    L = 0.1
    x, n = fil.x, fil.pcount
    g = x[fil.infront - 1]
    p = _params(velocity="BS", boundary="periodic", box_size=L)
    _, _, ubs, _ = port.velocity(p, fil)
    ref = np.zeros((n, 3))
    for i in range(n):
        for s in [np.array([a, b, c]) * L for a in (-1, 0, 1) for b in (-1, 0, 1)
                  for c in (-1, 0, 1)]:
            central = not s.any()
            for j in range(n):
                if j == fil.behind[i] - 1 or (central and j == i):
                    continue
                r = x[j] + s - x[i]
                e = g[j] - x[j]
                A, B, Cc = r @ r, 2 * r @ e, e @ e
                D = 4 * A * Cc - B * B
                ref[i] += (np.cross(r, e) * 9.97e-4 / (2 * math.pi * D) *
                           ((2 * Cc + B) / math.sqrt(A + B + Cc) - B / math.sqrt(A)))
    assert np.max(rel_err(ubs, ref)) < 1e-11


def test_translation_invariance_open(oracle_lib):
    from oracle import port
    fil = G.random_loops(0.2, 5e-4, n_loops=3, seed=25, wrap=False, rmax=0.01)
    fil.x *= 0.5
    p = _params(velocity="BS", boundary="open", box_size=0.4)
    u0, _, _, _ = port.velocity(p, fil)
    fil2 = Fil(fil.x + np.array([0.0123, -0.0071, 0.0033]), fil.infront, fil.behind)
    u1, _, _, _ = port.velocity(p, fil2)
    assert np.max(rel_err(u1, u0)) < 1e-9


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_ab_ladder_polynomial_exactness(order, oracle_lib):
    """The k-step branch of the AB ladder integrates a velocity polynomial of
    degree k-1 in time exactly (Euler, AB2, AB3, AB4 of timestep.f90:51-61)."""
    from oracle import port
    rng = np.random.Generator(np.random.PCG64(order))
    dt = 1e-3
    n = 5
    coef = rng.normal(size=(order, n, 3))

    def vel(t):
        return sum(coef[k] * t ** k for k in range(order))

    def integral(t0, t1):
        return sum(coef[k] * (t1 ** (k + 1) - t0 ** (k + 1)) / (k + 1) for k in range(order))

    t = 0.37
    u = vel(t)
    hist = [vel(t - dt * (k + 1)) if k < order - 1 else np.zeros((n, 3)) for k in range(3)]
    x0 = rng.normal(size=(n, 3))
    p = _params(dt=dt)
    inf = np.arange(2, n + 2, dtype=np.int32)
    inf[-1] = 1
    x1, u1, u2, u3 = port.ab_update(p, x0, inf, u, *hist)
    assert np.allclose(x1 - x0, integral(t, t + dt), rtol=1e-9, atol=1e-15)
    assert np.array_equal(u1, u) and np.array_equal(u2, hist[0]) and np.array_equal(u3, hist[1])


def test_planted_mistake_is_caught():
    """The parity measure used by the GPU tests flags a 1e-11 relative change
    of one point's velocity (checks the checker, not the product)."""
    rng = np.random.Generator(np.random.PCG64(3))
    ref = rng.normal(size=(100, 3))
    bad = ref.copy()
    bad[17] *= 1 + 1e-11
    assert np.max(rel_err(bad, ref)) > 1e-12
    assert np.max(rel_err(ref, ref)) == 0.0
