"""Adjoint (VJP) and tangent (JVP) of the direct-sum velocity, and multi-step
gradients through the PyTorch harness.

The finite differences use the oracle's velocity (bit-identical to the flang
build of the reference), so the GPU derivative is checked against the
reference's own function.  Bars: 1e-6 relative for derivatives against
central differences (h = 1e-7 cm), 1e-10 for the VJP/JVP duality, 1e-5 for
the mixed-precision adjoint against the FP64 one.
"""
import numpy as np
import pytest

from conftest import Fil

pytestmark = pytest.mark.gpu

H = 1e-7


def _case(boundary="periodic", order=2, n_loops=2, seed=51):
    from qvort_amd import geometry as G
    fil = G.random_loops(0.1, 5e-4, n_loops=n_loops, seed=seed, wrap=boundary == "periodic")
    kw = dict(velocity="BS", boundary=boundary, box_size=0.1, corea=1.3e-8)
    return fil, kw, order


def _oracle_u(kw, order, x, fil):
    from oracle import port
    p = port.make_params(deriv_order=order, **kw)
    u, _, _, _ = port.velocity(p, Fil(x, fil.infront, fil.behind), nthreads=16)
    return u


def _ctx(kw, order, fil, precision="fp64"):
    from qvort_amd import gpu
    c = gpu.Context(gpu.params(deriv_order=order, precision=precision, **kw))
    c.upload_filaments(fil)
    return c


@pytest.mark.parametrize("boundary,order", [("periodic", 2), ("open", 4), ("openx", 2)])
def test_vjp_against_oracle_differences(boundary, order):
    fil, kw, order = _case(boundary, order)
    print("vjp", boundary, order, fil.pcount, flush=True)
    rng = np.random.Generator(np.random.PCG64(1))
    w = rng.normal(size=fil.x.shape)
    with _ctx(kw, order, fil) as c:
        g = c.vjp(w)
    act = np.flatnonzero(fil.infront)
    ks = rng.choice(act, size=10, replace=False)
    scale = np.max(np.abs(g))
    for i in ks:
        for a in range(3):
            xp, xm = fil.x.copy(), fil.x.copy()
            xp[i, a] += H
            xm[i, a] -= H
            fd = (np.sum(w * _oracle_u(kw, order, xp, fil)) -
                  np.sum(w * _oracle_u(kw, order, xm, fil))) / (2 * H)
            assert abs(g[i, a] - fd) <= 1e-6 * max(abs(fd), 1e-3 * scale), (i, a, g[i, a], fd)


def test_jvp_against_differences_and_duality():
    fil, kw, order = _case()
    rng = np.random.Generator(np.random.PCG64(2))
    v = rng.normal(size=fil.x.shape)
    w = rng.normal(size=fil.x.shape)
    with _ctx(kw, order, fil) as c:
        jv = c.jvp(v)
        jtw = c.vjp(w)
    fd = (_oracle_u(kw, order, fil.x + H * v, fil) - _oracle_u(kw, order, fil.x - H * v, fil)) / (2 * H)
    assert np.linalg.norm(jv - fd) <= 1e-6 * np.linalg.norm(fd)
    lhs, rhs = np.sum(w * jv), np.sum(jtw * v)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_mixed_vjp_against_fp64():
    fil, kw, order = _case(n_loops=3, seed=53)
    rng = np.random.Generator(np.random.PCG64(3))
    w = rng.normal(size=fil.x.shape)
    with _ctx(kw, order, fil) as c:
        g64 = c.vjp(w)
    with _ctx(kw, order, fil, precision="fp32c") as c:
        g32 = c.vjp(w)
    e = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print("mixed vjp rel err", e, flush=True)
    assert e <= 1e-5


def test_tree_mode_rejected():
    from qvort_amd import gpu
    fil, kw, order = _case()
    kw["velocity"] = "Tree"
    with gpu.Context(gpu.params(tree_theta=0.4, **kw)) as c:
        c.upload_filaments(fil)
        with pytest.raises(gpu.QvortError) as ei:
            c.vjp(np.zeros(fil.x.shape))
        assert ei.value.status == gpu.QV_ERR_ARG


@pytest.mark.parametrize("objective", ["length", "energy"])
def test_multistep_gradient_against_differences(objective):
    """Kelvin-wave line (512 points, 16 waves), 40 AB steps checkpointed every
    10: autograd through the VJP against central differences of the rollout.

    The differences are Richardson-extrapolated over h and 2h, (4 D(h) -
    D(2h)) / 3, which removes their h^2 J'''/6 truncation term: with the
    curvature on the scale of the point spacing (2e-4 cm) that term alone is
    ~1e-12.  The three largest components must agree to 1e-6 relative; six
    random ones (typically 1e-3 of the largest: sums of contributions that
    cancel) to 1e-6 of max(|fd|, 1e-2 max|g|)."""
    from qvort_amd import adjoint as AD
    from qvort_amd import geometry as G
    fil = G.kelvin_line(512, 0.1, amp_sigma=2e-4, seed=61)
    eng = AD.Engine(fil, box_size=0.1, boundary="periodic")
    try:
        dt = 0.25 * AD.dt_max(0.1 / 512)
        nsteps = 40
        J, g = AD.gradient(objective, fil.x, eng, nsteps, dt, checkpoint_every=10)
        print(objective, "J", J, "dt", dt, flush=True)
        scale = np.max(np.abs(g))
        top = np.argsort(np.abs(g).ravel())[::-1][:3]
        comps = [(int(k) // 3, int(k) % 3, True) for k in top]
        rng = np.random.Generator(np.random.PCG64(4))
        comps += [(int(i), a, False) for i in rng.choice(fil.pcount, size=3, replace=False)
                  for a in (0, 2)]
        for i, a, strict in comps:
            d = []
            for h in (H, 2 * H):
                xp, xm = fil.x.copy(), fil.x.copy()
                xp[i, a] += h
                xm[i, a] -= h
                d.append((AD.objective_value(objective, xp, eng, nsteps, dt) -
                          AD.objective_value(objective, xm, eng, nsteps, dt)) / (2 * h))
            fd = (4 * d[0] - d[1]) / 3
            print(i, a, strict, g[i, a], fd, d[0], flush=True)
            floor = 0.0 if strict else 1e-2 * scale
            assert abs(g[i, a] - fd) <= 1e-6 * max(abs(fd), floor)
    finally:
        eng.close()


@pytest.mark.parametrize("objective", ["length", "energy"])
def test_kelvin_8192_200_steps_gradient(objective):
    """The adjoint configuration at its size: a periodic line of 8192 points
    along z (L = 0.1 cm) with 16 Kelvin waves of amplitude ~N(0, (2e-4 cm)^2),
    200 AB steps checkpointed every 20, the gradient of the objective against
    central differences with h = 1e-7 cm (Richardson-extrapolated over h and
    2h, as above) for the largest component and two others, to 1e-6; and the
    mixed-precision gradient (FP32 pair terms) within 1e-5 of the FP64 one
    relative to its largest component."""
    import time

    from qvort_amd import adjoint as AD
    from qvort_amd import geometry as G
    fil = G.kelvin_line(8192, 0.1, amp_sigma=2e-4, seed=20261202)
    eng = AD.Engine(fil, box_size=0.1, boundary="periodic")
    eng32 = AD.Engine(fil, box_size=0.1, boundary="periodic", precision="fp32c")
    try:
        dt = 0.25 * AD.dt_max(0.1 / 8192)
        nsteps = 200
        t0 = time.perf_counter()
        J, g = AD.gradient(objective, fil.x, eng, nsteps, dt, checkpoint_every=20)
        print(objective, "J", J, "dt", dt, f"gradient {time.perf_counter() - t0:.1f} s",
              flush=True)
        scale = np.max(np.abs(g))
        _, g32 = AD.gradient(objective, fil.x, eng32, nsteps, dt, checkpoint_every=20)
        d32 = np.max(np.abs(g32 - g)) / scale
        print("fp32c gradient vs fp64:", d32, flush=True)
        assert d32 <= 1e-5
        top = int(np.argmax(np.abs(g).ravel()))
        rng = np.random.Generator(np.random.PCG64(8))
        comps = [(top // 3, top % 3, True)]
        comps += [(int(i), 0, False) for i in rng.choice(fil.pcount, size=2, replace=False)]
        for i, a, strict in comps:
            d = []
            for h in (H, 2 * H):
                xp, xm = fil.x.copy(), fil.x.copy()
                xp[i, a] += h
                xm[i, a] -= h
                d.append((AD.objective_value(objective, xp, eng, nsteps, dt) -
                          AD.objective_value(objective, xm, eng, nsteps, dt)) / (2 * h))
            fd = (4 * d[0] - d[1]) / 3
            print(i, a, strict, g[i, a], fd, d[0], f"{time.perf_counter() - t0:.1f} s",
                  flush=True)
            floor = 0.0 if strict else 1e-2 * scale
            assert abs(g[i, a] - fd) <= 1e-6 * max(abs(fd), floor)
    finally:
        eng.close()
        eng32.close()
