"""The benchmark's configurations at their workloads, on one GPU.

* The 1M-point tangle (periodic cube L = 0.2 cm, delta = 5e-4 cm, seeded
  loops) run as the reference's main loop runs it: 100 steps of dt = 1e-6 s,
  FP32 far field, with the host topology step (pinsert, neighbour search,
  precon, premove) every step and the wrap.  Step 1 in FP64 against the
  oracle's velocity and AB update followed by the same host line
  operations with the oracle's neighbour search: identical topology
  (links, slot count, counters), positions within 1e-15 cm.  Over the 100
  steps: the loops stay consistent (infront(behind(i)) = i), positions and
  velocities finite, the counters never decrease.
* The ~100k-point periodic tangle at L = 0.1 cm, theta = 0.4: FP64 within
  1e-12 of the oracle at every point, FP32C within 1e-5 of FP64, equal
  interaction counts.
* The ensemble of 256 rings (points U{200..5000}, open box), 500 AB steps
  from captured HIP graphs: FP64 per system within 1e-12 of the
  single-system path (a sample of systems), FP32C velocity within 1e-5 of FP64, graph replay
  equal to eager steps, every state finite after 500 steps.
"""
import numpy as np
import pytest

from conftest import ROOT  # noqa: F401  (puts the repository on sys.path)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _links_ok(inf, beh):
    act = np.flatnonzero(inf)
    return (np.array_equal(beh[inf[act] - 1], act + 1)
            and np.array_equal(inf[beh[act] - 1], act + 1))


KW1M = dict(velocity="Tree", tree_theta=0.4, boundary="periodic", box_size=0.2, delta=5e-4,
            corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6)


@pytest.fixture(scope="module")
def tangle1m():
    from qvort_amd import geometry as G
    return G.random_loops(0.2, 5e-4, target_points=1_000_000, seed=20261202, spacing=0.8)


def test_tangle_1m_first_step_topology_vs_oracle(tangle1m):
    from oracle import port
    from qvort_amd import gpu, line
    fil = tangle1m
    p = gpu.params(precision="fp64", **KW1M)
    st = line.LineState.from_arrays(capacity=2 * fil.pcount + 64, x=fil.x,
                                    infront=fil.infront, behind=fil.behind)
    cnt = line.new_counts()
    _say("gpu step 1", fil.pcount)
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        c.velocity_only()
        c.ab_step()
        line.topology_step(c, p, st, cnt, search=True)
        c.enforce_periodic()
        x_gpu, _, _, _, _ = c.download()
    # the oracle: velocity, AB update, the step's ghosts, then the same host
    # line operations with the oracle's neighbour search on the same tree
    _say("oracle step 1")
    okw = {k: v for k, v in KW1M.items()}
    op = port.make_params(**okw)
    u, _, _, _ = port.velocity(op, fil, nthreads=16)
    z = np.zeros_like(fil.x)
    x1, u1, u2, u3 = port.ab_update(op, fil.x, fil.infront, u, z, z, z)
    gi, gb, gii, gbb = port.ghostp(op, fil)
    so = line.LineState.from_arrays(capacity=2 * fil.pcount + 64, x=x1, u=u, u1=u1, u2=u2,
                                    u3=u3, ghosti=gi, ghostb=gb, ghostii=gii, ghostbb=gbb,
                                    infront=fil.infront, behind=fil.behind)
    co = line.new_counts()
    line.pinsert(p, so, co)
    n = so.pcount
    cur = type(fil)(so.x.copy(), so.infront.copy(), so.behind.copy(), fil.loop_start,
                    fil.loop_len)
    cl, cd = port.closest(op, fil, cur, so.closest, so.closestd, nthreads=16)
    so.closest[:n] = cl
    so.closestd[:n] = cd
    line.precon(p, so, co)
    line.premove(p, so, co)
    _say("gpu", line.counts_dict(cnt))
    _say("oracle", line.counts_dict(co))
    assert st.pcount == so.pcount
    assert np.array_equal(st.infront, so.infront) and np.array_equal(st.behind, so.behind)
    a, b = line.counts_dict(cnt), line.counts_dict(co)
    for k in ("inserted", "recon_count", "remove_count", "small_loop_remove_count"):
        assert a[k] == b[k], k
    assert a["recon_count"] > 0 and a["remove_count"] > 0
    xo = so.x.copy()
    xo = port.enforce_boundary(op, xo, so.infront, so.behind, so.u1, so.u2)[0]
    act = so.infront != 0
    assert np.max(np.abs(x_gpu[act] - xo[act])) <= 1e-15


def test_tangle_1m_hundred_steps_with_reconnections(tangle1m):
    import time

    from qvort_amd import gpu, line
    fil = tangle1m
    p = gpu.params(precision="fp32c", **KW1M)
    st = line.LineState.from_arrays(capacity=2 * fil.pcount + 64, x=fil.x,
                                    infront=fil.infront, behind=fil.behind)
    cnt = line.new_counts()
    prev = line.counts_dict(cnt)
    t0 = time.perf_counter()
    with gpu.Context(p) as c:
        c.upload_filaments(fil)
        for it in range(1, 101):
            c.velocity_only()
            c.ab_step()
            line.topology_step(c, p, st, cnt, search=True)
            c.enforce_periodic()
            d = line.counts_dict(cnt)
            for k in ("recon_count", "remove_count", "small_loop_remove_count", "inserted"):
                assert d[k] >= prev[k], k
            prev = d
            if it % 20 == 0:
                _say("step", it, "points", int(np.count_nonzero(st.infront)), d,
                     f"{time.perf_counter() - t0:.1f} s")
        x, u, _, _, _ = c.download()
    act = st.infront != 0
    assert _links_ok(st.infront, st.behind)
    assert np.all(np.isfinite(x[act])) and np.all(np.isfinite(u[act]))
    assert np.all(np.abs(x[act]) <= 0.1)
    assert prev["recon_count"] > 0


def test_tangle_100k_fp64_vs_oracle_and_fp32c():
    from oracle import port
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    fil = G.random_loops(0.1, 5e-4, target_points=100_000, seed=20261203, spacing=0.8)
    kw = dict(velocity="Tree", tree_theta=0.4, boundary="periodic", box_size=0.1, delta=5e-4,
              corea=1.3e-8, quant_circ=9.97e-4)
    _say("points", fil.pcount, "loops", len(fil.loop_len))
    out = {}
    for prec in ("fp64", "fp32c"):
        with gpu.Context(gpu.params(precision=prec, **kw)) as c:
            c.upload_filaments(fil)
            out[prec] = c.velocity()
    u_o, _, _, ev_o = port.velocity(port.make_params(**kw), fil, nthreads=16)
    act = fil.infront != 0
    u64, ev64 = out["fp64"][0], out["fp64"][3]
    u32, ev32 = out["fp32c"][0], out["fp32c"][3]
    assert ev64 == ev_o and ev32 == ev_o
    e = np.linalg.norm(u64[act] - u_o[act], axis=1) / np.linalg.norm(u_o[act], axis=1)
    e32 = np.linalg.norm(u32[act] - u64[act], axis=1) / np.linalg.norm(u64[act], axis=1)
    _say("fp64 vs oracle max", e.max(), "fp32c vs fp64 max", e32.max(), "rms",
         np.sqrt(np.mean(e32 ** 2)))
    assert e.max() <= 1e-12
    assert e32.max() <= 1e-5


def test_ensemble_256_rings_500_steps():
    import math

    from qvort_amd import geometry as G
    from qvort_amd import gpu
    systems = G.ring_ensemble(256, seed=20261202)
    counts = np.array([s.pcount for s in systems])
    _say("systems", len(systems), "points", int(counts.sum()), "min/max", counts.min(),
         counts.max())
    spacing = min(float(np.min(np.linalg.norm(s.x[s.infront - 1] - s.x, axis=1)))
                  for s in systems)
    dmin = spacing / 2.0
    dt = 0.5 * abs(dmin ** 2 / (9.97e-4 * math.log(dmin / (math.pi * 1.3e-8))))
    kw = dict(velocity="BS", boundary="open", box_size=0.2, corea=1.3e-8, quant_circ=9.97e-4,
              dt=dt)
    with gpu.Batch(gpu.params(precision="fp64", **kw), systems) as b:
        u64 = b.velocity()
    # FP64 per system equals the single-system path, on a sample of systems
    for s in (0, 17, 101, 255):
        with gpu.Context(gpu.params(precision="fp64", **kw)) as c:
            c.upload_filaments(systems[s])
            u1 = c.velocity()[0]
        act = systems[s].infront != 0
        e = np.linalg.norm(u1[act] - u64[s][act], axis=1) / np.linalg.norm(u64[s][act], axis=1)
        assert e.max() <= 1e-12
    with gpu.Batch(gpu.params(precision="fp32c", **kw), systems) as b:
        b.tune(2)
        u32 = b.velocity()
        a = np.concatenate(u32)
        r = np.concatenate(u64)
        e = np.linalg.norm(a - r, axis=1) / np.linalg.norm(r, axis=1)
        _say("fp32c vs fp64 max", e.max())
        assert e.max() <= 1e-5
        x0, _ = b.download()
    # graph replay equals eager steps
    with gpu.Batch(gpu.params(precision="fp32c", **kw), systems) as be, \
            gpu.Batch(gpu.params(precision="fp32c", **kw), systems) as bg:
        be.set_variant(0)
        bg.set_variant(0)
        be.step(20, graph=False)
        bg.step(20, graph=True)
        xe, _ = be.download()
        xg, _ = bg.download()
        assert all(np.array_equal(p, q) for p, q in zip(xe, xg))
        ms = bg.step(480, graph=True)
        xf, uf = bg.download()
    _say("500 steps, last 480:", f"{ms:.1f} ms")
    assert all(np.all(np.isfinite(v)) for v in xf)
    assert all(np.all(np.isfinite(v)) for v in uf)
    # rings keep their size and move by their self-induced velocity
    for s in (0, 128, 255):
        moved = np.linalg.norm(xf[s].mean(axis=0) - x0[s].mean(axis=0))
        assert 0 < moved < 0.01
