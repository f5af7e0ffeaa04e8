"""A run with line operations on the GPU path: velocity and AB update on the
device, pinsert / precon / premove on the host, the neighbour search on the
device tree (qvort_amd.line.topology_step), against the reference build's
run of the same steps (tests/golden/line_*.npz): links, slot count and
counters exactly, positions to 1e-15 cm, AB history to 1e-12."""
import numpy as np
import pytest

from conftest import line_names, load_golden

pytestmark = pytest.mark.gpu

NAMES = line_names()


@pytest.mark.parametrize("name", NAMES)
def test_run_with_line_operations(name):
    from qvort_amd import gpu, line
    g = load_golden(name)
    meta = g["meta"]
    pr = dict(meta["params"])
    recon_shots = pr.pop("recon_shots")
    p = gpu.params(**pr)
    print("case", name, flush=True)
    st = line.LineState.from_arrays(capacity=2 * meta["pcount"] + 8, x=g["x0"],
                                    infront=g["infront"], behind=g["behind"])
    c = line.new_counts()
    with gpu.Context(p) as ctx:
        ctx.upload(g["x0"], g["infront"], g["behind"])
        for itime in range(1, meta["nsteps"] + 1):
            ctx.velocity_only()
            ctx.ab_step()
            line.topology_step(ctx, p, st, c, search=itime % recon_shots == 0)
            ctx.enforce_periodic()
        x, _, u1, u2, _ = ctx.download()
    n = int(g["final__pcount"])
    assert st.pcount == n and ctx.pcount == n
    assert np.array_equal(st.infront, g["final__infront"])
    assert np.array_equal(st.behind, g["final__behind"])
    act = g["final__infront"] != 0
    assert np.max(np.abs(x - g["final__x"])[act]) <= 1e-15
    for k, ref in ((1, g["final__u1"]), (2, g["final__u2"])):
        got = (u1 if k == 1 else u2)[act]
        nr = np.linalg.norm(ref[act], axis=1)
        m = nr > 0
        assert np.max(np.linalg.norm(got - ref[act], axis=1)[m] / nr[m]) <= 1e-12
    ref = meta["counters_final"]
    assert c.recon_count == ref["recon_count"]
    assert c.remove_count == ref["remove_count"]
    assert c.small_loop_remove_count == ref["small_loop_remove_count"]
    assert abs(c.total_length - ref["total_length"]) <= 1e-12 * ref["total_length"]


@pytest.mark.parametrize("name", NAMES)
def test_fortran_driver_with_line_operations(name, tmp_path):
    import os
    import subprocess

    from conftest import ROOT
    from qvort_amd import formats
    g = load_golden(name)
    meta = g["meta"]
    p = meta["params"]
    n, ns = meta["pcount"], meta["nsteps"]
    print("case", name, flush=True)
    lines = [f"nsteps {ns}", f"shots {ns}", f"dt {p['dt']!r}", f"delta {p['delta']!r}",
             f"box_size {p['box_size']!r}", f"velocity {p['velocity']}",
             f"boundary {p['boundary']}", f"quant_circ {p['quant_circ']!r}",
             f"corea {p['corea']!r}", f"tree_theta {p['tree_theta']!r}",
             f"deriv_order {p.get('deriv_order', 'second')}",
             f"recon_shots {p['recon_shots']}", "line_ops T", "initf state_file",
             "gpu_precision fp64"]
    (tmp_path / "run.in").write_text("\n".join(lines) + "\n")
    with open(tmp_path / "state.bin", "wb") as fh:
        np.array([1, n, ns, 0, 0], dtype="<i4").tofile(fh)
        g["x0"].astype("<f8").tofile(fh)
        g["infront"].astype("<i4").tofile(fh)
        g["behind"].astype("<i4").tofile(fh)
        np.zeros((3, n, 3)).astype("<f8").tofile(fh)
    r = subprocess.run([os.path.join(ROOT, "qvort_amd", "lib", "qv_run")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    d = formats.read_var_dat(tmp_path / "data" / "var.dat")
    f = d["f"]
    assert d["pcount"] == int(g["final__pcount"])
    assert d["recon_count"] == meta["counters_final"]["recon_count"]
    assert np.array_equal(f["infront"], g["final__infront"])
    assert np.array_equal(f["behind"], g["final__behind"])
    act = g["final__infront"] != 0
    assert np.max(np.abs(f["x"] - g["final__x"])[act]) <= 1e-15
    # u is the step's velocity as the host operations left it (interpolated
    # for inserted points, zero for removed ones): where the AB history was
    # not reset by pinsert, u = u1
    keep = act & np.any(g["final__u1"] != 0, axis=1)
    assert keep.sum() > 0
    ref = g["final__u1"][keep]
    assert np.max(np.linalg.norm(f["u"][keep] - ref, axis=1) / np.linalg.norm(ref, axis=1)) <= 1e-12
    assert not np.any(f["u"][~act])


def test_pinned_host_memory_round_trip():
    """qv_host_alloc gives page-locked memory that qv_upload / qv_download use
    as any host array; qv_host_free releases it (NULL is ignored)."""
    import ctypes as C
    from qvort_amd import geometry as G
    from qvort_amd import gpu
    L = gpu.lib()
    L.qv_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    L.qv_host_alloc.restype = C.c_int
    L.qv_host_free.argtypes = [C.c_void_p]
    L.qv_host_free.restype = C.c_int
    fil = G.random_loops(0.1, 5e-4, n_loops=3, seed=61)
    n = fil.pcount
    ptr = C.c_void_p()
    assert L.qv_host_alloc(24 * n, C.byref(ptr)) == 0 and ptr.value
    try:
        buf = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(n, 3))
        buf[:] = fil.x
        kw = dict(velocity="BS", tree_theta=0.0, boundary="periodic", box_size=0.1, corea=1.3e-8)
        with gpu.Context(gpu.params(**kw)) as c:
            c.upload(buf, fil.infront, fil.behind)
            buf[:] = 0.
            x = c.download(out=[buf] + [np.zeros((n, 3)) for _ in range(4)])[0]
        assert np.array_equal(buf, fil.x) and np.array_equal(x, fil.x)
    finally:
        assert L.qv_host_free(ptr) == 0
    assert L.qv_host_free(None) == 0


@pytest.mark.parametrize("name", NAMES)
def test_topology_upload_keeps_device_history(name):
    """qv_upload_topology after pinsert and after premove: the device's
    velocities and AB history equal the reference's after those operations
    (new points: u the mean of their neighbours' velocities in pinsert's
    order, u1 = u2 = 0; cleared slots: zeros; the rest untouched) with the
    host's copies of them never sent."""
    from qvort_amd import gpu, line
    g = load_golden(name)
    meta = g["meta"]
    pr = dict(meta["params"])
    pr.pop("recon_shots")
    p = gpu.params(**pr)
    print("case", name, flush=True)
    st = line.LineState.from_arrays(capacity=2 * meta["pcount"] + 8, x=g["x0"],
                                    infront=g["infront"], behind=g["behind"])
    c = line.new_counts()

    def check(ctx, stage):
        n = int(g[f"{stage}__pcount"])
        _, u, u1, u2, _ = ctx.download()
        assert ctx.pcount == n
        act = g[f"{stage}__infront"] != 0
        for got, key in ((u, "u"), (u1, "u1"), (u2, "u2")):
            ref = g[f"{stage}__{key}"]
            assert not np.any(got[~act]), (stage, key)
            nr = np.linalg.norm(ref[act], axis=1)
            z = nr == 0
            assert not np.any(got[act][z]), (stage, key)
            e = np.linalg.norm(got[act][~z] - ref[act][~z], axis=1) / nr[~z]
            assert e.size == 0 or e.max() <= 1e-12, (stage, key, e.max())

    with gpu.Context(p) as ctx:
        ctx.upload(g["x0"], g["infront"], g["behind"])
        ctx.velocity_only()
        ctx.ab_step()
        n = ctx.pcount
        st.pcount = n
        gpu._check(gpu.lib().qv_download(ctx._h, gpu._dp(st._x[:n]), None, None, None, None),
                   "qv_download")
        ctx.download_ghosts(out=[getattr(st, "_" + k)[:n]
                                 for k in ("ghosti", "ghostb", "ghostii", "ghostbb")])
        line.pinsert(p, st, c)
        ctx.upload_topology(st.x, st.infront, st.behind)
        check(ctx, "pinsert")
        if "premove__pcount" not in g:
            return  # recon_shots does not divide the first step
        ctx.closest(st.closest, st.closestd, inplace=True)
        line.precon(p, st, c)
        line.premove(p, st, c)
        # precon and premove only relink and clear: the device keeps its
        # positions and zeroes the cleared slots', which must give the host's
        ctx.upload_topology(None, st.infront, st.behind)
        check(ctx, "premove")
        x, _, _, _, _ = ctx.download()
        assert np.array_equal(x, st.x)
        # a slot that becomes a point needs its position
        if np.any(st.infront == 0):
            inf2, beh2 = st.infront.copy(), st.behind.copy()
            hole = int(np.flatnonzero(inf2 == 0)[0])
            inf2[hole] = beh2[hole] = hole + 1  # a one-point loop
            with pytest.raises(RuntimeError):
                ctx.upload_topology(None, inf2, beh2)
