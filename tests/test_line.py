"""The line operations (qv_line_pinsert / precon / premove, host code in
the library) against the reference build, operation by operation.

The fixtures (tests/golden/line_*.npz) hold the f(:) fields the operations
touch after each operation of a reference step (pmotion, pinsert, the
neighbour search, precon, premove): each operation here starts from the
reference's state before it and must give the reference's state after it,
bit for bit, and the module counters after the step.  Slots pinsert appends
have undefined closest/closestd in the reference until the search sets
them; those are not compared before the search.
"""
import numpy as np
import pytest

from conftest import line_names, load_golden

NAMES = line_names()
VEC = ("x", "u", "u1", "u2", "ghosti", "ghostb", "ghostii", "ghostbb")


def _params(p):
    from qvort_amd import gpu
    kw = {k: v for k, v in p.items() if k not in ("recon_shots", "runge_kutta")}
    return gpu.params(**kw)


def _stage(g, st):
    keys = [k[len(st) + 2:] for k in g if k.startswith(st + "__")]
    return {k: g[f"{st}__{k}"] for k in keys}


def _state(d, capacity):
    from qvort_amd import line
    arrs = {k: d[k] for k in VEC + ("infront", "behind", "closest", "closestd")}
    return line.LineState.from_arrays(capacity=capacity, **arrs)


def _compare(st, ref, n_defined=None):
    assert st.pcount == int(ref["pcount"])
    for k in VEC:
        assert np.array_equal(getattr(st, k), ref[k]), k
    assert np.array_equal(st.infront, ref["infront"])
    assert np.array_equal(st.behind, ref["behind"])
    m = st.pcount if n_defined is None else n_defined
    assert np.array_equal(st.closest[:m], ref["closest"][:m])
    assert np.array_equal(st.closestd[:m], ref["closestd"][:m])


def _loops_ok(st):
    act = np.flatnonzero(st.infront)
    assert np.array_equal(st.behind[st.infront[act] - 1], act + 1)
    assert np.array_equal(st.infront[st.behind[act] - 1], act + 1)
    seen = np.zeros(st.pcount, dtype=bool)
    for s in act:
        if seen[s]:
            continue
        n, k = 0, s
        while not seen[k]:
            seen[k] = True
            n += 1
            k = st.infront[k] - 1
        assert k == s and n >= 5


@pytest.mark.parametrize("name", NAMES)
def test_pinsert_matches_reference(name):
    from qvort_amd import line
    g = load_golden(name)
    meta = g["meta"]
    p = _params(meta["params"])
    before, after = _stage(g, "motion"), _stage(g, "pinsert")
    n0 = int(before["pcount"])
    st = _state(before, capacity=int(after["pcount"]) + 8)
    c = line.new_counts()
    line.pinsert(p, st, c)
    _compare(st, after, n_defined=n0)
    _loops_ok(st)
    ref = meta["counters1"]
    assert c.total_length == ref["total_length"] and c.avg_sep == ref["avg_sep"]
    assert c.inserted == st.pcount - n0 + int(np.count_nonzero(before["infront"] == 0)) \
        - int(np.count_nonzero(st.infront == 0))


@pytest.mark.parametrize("name", [n for n in NAMES
                                  if "precon" in load_golden(n)["meta"]["stages"]])
def test_precon_and_premove_match_reference(name):
    from qvort_amd import line
    g = load_golden(name)
    meta = g["meta"]
    p = _params(meta["params"])
    c = line.new_counts()
    searched = _stage(g, "search")
    st = _state(searched, capacity=int(searched["pcount"]))
    line.precon(p, st, c)
    _compare(st, _stage(g, "precon"))
    line.premove(p, st, c)
    _compare(st, _stage(g, "premove"))
    _loops_ok(st)
    ref = meta["counters1"]
    assert c.recon_count == ref["recon_count"]
    assert c.remove_count == ref["remove_count"]
    assert c.small_loop_remove_count == ref["small_loop_remove_count"]
    assert c.small_loop_remove_length == ref["small_loop_remove_length"]


def test_scene_exercises_every_operation():
    g = load_golden("line_scene_tree_periodic")
    ref = g["meta"]["counters1"]
    assert ref["recon_count"] >= 1 and ref["remove_count"] >= 1
    assert ref["small_loop_remove_count"] == 1
    before, after = _stage(g, "motion"), _stage(g, "pinsert")
    assert np.count_nonzero(after["infront"]) > np.count_nonzero(before["infront"])
    grow = load_golden("line_sparse_ring_bs_periodic")
    assert int(grow["pinsert__pcount"]) == 2 * int(grow["motion__pcount"])


def test_pinsert_capacity_and_empty_line_errors():
    from qvort_amd import gpu, line
    g = load_golden("line_sparse_ring_bs_periodic")
    p = _params(g["meta"]["params"])
    before = _stage(g, "motion")
    st = _state(before, capacity=int(before["pcount"]) + 3)
    with pytest.raises(gpu.QvortError) as ei:
        line.pinsert(p, st, line.new_counts())
    assert ei.value.status == gpu.QV_ERR_CAPACITY
    empty = line.LineState(4, 8)
    with pytest.raises(gpu.QvortError) as ei:
        line.pinsert(p, empty, line.new_counts())
    assert ei.value.status == gpu.QV_ERR_LINE_EMPTY
    bad = _state(before, capacity=int(before["pcount"]) * 2 + 8)
    bad.infront[3] = bad.pcount + 5
    with pytest.raises(gpu.QvortError) as ei:
        line.premove(p, bad, line.new_counts())
    assert ei.value.status == gpu.QV_ERR_ARG


def _big_state(seed):
    """A 120,000-point state of random loops (periodic cube L = 0.2 cm,
    delta = 5e-4 cm) with ghosts from its links, every 97th segment
    stretched past delta (pinsert inserts), every 89th point moved next to
    the point two ahead (premove removes) and closest / closestd pointing
    every 53rd point at a point moved next to it from elsewhere, closer
    than delta / 2 (precon tests them): above the size from which the tests are computed
    on several host threads."""
    from qvort_amd import geometry as G
    from qvort_amd import line
    fil = G.random_loops(0.2, 5e-4, target_points=120_000, seed=seed, spacing=0.8)
    n = fil.pcount
    x = fil.x.copy()
    inf, beh = fil.infront.copy(), fil.behind.copy()
    act = np.flatnonzero(inf)
    rng = np.random.default_rng(seed)
    for i in act[::97]:
        x[inf[i] - 1] += (x[inf[i] - 1] - x[i]) * 1.5
    for i in act[::89]:
        aa = inf[inf[i] - 1] - 1
        x[i] = x[aa] + 1e-5
    # every 53rd point gets a point of elsewhere moved next to it
    pairs = []
    for i in act[::53]:
        j = int(act[rng.integers(act.size)])
        if j in (i, inf[i] - 1, beh[i] - 1):
            continue
        x[j] = x[i] + np.array([0., 1e-4, 0.])
        pairs.append((i, j))
    L = 0.2

    def ghost(src):
        g = np.zeros_like(x)
        g[act] = x[src[act] - 1]
        d = x - g
        g[d > L / 2] += L
        g[d < -L / 2] -= L
        return g
    gi, gb = ghost(inf), ghost(beh)
    closest = np.zeros(n, np.int32)
    closestd = np.full(n, 100.)
    for i, j in pairs:
        closest[i] = j + 1
        closestd[i] = 1e-4
    return line.LineState.from_arrays(capacity=2 * n, pinned=False, x=x, infront=inf,
                                      behind=beh, ghosti=gi, ghostb=gb, closest=closest,
                                      closestd=closestd)


@pytest.mark.parametrize("threads", ["0", "1"])
def test_threaded_pre_pass_changes_nothing(monkeypatch, threads):
    """The line operations on a state past the threading threshold give the
    same state and counters with the tests computed beforehand on eight host
    threads as with every test evaluated inside the slot loop ("0", the
    plain sequential algorithm) or the pre-pass on one thread."""
    from qvort_amd import gpu, line
    p = gpu.params(velocity="Tree", tree_theta=0.4, boundary="periodic", box_size=0.2,
                   delta=5e-4, corea=1.3e-8, quant_circ=9.97e-4, dt=1e-6)
    runs = []
    for t in (threads, "8"):
        monkeypatch.setenv("QV_HOST_THREADS", t)
        st = _big_state(20261203)
        c = line.new_counts()
        line.pinsert(p, st, c)
        line.precon(p, st, c)
        line.premove(p, st, c)
        runs.append((st, line.counts_dict(c)))
    (a, ca), (b, cb) = runs
    assert ca == cb
    assert ca["inserted"] > 0 and ca["recon_count"] > 0 and ca["remove_count"] > 0
    assert a.pcount == b.pcount
    for k in VEC + ("infront", "behind", "closest", "closestd"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    _loops_ok(a)
