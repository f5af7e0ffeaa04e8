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
