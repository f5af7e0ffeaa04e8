"""ctypes binding of the C restatement (oracle/qv_oracle.c -> oracle/_ref/liboracle.so).

Test infrastructure only (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "liboracle.so")

VEL_BS, VEL_TREE = 1, 2
BC = {"open": 0, "periodic": 1, "openx": 2, "openxy": 3}
NODE_DOUBLES = 22


class QvoParams(C.Structure):
    _fields_ = [("quant_circ", C.c_double), ("corea", C.c_double), ("box_size", C.c_double),
                ("delta", C.c_double), ("tree_theta", C.c_double), ("dt", C.c_double),
                ("velocity", C.c_int32), ("boundary", C.c_int32),
                ("tree_extra_correction", C.c_int32), ("deriv_order", C.c_int32),
                ("remove_mean_pos", C.c_int32)]


def make_params(*, quant_circ=9.97e-4, corea=8.244023e-9, box_size=0.1, delta=5e-4,
                tree_theta=0.0, dt=1e-6, velocity="Tree", boundary="periodic",
                tree_extra_correction=True, deriv_order=2,
                remove_mean_pos=False) -> QvoParams:
    return QvoParams(quant_circ, corea, box_size, delta, tree_theta, dt,
                     VEL_TREE if velocity == "Tree" else VEL_BS, BC[boundary],
                     1 if tree_extra_correction else 0, int(deriv_order),
                     1 if remove_mean_pos else 0)


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "port"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i32, i64 = P(C.c_double), P(C.c_int32), P(C.c_int64)
        L.qvo_velocity.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, d, d, i64, C.c_int32]
        L.qvo_velocity.restype = C.c_int
        L.qvo_velocity_targets.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, i32, C.c_int32,
                                           d, i64, C.c_int32, d, d]
        L.qvo_velocity_targets.restype = C.c_int
        L.qvo_leaf_sum_targets.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, i32, C.c_int32,
                                           d, i64, C.c_int32]
        L.qvo_leaf_sum_targets.restype = C.c_int
        L.qvo_tree_dump.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, C.c_int64]
        L.qvo_tree_dump.restype = C.c_int64
        L.qvo_ghostp.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, d, d, d]
        L.qvo_ghostp.restype = C.c_int
        L.qvo_ab_update.argtypes = [P(QvoParams), C.c_int32, d, i32, d, d, d, d]
        L.qvo_ab_update.restype = C.c_int
        L.qvo_run.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, d, d, C.c_int32, d, C.c_int32,
                              i64, d]
        L.qvo_run.restype = C.c_int
        L.qvo_segment_velocity.argtypes = [P(QvoParams), d, d, d, d, C.c_int32, d]
        L.qvo_segment_velocity.restype = C.c_int
        L.qvo_enforce_periodic.argtypes = [P(QvoParams), C.c_int32, d, i32]
        L.qvo_enforce_periodic.restype = C.c_int
        L.qvo_enforce_boundary.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, d, i64, d]
        L.qvo_enforce_boundary.restype = C.c_int
        L.qvo_closest_tree.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, C.c_int32, d, i32,
                                       i32, i32, d, C.c_int32]
        L.qvo_closest_tree.restype = C.c_int
        L.qvo_closest_direct.argtypes = [C.c_int32, d, i32, i32, i32, d, C.c_int32]
        L.qvo_closest_direct.restype = C.c_int
        L.qvo_probe_velocity.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, C.c_int32, d, i32,
                                         d, C.c_int32, d, C.c_int32, d, C.c_int32]
        L.qvo_probe_velocity.restype = C.c_int
        L.qvo_run_rk3.argtypes = [P(QvoParams), C.c_int32, d, i32, i32, d, C.c_int32, d,
                                  C.c_int32]
        L.qvo_run_rk3.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _prep(fil):
    x = np.ascontiguousarray(fil.x, dtype=np.float64)
    inf = np.ascontiguousarray(fil.infront, dtype=np.int32)
    beh = np.ascontiguousarray(fil.behind, dtype=np.int32)
    return x, inf, beh


def velocity(params: QvoParams, fil, nthreads: int = 0):
    """(u, u_li, u_bs, evals) for every slot, as calc_velocity computes them."""
    x, inf, beh = _prep(fil)
    n = x.shape[0]
    u = np.zeros((n, 3)); uli = np.zeros((n, 3)); ubs = np.zeros((n, 3))
    ev = C.c_int64(0)
    rc = lib().qvo_velocity(C.byref(params), n, _d(x), _i(inf), _i(beh), _d(u), _d(uli), _d(ubs),
                            C.byref(ev), nthreads)
    if rc != 0:
        raise RuntimeError(f"qvo_velocity failed ({rc})")
    return u, uli, ubs, int(ev.value)


def velocity_targets(params: QvoParams, fil, targets, nthreads: int = 0):
    x, inf, beh = _prep(fil)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    u = np.zeros((t.shape[0], 3))
    ev = C.c_int64(0)
    bs, ws = C.c_double(0), C.c_double(0)
    rc = lib().qvo_velocity_targets(C.byref(params), x.shape[0], _d(x), _i(inf), _i(beh), _i(t),
                                    t.shape[0], _d(u), C.byref(ev), nthreads, C.byref(bs),
                                    C.byref(ws))
    if rc != 0:
        raise RuntimeError(f"qvo_velocity_targets failed ({rc})")
    return u, int(ev.value), bs.value, ws.value


def leaf_sum_targets(params: QvoParams, fil, targets, nthreads: int = 0):
    """(u, evals): the theta -> 0+ limit of the Tree velocity (every leaf
    with tree_walk's exclusions, summed in slot order) for the listed
    targets -- the reference of the multipole truncation error."""
    x, inf, beh = _prep(fil)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    u = np.zeros((t.shape[0], 3))
    ev = C.c_int64(0)
    rc = lib().qvo_leaf_sum_targets(C.byref(params), x.shape[0], _d(x), _i(inf), _i(beh), _i(t),
                                    t.shape[0], _d(u), C.byref(ev), nthreads)
    if rc != 0:
        raise RuntimeError(f"qvo_leaf_sum_targets failed ({rc})")
    return u, int(ev.value)


def tree_dump(params: QvoParams, fil) -> np.ndarray:
    x, inf, beh = _prep(fil)
    n = lib().qvo_tree_dump(C.byref(params), x.shape[0], _d(x), _i(inf), _i(beh), None, 0)
    if n < 0:
        raise RuntimeError("qvo_tree_dump failed")
    out = np.zeros((n, NODE_DOUBLES))
    lib().qvo_tree_dump(C.byref(params), x.shape[0], _d(x), _i(inf), _i(beh), _d(out), n)
    return out


def ghostp(params: QvoParams, fil):
    x, inf, beh = _prep(fil)
    n = x.shape[0]
    g = [np.zeros((n, 3)) for _ in range(4)]
    lib().qvo_ghostp(C.byref(params), n, _d(x), _i(inf), _i(beh), *[_d(a) for a in g])
    return g


def ab_update(params: QvoParams, x, infront, u, u1, u2, u3):
    """In-place AB ladder on copies; returns (x, u1, u2, u3)."""
    x = np.array(x, dtype=np.float64, order="C"); u1 = np.array(u1, dtype=np.float64, order="C")
    u2 = np.array(u2, dtype=np.float64, order="C"); u3 = np.array(u3, dtype=np.float64, order="C")
    inf = np.ascontiguousarray(infront, dtype=np.int32)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    lib().qvo_ab_update(C.byref(params), x.shape[0], _d(x), _i(inf), _d(uu), _d(u1), _d(u2), _d(u3))
    return x, u1, u2, u3


def run(params: QvoParams, fil, nsteps: int, u1=None, u2=None, u3=None, nthreads: int = 0,
        topology: bool = False):
    """nsteps of velocity, AB ladder (and remove_mean_pos) and the boundary
    step; returns (x, u1, u2, u3, u of the first step), with topology=True
    also the links after the run and boundary_loop_remove's count and
    length."""
    x, inf, beh = _prep(fil)
    x, inf, beh = x.copy(), inf.copy(), beh.copy()
    n = x.shape[0]
    u1 = np.zeros((n, 3)) if u1 is None else np.array(u1, dtype=np.float64, order="C")
    u2 = np.zeros((n, 3)) if u2 is None else np.array(u2, dtype=np.float64, order="C")
    u3 = np.zeros((n, 3)) if u3 is None else np.array(u3, dtype=np.float64, order="C")
    uf = np.zeros((n, 3))
    cnt, length = C.c_int64(0), C.c_double(0.0)
    rc = lib().qvo_run(C.byref(params), n, _d(x), _i(inf), _i(beh), _d(u1), _d(u2), _d(u3),
                       nsteps, _d(uf), nthreads, C.byref(cnt), C.byref(length))
    if rc != 0:
        raise RuntimeError(f"qvo_run failed ({rc})")
    if topology:
        return x, u1, u2, u3, uf, inf, beh, int(cnt.value), float(length.value)
    return x, u1, u2, u3, uf


def enforce_boundary(params: QvoParams, x, infront, behind, u1, u2):
    """The boundary step on copies: (x, infront, behind, u1, u2, count, length)."""
    x = np.array(x, dtype=np.float64, order="C")
    inf = np.array(infront, dtype=np.int32, order="C")
    beh = np.array(behind, dtype=np.int32, order="C")
    u1 = np.array(u1, dtype=np.float64, order="C")
    u2 = np.array(u2, dtype=np.float64, order="C")
    cnt, length = C.c_int64(0), C.c_double(0.0)
    lib().qvo_enforce_boundary(C.byref(params), x.shape[0], _d(x), _i(inf), _i(beh), _d(u1),
                               _d(u2), C.byref(cnt), C.byref(length))
    return x, inf, beh, u1, u2, int(cnt.value), float(length.value)


def segment_velocity(params: QvoParams, xi, xj, gj, shift=(0.0, 0.0, 0.0), form: int = 0):
    """Velocity at xi of the straight segment xj -> gj shifted by `shift`."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xi, xj, gj, shift)]
    u = np.zeros(3)
    lib().qvo_segment_velocity(C.byref(params), *[_d(v) for v in a], form, _d(u))
    return u


def closest(params: QvoParams, build, current, closest_in, closestd_in, nthreads: int = 0):
    """Neighbour search of the reconnection step: pclose_tree over the tree
    of ``build`` (tree_theta > 0) or pclose (tree_theta = 0) with the
    positions and links of ``current``; returns (closest, closestd)."""
    x, inf, beh = _prep(current)
    n = x.shape[0]
    cl = np.array(closest_in, dtype=np.int32, order="C").copy()
    cd = np.array(closestd_in, dtype=np.float64, order="C").copy()
    assert cl.shape == (n,) and cd.shape == (n,)
    if params.tree_theta > 0:
        xb, ib, bb = _prep(build)
        rc = lib().qvo_closest_tree(C.byref(params), xb.shape[0], _d(xb), _i(ib), _i(bb), n,
                                    _d(x), _i(inf), _i(beh), _i(cl), _d(cd), nthreads)
    else:
        rc = lib().qvo_closest_direct(n, _d(x), _i(inf), _i(beh), _i(cl), _d(cd), nthreads)
    if rc != 0:
        raise RuntimeError(f"closest search failed ({rc})")
    return cl, cd


def probe_velocity(params: QvoParams, build, current, ghosti, probes, ref_scaling: bool = True,
                   nthreads: int = 0) -> np.ndarray:
    """u at probe points as mesh_velocity sums it: biot_savart_general(_shift)
    over ``current`` with the ghosts ``ghosti`` (BS), or tree_walk_general over
    the tree of ``build`` (Tree; ref_scaling keeps the reference's factors)."""
    x, inf, _ = _prep(current)
    xb, ib, bb = _prep(build)
    gi = np.ascontiguousarray(ghosti, dtype=np.float64)
    y = np.ascontiguousarray(probes, dtype=np.float64)
    u = np.zeros_like(y)
    rc = lib().qvo_probe_velocity(C.byref(params), xb.shape[0], _d(xb), _i(ib), _i(bb),
                                  x.shape[0], _d(x), _i(inf), _d(gi), y.shape[0], _d(y),
                                  1 if ref_scaling else 0, _d(u), nthreads)
    if rc != 0:
        raise RuntimeError(f"probe velocity failed ({rc})")
    return u


def run_rk3(params: QvoParams, fil, nsteps: int, u1=None, nthreads: int = 0):
    """nsteps of pmotion_RK3 + the boundary step; returns (x, u1, u_last)."""
    x, inf, beh = _prep(fil)
    x, inf, beh = x.copy(), inf.copy(), beh.copy()
    n = x.shape[0]
    u1 = np.zeros((n, 3)) if u1 is None else np.array(u1, dtype=np.float64, order="C")
    ul = np.zeros((n, 3))
    rc = lib().qvo_run_rk3(C.byref(params), n, _d(x), _i(inf), _i(beh), _d(u1), nsteps,
                           _d(ul), nthreads)
    if rc != 0:
        raise RuntimeError(f"qvo_run_rk3 failed ({rc})")
    return x, u1, ul
