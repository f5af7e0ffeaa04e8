"""ctypes binding of libqvort_gpu.so (C ABI: include/qvort_gpu.h).

The names follow the reference's hot-path procedures: ``upload`` hands over
f(:)%x / infront / behind (cdata.f90:20-41), ``velocity`` runs ghostp +
construct_tree + the velocity loop of pmotion (timestep.f90:41-47),
``ab_step`` the Adams-Bashforth ladder (timestep.f90:48-71), ``tree_dump``
returns the tree node by node in tree_walk's order.  Errors raise
:class:`QvortError` carrying the library's status; the singular-leaf case
carries the reference's fatal_error message (tree.f90:373-375).

There is no fallback: if the library is missing or no GPU is visible, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libqvort_gpu.so")

VEL = {"BS": 1, "Tree": 2}
BC = {"open": 0, "periodic": 1, "openx": 2, "openxy": 3}
PREC = {"fp64": 0, "fp32c": 1}

QV_OK, QV_ERR_ARG, QV_ERR_HIP, QV_ERR_SINGULAR, QV_ERR_TREE_DEPTH, QV_ERR_STATE, \
    QV_ERR_NO_DEVICE, QV_ERR_LINE_EMPTY, QV_ERR_CAPACITY, QV_ERR_SCRATCH = range(10)
QV_PROBE_GHOSTP, QV_PROBE_REFERENCE_SCALE, QV_PROBE_EMPTY_CORE = 1, 2, 4
QV_FLAG_REMOVE_MEAN_POS = 1

EXPORTS = ["qv_create", "qv_destroy", "qv_set_params", "qv_upload", "qv_upload_topology",
           "qv_upload_history",
           "qv_velocity", "qv_ab_step", "qv_rk3_step", "qv_enforce_periodic",
           "qv_boundary_removed", "qv_step", "qv_download", "qv_download_links",
           "qv_tree_dump", "qv_closest", "qv_probe_velocity", "qv_vjp", "qv_jvp", "qv_set_stream", "qv_velocity_device",
           "qv_vjp_device", "qv_jvp_device", "qv_vjp_device_noalloc", "qv_get_timings",
           "qv_batch_create",
           "qv_batch_destroy", "qv_batch_upload", "qv_batch_velocity", "qv_batch_step",
           "qv_batch_download", "qv_batch_tune", "qv_batch_set_variant", "qv_batch_info",
           "qv_download_ghosts", "qv_line_pinsert", "qv_line_precon", "qv_line_premove",
           "qv_host_alloc", "qv_host_free",
           "qv_status_string"]


class QvParams(C.Structure):
    _fields_ = [("quant_circ", C.c_double), ("corea", C.c_double), ("box_size", C.c_double),
                ("delta", C.c_double), ("tree_theta", C.c_double), ("dt", C.c_double),
                ("velocity", C.c_int32), ("boundary", C.c_int32),
                ("tree_extra_correction", C.c_int32), ("deriv_order", C.c_int32),
                ("precision", C.c_int32), ("flags", C.c_int32)]


class QvTimings(C.Structure):
    _fields_ = [("ghost_lia_ms", C.c_float), ("build_ms", C.c_float), ("walk_ms", C.c_float),
                ("total_ms", C.c_float), ("nodes", C.c_int64), ("evals", C.c_int64),
                ("visits", C.c_int64), ("leaf_evals", C.c_int64), ("cell_evals", C.c_int64),
                ("wave_steps", C.c_int64)]


class QvLineState(C.Structure):
    _fields_ = [("pcount", C.c_int32), ("capacity", C.c_int32)] + \
        [(n, C.POINTER(C.c_double)) for n in ("x", "u", "u1", "u2", "u3", "ghosti", "ghostb",
                                               "ghostii", "ghostbb")] + \
        [(n, C.POINTER(C.c_int32)) for n in ("infront", "behind", "closest")] + \
        [("closestd", C.POINTER(C.c_double))]


class QvLineCounts(C.Structure):
    _fields_ = [("inserted", C.c_int64), ("recon_count", C.c_int64),
                ("remove_count", C.c_int64), ("small_loop_remove_count", C.c_int64),
                ("total_length", C.c_double), ("avg_sep", C.c_double),
                ("small_loop_remove_length", C.c_double)]


class QvortError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {status_string(status)} (status {status})")


def params(*, quant_circ=9.97e-4, corea=8.244023e-9, box_size=0.1, delta=5e-4, tree_theta=0.0,
           dt=1e-6, velocity="Tree", boundary="periodic", tree_extra_correction=True,
           deriv_order=2, precision="fp64", remove_mean_pos=False) -> QvParams:
    """run.in parameters of the hot path; defaults are the reference's
    (cdata.f90:196-416) where it has one."""
    if isinstance(deriv_order, str):
        deriv_order = {"second": 2, "fourth": 4}[deriv_order]
    return QvParams(quant_circ, corea, box_size, delta, tree_theta, dt, VEL[velocity],
                    BC[boundary], 1 if tree_extra_correction else 0, int(deriv_order),
                    PREC[precision], QV_FLAG_REMOVE_MEAN_POS if remove_mean_pos else 0)


_lib = None


def _one_hip_runtime() -> None:
    """PyTorch ships its own libamdhip64.so.7; a process that loads that one
    and ROCm's copy holds two HIP runtimes, and the second to initialise does
    not see the device.  Loading and initialising torch (when installed)
    before the library makes the library bind torch's runtime, so the ctypes
    calls and torch tensors share one runtime, in either order of use."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `python -m qvort_amd.build`")
        _one_hip_runtime()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i32, i64 = P(C.c_double), P(C.c_int32), P(C.c_int64)
        vp = C.c_void_p
        L.qv_create.argtypes = [P(QvParams), P(vp)]
        L.qv_destroy.argtypes = [vp]
        L.qv_set_params.argtypes = [vp, P(QvParams)]
        L.qv_upload.argtypes = [vp, C.c_int32, d, i32, i32]
        L.qv_upload_history.argtypes = [vp, d, d, d]
        L.qv_upload_topology.argtypes = [vp, C.c_int32, d, i32, i32]
        L.qv_velocity.argtypes = [vp, d, d, d, i64]
        L.qv_ab_step.argtypes = [vp]
        L.qv_rk3_step.argtypes = [vp]
        L.qv_enforce_periodic.argtypes = [vp]
        L.qv_boundary_removed.argtypes = [vp, i64, d]
        L.qv_step.argtypes = [vp, C.c_int32]
        L.qv_download.argtypes = [vp, d, d, d, d, d]
        L.qv_download_links.argtypes = [vp, i32, i32]
        L.qv_tree_dump.argtypes = [vp, d, C.c_int64, i64]
        L.qv_get_timings.argtypes = [vp, P(QvTimings)]
        L.qv_closest.argtypes = [vp, i32, d]
        L.qv_probe_velocity.argtypes = [vp, C.c_int32, d, d, d, C.c_int32]
        L.qv_vjp.argtypes = [vp, d, d]
        L.qv_jvp.argtypes = [vp, d, d]
        L.qv_set_stream.argtypes = [vp, vp]
        L.qv_velocity_device.argtypes = [vp, vp, vp]
        L.qv_vjp_device.argtypes = [vp, vp, vp, vp]
        L.qv_jvp_device.argtypes = [vp, vp, vp, vp]
        L.qv_vjp_device_noalloc.argtypes = [vp, vp, vp, vp]
        L.qv_batch_create.argtypes = [P(QvParams), C.c_int32, i32, P(vp)]
        L.qv_batch_destroy.argtypes = [vp]
        L.qv_batch_upload.argtypes = [vp, d, i32, i32]
        L.qv_batch_velocity.argtypes = [vp, d]
        L.qv_batch_step.argtypes = [vp, C.c_int32, C.c_int32]
        L.qv_batch_download.argtypes = [vp, d, d]
        L.qv_batch_tune.argtypes = [vp, C.c_int32, i32, P(C.c_float)]
        L.qv_batch_set_variant.argtypes = [vp, C.c_int32]
        L.qv_batch_info.argtypes = [vp, i32, i32, i32, i32, P(C.c_float)]
        L.qv_download_ghosts.argtypes = [vp, d, d, d, d]
        for name in ("qv_line_pinsert", "qv_line_precon", "qv_line_premove"):
            getattr(L, name).argtypes = [P(QvParams), P(QvLineState), P(QvLineCounts)]
        L.qv_status_string.argtypes = [C.c_int]
        L.qv_status_string.restype = C.c_char_p
        for name in EXPORTS[:-1]:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def status_string(status: int) -> str:
    try:
        return lib().qv_status_string(status).decode()
    except FileNotFoundError:
        return "library missing"


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(rc: int, where: str) -> None:
    if rc != QV_OK:
        raise QvortError(rc, where)


class Context:
    """One device context: f(:) state resident in HBM between calls."""

    def __init__(self, p: QvParams):
        self._p = p
        self._h = C.c_void_p()
        _check(lib().qv_create(C.byref(p), C.byref(self._h)), "qv_create")
        self.pcount = 0

    def close(self) -> None:
        if self._h:
            lib().qv_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: QvParams) -> None:
        _check(lib().qv_set_params(self._h, C.byref(p)), "qv_set_params")
        self._p = p

    def upload(self, x, infront, behind) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        inf = np.ascontiguousarray(infront, dtype=np.int32)
        beh = np.ascontiguousarray(behind, dtype=np.int32)
        n = inf.shape[0]
        if x.shape != (n, 3) or beh.shape != (n,):
            raise ValueError("x must be (pcount,3), links (pcount,)")
        _check(lib().qv_upload(self._h, n, _dp(x), _ip(inf), _ip(beh)), "qv_upload")
        self.pcount = n

    def upload_filaments(self, fil) -> None:
        self.upload(fil.x, fil.infront, fil.behind)

    def upload_topology(self, x, infront, behind) -> None:
        """Positions and links after host line operations, keeping the
        device's velocities and AB history (qv_upload_topology).  x=None
        keeps the device's positions (no slot may become a point; cleared
        slots get x = 0 on the device)."""
        inf = np.ascontiguousarray(infront, dtype=np.int32)
        beh = np.ascontiguousarray(behind, dtype=np.int32)
        n = inf.shape[0]
        if x is not None:
            x = np.ascontiguousarray(x, dtype=np.float64)
            if x.shape != (n, 3):
                raise ValueError("x must be (pcount,3)")
        if beh.shape != (n,):
            raise ValueError("links must be (pcount,)")
        _check(lib().qv_upload_topology(self._h, n, None if x is None else _dp(x), _ip(inf),
                                        _ip(beh)),
               "qv_upload_topology")
        self.pcount = n

    def upload_history(self, u1=None, u2=None, u3=None) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (u1, u2, u3)]
        _check(lib().qv_upload_history(self._h, *[_dp(a) for a in arrs]), "qv_upload_history")

    def velocity(self):
        """(u, u_s_LI, u_s_BS, evals) of every slot."""
        n = self.pcount
        u, uli, ubs = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        ev = C.c_int64(0)
        _check(lib().qv_velocity(self._h, _dp(u), _dp(uli), _dp(ubs), C.byref(ev)),
               "qv_velocity")
        return u, uli, ubs, int(ev.value)

    def velocity_only(self) -> None:
        """Velocity stage on the device without copying results back."""
        _check(lib().qv_velocity(self._h, None, None, None, None), "qv_velocity")

    def ab_step(self) -> None:
        _check(lib().qv_ab_step(self._h), "qv_ab_step")

    def rk3_step(self) -> None:
        """pmotion_RK3: three velocity stages with the low-storage update."""
        _check(lib().qv_rk3_step(self._h), "qv_rk3_step")

    def enforce_periodic(self) -> None:
        _check(lib().qv_enforce_periodic(self._h), "qv_enforce_periodic")

    def boundary_removed(self) -> tuple[int, float]:
        """(boundary_loop_remove_count, _length) of this context's boundary steps."""
        n, length = C.c_int64(0), C.c_double(0.0)
        _check(lib().qv_boundary_removed(self._h, C.byref(n), C.byref(length)),
               "qv_boundary_removed")
        return int(n.value), float(length.value)

    def step(self, nsteps: int) -> None:
        _check(lib().qv_step(self._h, nsteps), "qv_step")

    def _outs(self, count: int, out):
        n = self.pcount
        if out is None:
            return [np.zeros((n, 3)) for _ in range(count)]
        if len(out) != count or any(a.shape != (n, 3) or a.dtype != np.float64
                                    or not a.flags.c_contiguous for a in out):
            raise ValueError(f"out must be {count} C-contiguous float64 ({n},3) arrays")
        return list(out)

    def download(self, out=None):
        """(x, u, u1, u2, u3); out: five (pcount,3) float64 arrays to fill."""
        arrs = self._outs(5, out)
        _check(lib().qv_download(self._h, *[_dp(a) for a in arrs]), "qv_download")
        return tuple(arrs)

    def download_links(self, infront=None, behind=None):
        """(infront, behind) of the device state; the given C-contiguous
        int32 (pcount,) arrays are filled in place when passed."""
        n = self.pcount
        arrs = []
        for a in (infront, behind):
            if a is None:
                a = np.zeros(n, dtype=np.int32)
            elif a.dtype != np.int32 or not a.flags.c_contiguous or a.shape[0] < n:
                raise ValueError("links must be C-contiguous int32 arrays of >= pcount entries")
            arrs.append(a)
        _check(lib().qv_download_links(self._h, *[a.ctypes.data_as(C.POINTER(C.c_int32))
                                                  for a in arrs]), "qv_download_links")
        return arrs[0][:n], arrs[1][:n]

    def download_ghosts(self, out=None):
        """(ghosti, ghostb, ghostii, ghostbb) of the last ghostp; out: four
        (pcount,3) float64 arrays to fill."""
        arrs = self._outs(4, out)
        _check(lib().qv_download_ghosts(self._h, *[_dp(a) for a in arrs]), "qv_download_ghosts")
        return tuple(arrs)

    def tree_dump(self) -> np.ndarray:
        nn = C.c_int64(0)
        _check(lib().qv_tree_dump(self._h, None, 0, C.byref(nn)), "qv_tree_dump")
        out = np.zeros((nn.value, 22))
        _check(lib().qv_tree_dump(self._h, _dp(out), nn.value, C.byref(nn)), "qv_tree_dump")
        return out

    def closest(self, closest_in=None, closestd_in=None, inplace: bool = False):
        """Neighbour search of the reconnection step (pclose_tree with
        tree_theta > 0, else pclose) on the current slots: (closest, closestd),
        from the f(:)%closest / %closestd values given (zeros if None).  With
        inplace=True the given arrays (C-contiguous int32 and float64, one
        entry per slot, e.g. pinned host memory) receive the results."""
        n = self.pcount
        if inplace:
            cl, cd = closest_in, closestd_in
            if not (isinstance(cl, np.ndarray) and isinstance(cd, np.ndarray)
                    and cl.dtype == np.int32 and cd.dtype == np.float64
                    and cl.flags.c_contiguous and cd.flags.c_contiguous):
                raise ValueError("inplace needs C-contiguous int32 and float64 arrays")
        else:
            cl = np.zeros(n, dtype=np.int32) if closest_in is None else \
                np.array(closest_in, dtype=np.int32, order="C")
            cd = np.zeros(n) if closestd_in is None else \
                np.array(closestd_in, dtype=np.float64, order="C")
        if cl.shape != (n,) or cd.shape != (n,):
            raise ValueError("closest/closestd must have one entry per slot")
        _check(lib().qv_closest(self._h, _ip(cl), _dp(cd)), "qv_closest")
        return cl, cd

    def probe_velocity(self, y, grad: bool = False, ghostp: bool = False,
                       reference_scale: bool = False, empty_core: bool = False):
        """Superfluid velocity at probe points y (n,3), as mesh_velocity sums
        it; with grad=True also du_k/dy_m as an (n,3,3) array (BS only);
        empty_core=True zeroes both within corea of any slot's position, as
        velocity_quasip does (BS only)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        n = y.shape[0]
        if y.shape != (n, 3):
            raise ValueError("probes must be (n,3)")
        u = np.zeros((n, 3))
        g = np.zeros((n, 3, 3)) if grad else None
        flags = (QV_PROBE_GHOSTP if ghostp else 0) | \
            (QV_PROBE_REFERENCE_SCALE if reference_scale else 0) | \
            (QV_PROBE_EMPTY_CORE if empty_core else 0)
        _check(lib().qv_probe_velocity(self._h, n, _dp(y), _dp(u), _dp(g), flags),
               "qv_probe_velocity")
        return (u, g) if grad else u

    def vjp(self, w) -> np.ndarray:
        """xbar = (dU/dx)^T w at the device positions (direct-sum velocity)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros_like(w)
        _check(lib().qv_vjp(self._h, _dp(w), _dp(out)), "qv_vjp")
        return out

    def jvp(self, v) -> np.ndarray:
        """jv = (dU/dx) v at the device positions (direct-sum velocity)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(v)
        _check(lib().qv_jvp(self._h, _dp(v), _dp(out)), "qv_jvp")
        return out

    def set_stream(self, stream_handle: int | None) -> None:
        _check(lib().qv_set_stream(self._h, stream_handle), "qv_set_stream")

    def velocity_device(self, d_x: int | None, d_u: int) -> None:
        _check(lib().qv_velocity_device(self._h, d_x, d_u), "qv_velocity_device")

    def vjp_device(self, d_x: int | None, d_w: int, d_out: int) -> None:
        _check(lib().qv_vjp_device(self._h, d_x, d_w, d_out), "qv_vjp_device")

    def vjp_device_noalloc(self, d_x: int | None, d_w: int, d_out: int) -> None:
        _check(lib().qv_vjp_device_noalloc(self._h, d_x, d_w, d_out), "qv_vjp_device_noalloc")

    def jvp_device(self, d_x: int | None, d_v: int, d_out: int) -> None:
        _check(lib().qv_jvp_device(self._h, d_x, d_v, d_out), "qv_jvp_device")

    def timings(self) -> dict:
        t = QvTimings()
        _check(lib().qv_get_timings(self._h, C.byref(t)), "qv_get_timings")
        return {k: getattr(t, k) for k, _ in QvTimings._fields_}


class Batch:
    """An ensemble of independent filament systems (velocity='BS') advanced
    together, one persistent launch per AB step (qv_batch_* entry points).
    ``systems`` are filament records (x, infront, behind with 1-based links
    within each system)."""

    def __init__(self, p: QvParams, systems):
        self._h = C.c_void_p()
        self.counts = np.array([s.x.shape[0] for s in systems], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        rc = lib().qv_batch_create(C.byref(p), len(systems), _ip(self.counts),
                                   C.byref(self._h))
        if rc != QV_OK:
            self.close()
            raise QvortError(rc, "qv_batch_create")
        self.upload(systems)

    @property
    def ntot(self) -> int:
        return int(self.offsets[-1])

    def close(self) -> None:
        if self._h:
            lib().qv_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, systems) -> None:
        x = np.ascontiguousarray(np.concatenate([s.x for s in systems]), dtype=np.float64)
        inf = np.ascontiguousarray(np.concatenate([s.infront for s in systems]), dtype=np.int32)
        beh = np.ascontiguousarray(np.concatenate([s.behind for s in systems]), dtype=np.int32)
        if x.shape != (self.ntot, 3) or inf.shape != (self.ntot,) or beh.shape != (self.ntot,):
            raise ValueError("systems do not match the counts of this ensemble")
        _check(lib().qv_batch_upload(self._h, _dp(x), _ip(inf), _ip(beh)), "qv_batch_upload")

    def split(self, a: np.ndarray) -> list:
        return [a[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.counts))]

    def velocity(self) -> list:
        u = np.zeros((self.ntot, 3))
        _check(lib().qv_batch_velocity(self._h, _dp(u)), "qv_batch_velocity")
        return self.split(u)

    def step(self, nsteps: int, graph: bool = True) -> float:
        """nsteps AB steps of every system; returns the device time in ms."""
        _check(lib().qv_batch_step(self._h, nsteps, 1 if graph else 0), "qv_batch_step")
        return self.info()["last_ms"]

    def download(self):
        """(positions, last velocity), each a list of (n_s,3) arrays."""
        x, u = np.zeros((self.ntot, 3)), np.zeros((self.ntot, 3))
        _check(lib().qv_batch_download(self._h, _dp(x), _dp(u)), "qv_batch_download")
        return self.split(x), self.split(u)

    def tune(self, trial_steps: int = 3):
        """Time every tile shape and keep the fastest: (chosen, ms per shape)."""
        nv = self.info()["nvariants"]
        ms = (C.c_float * nv)()
        ch = C.c_int32(-1)
        _check(lib().qv_batch_tune(self._h, trial_steps, C.byref(ch), ms), "qv_batch_tune")
        return int(ch.value), [float(v) for v in ms]

    def set_variant(self, v: int) -> None:
        _check(lib().qv_batch_set_variant(self._h, v), "qv_batch_set_variant")

    def info(self) -> dict:
        nv, v, tile, nu = (C.c_int32() for _ in range(4))
        ms = C.c_float()
        _check(lib().qv_batch_info(self._h, C.byref(nv), C.byref(v), C.byref(tile), C.byref(nu),
                                   C.byref(ms)), "qv_batch_info")
        return {"nvariants": nv.value, "variant": v.value, "tile": tile.value,
                "nunits": nu.value, "last_ms": ms.value}
