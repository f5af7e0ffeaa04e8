"""The line operations of a step (run.f90:60-72) on numpy arrays, through
the library's host entry points qv_line_pinsert / qv_line_precon /
qv_line_premove (include/qvort_gpu.h).

:class:`LineState` holds the f(:) fields they read and write, with spare
slots for the points pinsert appends; :func:`topology_step` runs a step's
line operations in the reference's order with the neighbour search on the
device tree of the step's velocity stage.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gpu

VEC_FIELDS = ("x", "u", "u1", "u2", "u3", "ghosti", "ghostb", "ghostii", "ghostbb")
INT_FIELDS = ("infront", "behind", "closest")


def _pinned_allocator():
    """numpy arrays in page-locked host memory (torch's pinned allocator) when
    a GPU is visible, so that the state's copies to and from the device run
    at DMA speed; None otherwise."""
    try:
        import torch
    except ImportError:
        return None
    if not torch.cuda.is_available():
        return None
    keep = []

    def alloc(shape, dtype):
        t = torch.zeros(shape, dtype={np.float64: torch.float64, np.int32: torch.int32}[dtype],
                        pin_memory=True)
        keep.append(t)
        return t.numpy()
    alloc.keep = keep
    return alloc


class LineState:
    """Host arrays of `capacity` slots, of which the first `pcount` are f(:);
    page-locked when a GPU is visible (pinned=None) or asked for."""

    def __init__(self, pcount: int, capacity: int | None = None, pinned: bool | None = None):
        cap = max(pcount, capacity if capacity is not None else pcount + pcount // 8 + 64)
        self.pcount = int(pcount)
        self.capacity = int(cap)
        alloc = _pinned_allocator() if pinned in (None, True) else None
        if pinned and alloc is None:
            raise RuntimeError("pinned host memory needs torch with a visible GPU")
        if alloc is None:
            def alloc(shape, dtype):
                return np.zeros(shape, dtype=dtype)
        self._alloc = alloc  # keeps the pinned buffers alive
        for n in VEC_FIELDS:
            setattr(self, "_" + n, alloc((cap, 3), np.float64))
        for n in INT_FIELDS:
            setattr(self, "_" + n, alloc((cap,), np.int32))
        self._closestd = alloc((cap,), np.float64)

    @classmethod
    def from_arrays(cls, capacity: int | None = None, pinned: bool | None = None,
                    **arrays) -> "LineState":
        n = np.asarray(arrays["x"]).shape[0]
        st = cls(n, capacity, pinned)
        for name, a in arrays.items():
            if a is not None:
                getattr(st, "_" + name)[:n] = a
        return st

    def __getattr__(self, name):
        if name in VEC_FIELDS or name in INT_FIELDS or name == "closestd":
            return self.__dict__["_" + name][:self.pcount]
        raise AttributeError(name)

    def _struct(self) -> gpu.QvLineState:
        st = gpu.QvLineState()
        st.pcount = self.pcount
        st.capacity = self.capacity
        for n in VEC_FIELDS + ("closestd",):
            setattr(st, n, getattr(self, "_" + n).ctypes.data_as(C.POINTER(C.c_double)))
        for n in INT_FIELDS:
            setattr(st, n, getattr(self, "_" + n).ctypes.data_as(C.POINTER(C.c_int32)))
        return st

    def _call(self, fn: str, p: gpu.QvParams, counts: gpu.QvLineCounts) -> None:
        st = self._struct()
        rc = getattr(gpu.lib(), fn)(C.byref(p), C.byref(st), C.byref(counts))
        self.pcount = st.pcount
        gpu._check(rc, fn)


def new_counts() -> gpu.QvLineCounts:
    return gpu.QvLineCounts()


def counts_dict(c: gpu.QvLineCounts) -> dict:
    return {n: getattr(c, n) for n, _ in gpu.QvLineCounts._fields_}


def pinsert(p: gpu.QvParams, st: LineState, counts: gpu.QvLineCounts) -> None:
    """pinsert (line.f90:18-108)."""
    st._call("qv_line_pinsert", p, counts)


def precon(p: gpu.QvParams, st: LineState, counts: gpu.QvLineCounts) -> None:
    """precon_original (reconnection.f90:145-256) with loop_killer."""
    st._call("qv_line_precon", p, counts)


def premove(p: gpu.QvParams, st: LineState, counts: gpu.QvLineCounts) -> None:
    """premove (line.f90:114-161) with loop_killer."""
    st._call("qv_line_premove", p, counts)


def topology_step(ctx: gpu.Context, p: gpu.QvParams, st: LineState,
                  counts: gpu.QvLineCounts, *, search: bool) -> None:
    """The line operations after a step's velocity and AB update
    (run.f90:60-72): the positions and the step's ghosts come back from the
    device, pinsert runs on the host, and when `search` is set (recon_shots
    divides itime) the device searches the step's tree for every slot's
    closest point before precon and premove run on the host.  Positions and
    links go back with qv_upload_topology after pinsert and after premove,
    and the device gives new and cleared slots their velocities and AB
    history; the caller then runs the boundary step (qv_enforce_periodic) as
    the reference does after its ghostp.  The host owns the topology: `st`
    holds the current links (and closest/closestd) between calls; its
    velocity and history arrays are not kept up to date (the device's are
    the state), nor are its positions after precon and premove (the device
    zeroes the slots they clear); in openx / openxy boxes the links come
    back first, since the
    device's boundary step changes them when a loop leaves through an open
    face."""
    n = ctx.pcount
    if n > st.capacity:
        raise ValueError("the line state holds fewer slots than the device")
    st.pcount = n
    if p.boundary in (gpu.BC["openx"], gpu.BC["openxy"]):
        # the device's boundary step may have removed loops at an open face
        ctx.download_links(st._infront, st._behind)
    # straight into the state's arrays (their first n rows are contiguous):
    # the positions and the ghosts the operations read (ghostii and ghostbb
    # only for fourth-order derivatives); velocities and AB history stay on
    # the device, which applies pinsert's and clear_particle's values to
    # them (qv_upload_topology), so the host's copies are not read
    _dp = gpu._dp
    gpu._check(gpu.lib().qv_download(ctx._h, _dp(st._x[:n]), None, None, None, None),
               "qv_download")
    g = [st._ghosti[:n], st._ghostb[:n]]
    g += [st._ghostii[:n], st._ghostbb[:n]] if p.deriv_order == 4 else [None, None]
    gpu._check(gpu.lib().qv_download_ghosts(ctx._h, *[None if a is None else _dp(a) for a in g]),
               "qv_download_ghosts")
    pinsert(p, st, counts)
    ctx.upload_topology(st.x, st.infront, st.behind)
    if search:
        ctx.closest(st.closest, st.closestd, inplace=True)
        precon(p, st, counts)
        premove(p, st, counts)
        # they only relink and clear slots: the device keeps its positions
        ctx.upload_topology(None, st.infront, st.behind)
