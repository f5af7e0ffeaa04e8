"""Qvort's on-disk formats, read and written without the Fortran runtime.

* ``data/var.dat`` — the restart file of data_dump / data_restore
  (output.f90:272-285, initial.f90:500-529): Fortran sequential unformatted
  records, each framed by its byte length as a little-endian int32 before and
  after, holding pcount, recon_count, itime, t, the whole f(:) array, then
  quasi_pcount (and the quasi particles if > 0) and part_count (and the
  particles if > 0).  f(:) is written component by component in the order of
  ``type qvort`` (cdata.f90:20-38), with no padding: 460 bytes per point.
* ``data/var%04d.log`` — printf's snapshot (output.f90:116-143): stream
  access, t, pcount, then x(1..3), infront, u(1..3), u_mf(1..3), curv,
  stretch, f_mf(1..3), one array over all slots each.
* ``data/tree_mesh%03d.log`` — construct_tree's leaf print (tree.f90:116-123):
  one list-directed record per cell that is not split (pcount <= 1), in the
  order the cells are created: posx, posx+width, posy, posy+width, posz,
  posz+width, pcount (the Fortran runtime may wrap a record over lines).
* ``data/mesh%03d.dat`` — print_mesh's velocity grid (output.f90:305-321):
  stream access, t, the x coordinates of the n grid planes, then u_norm(1),
  u_norm(2), u_norm(3), u_sup(1), u_sup(2), u_sup(3), each over the n**3
  points of mesh(k,j,i) in Fortran order (k fastest), where mesh(k,j,i) sits
  at (x_i, y_j, z_k) (setup_mesh, initial.f90:550-571).

Records longer than 2**31 - 1 bytes (f(:) of more than 4.6M points) need the
compiler's subrecord convention and are refused.
"""
from __future__ import annotations

import struct

import numpy as np

#: type qvort (cdata.f90:20-38): name, numpy type, count
QVORT_FIELDS = [
    ("x", "<f8", 3), ("x_interp", "<f8", 3), ("u", "<f8", 3), ("u1", "<f8", 3),
    ("u2", "<f8", 3), ("u3", "<f8", 3), ("acc", "<f8", 3), ("u_sup", "<f8", 3),
    ("u_mf", "<f8", 3), ("u_s_LI", "<f8", 3), ("u_s_BS", "<f8", 3), ("ghosti", "<f8", 3),
    ("ghostb", "<f8", 3), ("ghostii", "<f8", 3), ("ghostbb", "<f8", 3), ("f_mf", "<f8", 3),
    ("stretch", "<f8", 1), ("curv", "<f8", 1), ("infront", "<i4", 1), ("behind", "<i4", 1),
    ("closest", "<i4", 1), ("closestd", "<f8", 1), ("closestd_loop", "<f8", 1),
    ("pinnedi", "<i4", 1), ("pinnedb", "<i4", 1), ("delta", "<f8", 1), ("t_recon", "<f8", 2),
]
QVORT_DTYPE = np.dtype([(n, t) if c == 1 else (n, t, (c,)) for n, t, c in QVORT_FIELDS])
assert QVORT_DTYPE.itemsize == 460

_MAX_RECORD = 2**31 - 1


class FormatError(ValueError):
    pass


def _records(raw: bytes):
    off = 0
    while off < len(raw):
        if off + 4 > len(raw):
            raise FormatError("truncated record marker")
        (n,) = struct.unpack_from("<i", raw, off)
        if n < 0:
            raise FormatError("subrecords (records over 2 GiB) are not supported")
        end = off + 4 + n
        if end + 4 > len(raw) or struct.unpack_from("<i", raw, end)[0] != n:
            raise FormatError(f"record at byte {off}: markers do not match")
        yield raw[off + 4:end]
        off = end + 4


def _record(payload: bytes) -> bytes:
    if len(payload) > _MAX_RECORD:
        raise FormatError("records over 2 GiB are not supported")
    m = struct.pack("<i", len(payload))
    return m + payload + m


def new_records(pcount: int) -> np.ndarray:
    """f(:) of pcount zeroed points (logicals .false.)."""
    return np.zeros(pcount, dtype=QVORT_DTYPE)


def read_var_dat(path) -> dict:
    """data_restore's reads (initial.f90:506-525).  Returns pcount,
    recon_count, itime, t, f (a structured array of QVORT_DTYPE),
    quasi_pcount, part_count and the raw bytes of the quasi-particle and
    particle records when present."""
    with open(path, "rb") as fh:
        return read_var_dat_bytes(fh.read())


def read_var_dat_bytes(raw: bytes) -> dict:
    """read_var_dat on the bytes of a var.dat."""
    recs = list(_records(raw))
    if len(recs) < 7:
        raise FormatError(f"{len(recs)} records, var.dat has at least 7")
    pcount = struct.unpack("<i", recs[0])[0]
    out = {"pcount": pcount, "recon_count": struct.unpack("<i", recs[1])[0],
           "itime": struct.unpack("<i", recs[2])[0], "t": struct.unpack("<d", recs[3])[0]}
    if len(recs[4]) != pcount * QVORT_DTYPE.itemsize:
        raise FormatError(f"f(:) record of {len(recs[4])} bytes for pcount {pcount}")
    out["f"] = np.frombuffer(recs[4], dtype=QVORT_DTYPE).copy()
    k = 5
    out["quasi_pcount"] = struct.unpack("<i", recs[k])[0]
    k += 1
    if out["quasi_pcount"] > 0:
        out["quasi_raw"] = recs[k]
        k += 1
    out["part_count"] = struct.unpack("<i", recs[k])[0]
    k += 1
    if out["part_count"] > 0:
        out["part_raw"] = recs[k]
        k += 1
    if k != len(recs):
        raise FormatError("trailing records")
    return out


def write_var_dat(path, f: np.ndarray, *, itime: int, t: float, recon_count: int = 0) -> None:
    """data_dump's writes (output.f90:274-283) with no quasi particles and no
    particles."""
    f = np.ascontiguousarray(f, dtype=QVORT_DTYPE)
    parts = [struct.pack("<i", f.shape[0]), struct.pack("<i", recon_count),
             struct.pack("<i", itime), struct.pack("<d", t), f.tobytes(),
             struct.pack("<i", 0), struct.pack("<i", 0)]
    with open(path, "wb") as fh:
        for p in parts:
            fh.write(_record(p))


def records_from_state(x, infront, behind, u=None, u1=None, u2=None, u3=None) -> np.ndarray:
    """f(:) holding a filament state; the other components are zero."""
    x = np.asarray(x, dtype=np.float64)
    f = new_records(x.shape[0])
    f["x"] = x
    f["infront"] = infront
    f["behind"] = behind
    for name, a in (("u", u), ("u1", u1), ("u2", u2), ("u3", u3)):
        if a is not None:
            f[name] = a
    return f


VAR_LOG_FIELDS = ["x1", "x2", "x3", "infront", "u1", "u2", "u3", "u_mf1", "u_mf2", "u_mf3",
                  "curv", "stretch", "f_mf1", "f_mf2", "f_mf3"]


def read_var_log(path) -> dict:
    """printf's snapshot (output.f90:122-139): t, pcount and 15 slot arrays;
    returns t, pcount, x (pcount, 3), infront, u (pcount, 3), u_mf, curv,
    stretch, f_mf."""
    with open(path, "rb") as fh:
        raw = fh.read()
    t = struct.unpack_from("<d", raw, 0)[0]
    n = struct.unpack_from("<i", raw, 8)[0]
    if len(raw) != 12 + n * (14 * 8 + 4):
        raise FormatError(f"{len(raw)} bytes for pcount {n}")
    off = 12
    cols = {}
    for name in VAR_LOG_FIELDS:
        dt = "<i4" if name == "infront" else "<f8"
        cols[name] = np.frombuffer(raw, dtype=dt, count=n, offset=off).copy()
        off += cols[name].nbytes
    return {"t": t, "pcount": n,
            "x": np.stack([cols["x1"], cols["x2"], cols["x3"]], axis=1),
            "infront": cols["infront"],
            "u": np.stack([cols["u1"], cols["u2"], cols["u3"]], axis=1),
            "u_mf": np.stack([cols["u_mf1"], cols["u_mf2"], cols["u_mf3"]], axis=1),
            "curv": cols["curv"], "stretch": cols["stretch"],
            "f_mf": np.stack([cols["f_mf1"], cols["f_mf2"], cols["f_mf3"]], axis=1)}


def write_var_log(path, t: float, x, infront, u, u_mf=None, curv=None, stretch=None,
                  f_mf=None) -> None:
    x = np.asarray(x, dtype="<f8")
    n = x.shape[0]
    z3 = np.zeros((n, 3))
    z = np.zeros(n)
    u = np.asarray(u, dtype="<f8")
    u_mf = z3 if u_mf is None else np.asarray(u_mf, dtype="<f8")
    f_mf = z3 if f_mf is None else np.asarray(f_mf, dtype="<f8")
    with open(path, "wb") as fh:
        fh.write(struct.pack("<d", t))
        fh.write(struct.pack("<i", n))
        for a in (x[:, 0], x[:, 1], x[:, 2]):
            np.ascontiguousarray(a, dtype="<f8").tofile(fh)
        np.ascontiguousarray(infront, dtype="<i4").tofile(fh)
        for a in (u[:, 0], u[:, 1], u[:, 2], u_mf[:, 0], u_mf[:, 1], u_mf[:, 2],
                  z if curv is None else curv, z if stretch is None else stretch,
                  f_mf[:, 0], f_mf[:, 1], f_mf[:, 2]):
            np.ascontiguousarray(a, dtype="<f8").tofile(fh)


def read_tree_mesh(path_or_text) -> np.ndarray:
    """construct_tree's leaf print as rows of posx, posx+width, posy,
    posy+width, posz, posz+width, pcount (float64)."""
    text = path_or_text
    if "\n" not in str(path_or_text):
        with open(path_or_text) as fh:
            text = fh.read()
    # list-directed output may continue a record on further lines (the
    # runtime wraps long records), so the file is read as a stream of values
    vals = text.split()
    if len(vals) % 7:
        raise FormatError("tree_mesh records have 7 values")
    return np.array([float(v) for v in vals], dtype=np.float64).reshape(-1, 7)


def tree_mesh_rows(nodes: np.ndarray) -> np.ndarray:
    """The leaf print's rows from a node dump (22 values per node in DFS
    order: level, pcount, leaf j, posx, posy, posz, width, ...): the cells
    with pcount <= 1, in DFS order, which is the order construct_tree
    creates (and prints) them."""
    leaf = nodes[nodes[:, 1] <= 1]
    px, py, pz, w = leaf[:, 3], leaf[:, 4], leaf[:, 5], leaf[:, 6]
    return np.stack([px, px + w, py, py + w, pz, pz + w, leaf[:, 1]], axis=1)


def mesh_points(mesh_size: int, box_size: float) -> np.ndarray:
    """setup_mesh's grid (initial.f90:545-571) as (n**3, 3) points in the
    order of print_mesh's arrays: point k + n*j + n*n*i (0-based) is
    mesh(k+1,j+1,i+1) at (c_i, c_j, c_k), c_m = delta*(2m+1)/2 - L/2 with
    delta = L/n, evaluated in that operation order."""
    n = int(mesh_size)
    if n <= 0 or n % 2:
        raise ValueError("mesh size must be a multiple of 2")
    L = float(box_size)
    d = L / n
    c = d * np.arange(1, 2 * n, 2, dtype=np.float64) / 2.0 - (L / 2.0)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # flat index k + n*j + n*n*i: i slowest, k fastest
    return np.stack([c[ii], c[jj], c[kk]], axis=-1).reshape(-1, 3)


def _mesh_size_from_bytes(nbytes: int) -> int:
    # 8 (t) + 8 n (plane coordinates) + 48 n**3 (six fields)
    n = int(round(((nbytes - 8) / 48.0) ** (1.0 / 3.0)))
    for m in (n - 1, n, n + 1):
        if m > 0 and 8 + 8 * m + 48 * m ** 3 == nbytes:
            return m
    raise FormatError(f"{nbytes} bytes is no mesh file")


def read_mesh_dat(path) -> dict:
    """print_mesh's file: t, mesh_size n, x (n plane coordinates), and
    u_norm, u_sup as (n**3, 3) arrays in the point order of mesh_points."""
    with open(path, "rb") as fh:
        raw = fh.read()
    n = _mesh_size_from_bytes(len(raw))
    t = struct.unpack_from("<d", raw, 0)[0]
    x = np.frombuffer(raw, dtype="<f8", count=n, offset=8).copy()
    f = np.frombuffer(raw, dtype="<f8", offset=8 + 8 * n).reshape(6, n ** 3)
    return {"t": t, "mesh_size": n, "x": x, "u_norm": f[0:3].T.copy(),
            "u_sup": f[3:6].T.copy()}


def write_mesh_dat(path, t: float, mesh_size: int, box_size: float, u_sup, u_norm=None) -> None:
    """print_mesh's file for the grid of mesh_points(mesh_size, box_size)."""
    n = int(mesh_size)
    pts = mesh_points(n, box_size)
    u_sup = np.asarray(u_sup, dtype="<f8").reshape(n ** 3, 3)
    u_norm = np.zeros_like(u_sup) if u_norm is None else np.asarray(u_norm, "<f8").reshape(-1, 3)
    # mesh(n/2, n/2, 1:n)%x(1): x_i over i
    xs = pts[(n // 2 - 1) + n * (n // 2 - 1) + n * n * np.arange(n), 0]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<d", t))
        np.ascontiguousarray(xs, dtype="<f8").tofile(fh)
        for a in (u_norm, u_sup):
            for c in range(3):
                np.ascontiguousarray(a[:, c], dtype="<f8").tofile(fh)
