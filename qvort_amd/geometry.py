"""Seeded filament geometries for the workloads this package measures.

Every generator returns a :class:`Filaments` record in Qvort's data model
(cdata.f90:20-41): slot positions ``x`` of shape (pcount, 3), FP64, and the
1-based links ``infront``/``behind`` (0 marks an empty slot).  The reference
builds its initial conditions with Fortran's ``random_number``
(initial_loop.f90:848-983), whose stream differs between compilers, so the
inputs here come from numpy's PCG64 with a stated seed and are handed to the
reference, the oracle and the GPU path as explicit positions.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GAMMA_HE4 = 9.97e-4      # quant_circ default, cdata.f90:202
COREA_DEFAULT = 8.244023e-9  # corea default, cdata.f90:203


@dataclasses.dataclass
class Filaments:
    x: np.ndarray          # (pcount, 3) float64
    infront: np.ndarray    # (pcount,) int32, 1-based, 0 = empty
    behind: np.ndarray     # (pcount,) int32, 1-based, 0 = empty
    loop_start: np.ndarray  # first slot (0-based) of each loop
    loop_len: np.ndarray    # points per loop

    @property
    def pcount(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_active(self) -> int:
        return int(np.count_nonzero(self.infront))

    def copy(self) -> "Filaments":
        return Filaments(self.x.copy(), self.infront.copy(), self.behind.copy(),
                         self.loop_start.copy(), self.loop_len.copy())


def _link_loops(lengths) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    lengths = np.asarray(lengths, dtype=np.int64)
    n = int(lengths.sum())
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
    idx = np.arange(n, dtype=np.int64)
    loop_of = np.repeat(np.arange(len(lengths)), lengths)
    s = starts[loop_of]
    m = lengths[loop_of]
    local = idx - s
    infront = s + (local + 1) % m + 1
    behind = s + (local - 1) % m + 1
    return infront.astype(np.int32), behind.astype(np.int32), starts


def wrap_periodic(x: np.ndarray, box_size: float) -> np.ndarray:
    """Wrap into [-L/2, L/2) per axis (the tree keeps only points inside)."""
    half = box_size / 2.0
    y = np.mod(x + half, box_size) - half
    y[y >= half] -= box_size
    y[y < -half] += box_size
    return y


def single_ring(n: int, radius: float, centre=(0.0, 0.0, 0.0)) -> Filaments:
    """One ring in the x-y plane, points at angles pi(2i-1)/n (the placement
    of setup_single_loop, initial_loop.f90:26-29), circulation such that the
    ring self-propagates along -z."""
    i = np.arange(1, n + 1, dtype=np.float64)
    ang = math.pi * (2.0 * i - 1.0) / n
    x = np.stack([radius * np.sin(ang), radius * np.cos(ang), np.zeros(n)], axis=1)
    x += np.asarray(centre, dtype=np.float64)
    infront, behind, starts = _link_loops([n])
    return Filaments(x, infront, behind, starts, np.array([n]))


def ring_speed(radius: float, gamma: float = GAMMA_HE4, corea: float = COREA_DEFAULT) -> float:
    """Translation speed of a thin ring, Gamma/(4 pi R)(ln(8R/a) - 1/2),
    as printed by setup_single_loop (initial_loop.f90:17)."""
    return gamma / (4.0 * math.pi * radius) * (math.log(8.0 * radius / corea) - 0.5)


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Rotation taking e_z to a uniformly distributed unit normal."""
    nrm = rng.normal(size=3)
    nrm /= np.linalg.norm(nrm)
    helper = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(nrm, helper)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)
    phase = rng.uniform(0.0, 2.0 * math.pi)
    c, s = math.cos(phase), math.sin(phase)
    a = c * e1 + s * e2
    b = -s * e1 + c * e2
    return np.stack([a, b, nrm], axis=1)


def random_loops(box_size: float, delta: float, *, n_loops: int | None = None,
                 target_points: int | None = None, rmin: float = 0.005,
                 rmax: float = 0.02, spacing: float = 0.8, seed: int = 20261202,
                 wrap: bool = True) -> Filaments:
    """Seeded random loops: radius U[rmin, rmax], centre uniform in the cube,
    normal uniform on the sphere, points equally spaced at ``spacing*delta``.
    Either ``n_loops`` or ``target_points`` (loops are added until the point
    count reaches it) must be given."""
    if (n_loops is None) == (target_points is None):
        raise ValueError("give exactly one of n_loops / target_points")
    rng = np.random.Generator(np.random.PCG64(seed))
    pts, lengths = [], []
    total = 0
    k = 0
    while True:
        if n_loops is not None and k >= n_loops:
            break
        if target_points is not None and total >= target_points:
            break
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(-box_size / 2.0, box_size / 2.0, size=3)
        rot = _random_rotation(rng)
        m = max(5, int(math.floor(2.0 * math.pi * r / (spacing * delta))))
        ang = 2.0 * math.pi * np.arange(m) / m
        local = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(m)], axis=1)
        pts.append(local @ rot.T + c)
        lengths.append(m)
        total += m
        k += 1
    x = np.concatenate(pts, axis=0)
    if wrap:
        x = wrap_periodic(x, box_size)
    infront, behind, starts = _link_loops(lengths)
    return Filaments(np.ascontiguousarray(x), infront, behind, starts,
                     np.asarray(lengths, dtype=np.int64))


def kelvin_line(n: int, box_size: float, *, n_waves: int = 16, amp_sigma: float = 2e-4,
                seed: int = 20261202) -> Filaments:
    """Straight periodic line along z through the box, perturbed by a seeded
    sum of Kelvin waves with wavenumbers 2 pi k / L, k = 1..n_waves, and
    amplitudes ~ N(0, amp_sigma^2) in x and y."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = box_size
    z = -L / 2.0 + (np.arange(n) + 0.5) * (L / n)
    ax = rng.normal(0.0, amp_sigma, size=n_waves)
    ay = rng.normal(0.0, amp_sigma, size=n_waves)
    px = rng.uniform(0.0, 2.0 * math.pi, size=n_waves)
    py = rng.uniform(0.0, 2.0 * math.pi, size=n_waves)
    k = 2.0 * math.pi * np.arange(1, n_waves + 1) / L
    xx = (ax[None, :] * np.cos(k[None, :] * z[:, None] + px[None, :])).sum(axis=1)
    yy = (ay[None, :] * np.cos(k[None, :] * z[:, None] + py[None, :])).sum(axis=1)
    x = np.stack([xx, yy, z], axis=1)
    infront, behind, starts = _link_loops([n])
    return Filaments(np.ascontiguousarray(x), infront, behind, starts, np.array([n]))


def ring_ensemble(n_systems: int, *, nmin: int = 200, nmax: int = 5000, rmin: float = 0.02,
                  rmax: float = 0.06, helix_amp: float = 0.01, helix_waves: int = 8,
                  seed: int = 20261202) -> list[Filaments]:
    """Independent rings (one system each) of uneven point counts, each with a
    seeded helical perturbation of relative amplitude ``helix_amp``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n_systems):
        n = int(rng.integers(nmin, nmax + 1))
        r = rng.uniform(rmin, rmax)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        t = 2.0 * math.pi * np.arange(n) / n
        a = helix_amp * r
        rr = r + a * np.cos(helix_waves * t + ph)
        zz = a * np.sin(helix_waves * t + ph)
        x = np.stack([rr * np.cos(t), rr * np.sin(t), zz], axis=1)
        infront, behind, starts = _link_loops([n])
        out.append(Filaments(np.ascontiguousarray(x), infront, behind, starts, np.array([n])))
    return out


def combine(*fils: Filaments) -> Filaments:
    """The filaments of several records in one f(:), in the order given
    (links renumbered; empty slots stay empty)."""
    xs, infs, behs, starts, lens = [], [], [], [], []
    off = 0
    for f in fils:
        xs.append(f.x)
        infs.append(np.where(f.infront != 0, f.infront + off, 0).astype(np.int32))
        behs.append(np.where(f.behind != 0, f.behind + off, 0).astype(np.int32))
        starts.append(f.loop_start + off)
        lens.append(f.loop_len)
        off += f.pcount
    return Filaments(np.concatenate(xs), np.concatenate(infs), np.concatenate(behs),
                     np.concatenate(starts), np.concatenate(lens))


def with_holes(fil: Filaments, n_holes: int, seed: int = 7) -> Filaments:
    """Insert empty slots (infront=behind=0) at seeded positions, renumbering
    the links, as Qvort's f(:) holds after point removals (line.f90:114-161)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = fil.pcount
    total = n + n_holes
    hole_pos = np.sort(rng.choice(total, size=n_holes, replace=False))
    is_hole = np.zeros(total, dtype=bool)
    is_hole[hole_pos] = True
    new_of_old = np.flatnonzero(~is_hole)  # new 0-based slot of old point k
    x = np.zeros((total, 3))
    x[new_of_old] = fil.x
    infront = np.zeros(total, dtype=np.int32)
    behind = np.zeros(total, dtype=np.int32)
    infront[new_of_old] = new_of_old[fil.infront - 1] + 1
    behind[new_of_old] = new_of_old[fil.behind - 1] + 1
    return Filaments(x, infront, behind, new_of_old[fil.loop_start], fil.loop_len.copy())


def reconnection_scene(delta: float = 5e-4, radius: float = 0.01, gap: float = 0.3,
                       seed: int = 81) -> Filaments:
    """A state that exercises every line operation of one step.

    Ring A (x-y plane, centre 0) and ring B (x-z plane) pass within
    gap * delta of each other at (radius, 0, 0), with crossing tangents, so
    the pair reconnects (closestd < delta/2, line length shrinks).  Ring A
    has one empty slot on its far side, a gap of two spacings that pinsert
    fills; ring B has an extra point half a spacing after its far point,
    which premove removes; a 5-point loop of radius delta/2 (its points two
    apart are closer than delta) loses a point to premove and is then removed
    by loop_killer.  Spacing 0.8 delta; positions carry a seeded jitter of
    1e-3 delta."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 2 * int(round(math.pi * radius / (0.8 * delta)))
    th = 2.0 * math.pi * np.arange(n) / n
    ring_a = np.stack([radius * np.cos(th), radius * np.sin(th), np.zeros(n)], axis=1)
    cx = 2.0 * radius + gap * delta
    ring_b = np.stack([cx - radius * np.cos(th), np.zeros(n), -radius * np.sin(th)], axis=1)
    far = n // 2
    ph = 0.5 * (th[far] + th[far + 1])
    extra = np.array([[cx - radius * math.cos(ph), 0.0, -radius * math.sin(ph)]])
    r_small = 0.5 * delta
    ts = 2.0 * math.pi * np.arange(5) / 5
    small = np.stack([r_small * np.cos(ts), r_small * np.sin(ts), np.full(5, 0.03)], axis=1)
    x = np.concatenate([ring_a, ring_b, extra, small])
    x += rng.normal(scale=1e-3 * delta, size=x.shape)
    m = x.shape[0]
    inf = np.zeros(m, dtype=np.int32)
    beh = np.zeros(m, dtype=np.int32)

    def chain(slots):
        for k, s in enumerate(slots):
            inf[s] = slots[(k + 1) % len(slots)] + 1
            beh[s] = slots[k - 1] + 1

    chain([k for k in range(n) if k != far])
    b_slots = list(range(n, 2 * n))
    b_slots.insert(far + 1, 2 * n)
    chain(b_slots)
    chain(list(range(2 * n + 1, 2 * n + 6)))
    x[far] = 0.0
    starts = np.array([0, n, 2 * n + 1])
    lens = np.array([n - 1, n + 1, 5])
    return Filaments(x, inf, beh, starts, lens)
