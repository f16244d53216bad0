"""Symmetry descriptors and the orbit tables of symmetry-reduced assembly.

Mirrors the reference's symmetry types and the orbit bookkeeping its reduced
Kress workspaces build (host work, O(N log N)):
  * Reflection / Rotation, XReflection, YReflection, XYReflection
    (src/utils/symmetry.jl:14-47, :97-131) and the characters
    chi_l = exp(2 pi i (m mod n) l / n) of a rotation irrep (_rotation_tables,
    symmetry.jl:64-82);
  * CFIE_kress: node maps matched by coordinates (build_symmetry_maps with
    match_index for reflections, build_rotation_maps_components for rotations,
    symmetry.jl:590-766), orbits [g, image(g), ...] with scales
    [1, parity] / [1, px, py, px py] / chi_l (apply_symmetry_orbit_from_maps,
    symmetry.jl:776-802), representatives in order of first visit
    (build_CFIE_kress_reduced_workspace, cheb-CFIE:475-505);
  * DLP_kress: index orbits on the periodic grid (periodic_symmetry_index_orbits,
    symmetry.jl:214-298): both single reflections use q -> N - q + 1 and
    Ifund = 1..N/2, the origin reflection q -> {q, N/2 - q + 1, N - q + 1,
    q + N/2} with Ifund = 1..N/4, rotations q -> q + l N/n with Ifund = 1..N/n.
  * The node-count rule: N even, divisible by 4 for a reflection and by
    lcm(2, n) for a rotation (cfie.jl:538-549, dlp_kress.jl:363-371).
Indices here are 0-based; the reference's are 1-based.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

TWO_PI = 2.0 * math.pi


@dataclass(frozen=True)
class Reflection:
    """Reflection(sym_map, parity, axis) (symmetry.jl:14-18).  axis is
    "y_axis" (x -> -x), "x_axis" (y -> -y) or "origin" (both; parity is then
    the pair (px, py))."""
    parity: Union[int, Tuple[int, int]]
    axis: str

    @property
    def order(self) -> int:
        return 4 if self.axis == "origin" else 2


@dataclass(frozen=True)
class Rotation:
    """Rotation(n, m; center) (symmetry.jl:40-47): C_n, irrep m mod n."""
    n: int
    m: int = 0
    center: Tuple[float, float] = (0.0, 0.0)

    def __post_init__(self):
        if self.n < 2:
            raise ValueError("rotation order n must be >= 2")
        object.__setattr__(self, "m", self.m % self.n)
        object.__setattr__(self, "center", (float(self.center[0]), float(self.center[1])))

    @property
    def order(self) -> int:
        return self.n


def XReflection(parity: int) -> Reflection:
    """Reflection over the y axis, x -> -x (symmetry.jl:97-99)."""
    return Reflection(int(parity), "y_axis")


def YReflection(parity: int) -> Reflection:
    """Reflection over the x axis, y -> -y (symmetry.jl:112-114)."""
    return Reflection(int(parity), "x_axis")


def XYReflection(parity_x: int, parity_y: int) -> Reflection:
    """Both reflections (symmetry.jl:128-130)."""
    return Reflection((int(parity_x), int(parity_y)), "origin")


def rotation_characters(n: int, m: int) -> np.ndarray:
    """chi_l = cis(2 pi (m mod n) l / n), l = 0..n-1 (_rotation_tables,
    symmetry.jl:64-82)."""
    l = np.arange(n, dtype=np.float64)
    return np.exp(1j * (TWO_PI * ((m % n) * l) / n))


def node_divisor(symmetry) -> int:
    """The divisibility the node count N must meet (cfie.jl:538-549,
    dlp_kress.jl:363-371)."""
    if symmetry is None:
        return 2
    if isinstance(symmetry, Rotation):
        return 2 * symmetry.n // math.gcd(2, symmetry.n)
    return 4


def _scales(sym) -> np.ndarray:
    if isinstance(sym, Rotation):
        return rotation_characters(sym.n, sym.m)
    if sym.axis == "origin":
        px, py = sym.parity
        return np.array([1, px, py, px * py], dtype=np.complex128)
    return np.array([1, sym.parity], dtype=np.complex128)


# ---------------------------------------------------------------------------
# coordinate maps (CFIE_kress)
# ---------------------------------------------------------------------------
def _match(targets: np.ndarray, xy: np.ndarray, tol: float) -> np.ndarray:
    """match_index for every target point (symmetry.jl:590-609): the unique
    node within tol; no match or an ambiguous one is an error."""
    from scipy.spatial import cKDTree
    tree = cKDTree(xy)
    d, idx = tree.query(targets, k=2)
    if np.any(d[:, 0] >= tol):
        raise ValueError(f"no match (dmin={float(np.max(d[:, 0]))})")
    if np.any(d[:, 1] < tol):
        raise ValueError("ambiguous match")
    return idx[:, 0].astype(np.int64)


def _flatten(pts) -> Tuple[np.ndarray, np.ndarray]:
    comps = [pts] if hasattr(pts, "xy") else list(pts)
    xy = np.concatenate([p.xy for p in comps])
    offs = np.concatenate([[0], np.cumsum([len(p) for p in comps])]).astype(np.int64)
    return xy, offs


def _rotation_maps(pts, sym: Rotation, tol: float) -> List[np.ndarray]:
    """build_rotation_maps_components (symmetry.jl:648-730): for l = 1..n-1,
    rotate each component and find the component and the single cyclic shift
    it lands on; returns the global permutation of each rotation power."""
    comps = [pts] if hasattr(pts, "xy") else list(pts)
    xy, offs = _flatten(pts)
    cx, cy = sym.center
    n = sym.n
    th = TWO_PI / n
    c1, s1 = math.cos(th), math.sin(th)
    cos_t, sin_t = [1.0], [0.0]
    for _ in range(1, n):  # the recurrence of _rotation_tables
        c, s = cos_t[-1], sin_t[-1]
        cos_t.append(c * c1 - s * s1)
        sin_t.append(s * c1 + c * s1)
    maps = [np.arange(len(xy), dtype=np.int64)]
    for l in range(1, n):
        c, s = cos_t[l], sin_t[l]
        mp = np.empty(len(xy), dtype=np.int64)
        for a, pa in enumerate(comps):
            dx, dy = pa.xy[:, 0] - cx, pa.xy[:, 1] - cy
            rx, ry = cx + c * dx - s * dy, cy + s * dx + c * dy
            found = False
            for b, pb in enumerate(comps):
                if len(pb) != len(pa):
                    continue
                d0 = (pb.xy[:, 0] - rx[0]) ** 2 + (pb.xy[:, 1] - ry[0]) ** 2
                jb = int(np.argmin(d0))
                if d0[jb] > tol * tol:
                    continue
                jj = (np.arange(len(pa)) + jb) % len(pb)
                d = (pb.xy[jj, 0] - rx) ** 2 + (pb.xy[jj, 1] - ry) ** 2
                if np.any(d > tol * tol):
                    continue
                mp[offs[a]:offs[a] + len(pa)] = offs[b] + jj
                found = True
                break
            if not found:
                raise ValueError(f"Could not find rotated target component for component {a + 1}")
        maps.append(mp)
    return maps


def build_symmetry_maps(pts, sym, tol: float = 1e-10) -> List[np.ndarray]:
    """Node maps of the group elements in the order of the orbit scales
    (build_symmetry_maps, symmetry.jl:746-766): identity first."""
    xy, _ = _flatten(pts)
    ident = np.arange(len(xy), dtype=np.int64)
    if isinstance(sym, Reflection):
        mx = lambda: _match(np.column_stack([-xy[:, 0], xy[:, 1]]), xy, tol)
        my = lambda: _match(np.column_stack([xy[:, 0], -xy[:, 1]]), xy, tol)
        if sym.axis == "y_axis":
            return [ident, mx()]
        if sym.axis == "x_axis":
            return [ident, my()]
        if sym.axis == "origin":
            mxy = _match(-xy, xy, tol)
            return [ident, mx(), my(), mxy]
        raise ValueError(f"Unknown reflection axis {sym.axis}")
    if isinstance(sym, Rotation):
        return _rotation_maps(pts, sym, max(1e-8, tol))
    raise ValueError(f"Unknown symmetry type {type(sym).__name__}")


def cfie_orbits(pts, sym, tol: float = 1e-10) -> Tuple[np.ndarray, np.ndarray]:
    """Orbit table of the reduced CFIE workspace (cheb-CFIE:475-505): nodes
    visited in order, each unvisited node g opens orbit [map_l(g)]_l.
    Returns img (m x G int64, column 0 = Ifund) and scale (m x G complex)."""
    maps = build_symmetry_maps(pts, sym, tol)
    sc = _scales(sym)
    N = len(maps[0])
    M = np.stack(maps, axis=1)  # M[g, l] = image l of node g
    visited = np.zeros(N, dtype=bool)
    reps = []
    for g in range(N):
        if visited[g]:
            continue
        reps.append(g)
        visited[M[g]] = True
    img = M[np.array(reps, dtype=np.int64)]
    return img, np.tile(sc, (len(reps), 1))


def dlp_orbits(N: int, sym) -> Tuple[np.ndarray, np.ndarray]:
    """periodic_symmetry_index_orbits (symmetry.jl:214-298) in 0-based form."""
    if isinstance(sym, Rotation):
        n = sym.n
        if N % n:
            raise ValueError(f"N = {N} is not divisible by the rotation order {n}")
        f = np.arange(N // n, dtype=np.int64)
        img = np.stack([(f + l * (N // n)) % N for l in range(n)], axis=1)
    elif sym.axis in ("y_axis", "x_axis"):
        if N % 2:
            raise ValueError("a reflection needs an even N")
        f = np.arange(N // 2, dtype=np.int64)
        img = np.stack([f, (N - 1 - f) % N], axis=1)  # q1 -> N - q1 + 1
    elif sym.axis == "origin":
        if N % 4:
            raise ValueError("the origin reflection needs N divisible by 4")
        f = np.arange(N // 4, dtype=np.int64)
        img = np.stack([f, (N // 2 - 1 - f) % N, (N - 1 - f) % N, (f + N // 2) % N], axis=1)
    else:
        raise ValueError(f"Unsupported reflection axis {sym.axis}")
    sc = _scales(sym)
    return img, np.tile(sc, (img.shape[0], 1))


def orbits_for(solver, pts) -> Optional[Tuple[np.ndarray, np.ndarray]]:
    """The orbit table the reference's reduced workspace of `solver` holds
    (None without symmetry)."""
    sym = getattr(solver, "symmetry", None)
    if sym is None:
        return None
    from . import _lib
    if solver.op == _lib.QB_DLP_KRESS:
        comps = [pts] if hasattr(pts, "xy") else list(pts)
        return dlp_orbits(int(sum(len(p) for p in comps)), sym)
    return cfie_orbits(pts, sym)
