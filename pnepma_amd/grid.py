"""Moments, even/odd classes and the staggered grid (host side).

Mirrors the bookkeeping of pn_solver/moment_specification.jl, pn_solver/pn_variable.jl
and starmap/src/staggered_grid.jl, starmap/src/even_odd_classification.jl.

Classes are indexed 0..7 in the order eee, eeo, eoe, eoo, oee, oeo, ooe, ooo
(``all_eos()``), i.e. index = 4*odd_x + 2*odd_y + odd_z.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

CLASS_NAMES = ("eee", "eeo", "eoe", "eoo", "oee", "oeo", "ooe", "ooo")


def x_odd(k: int) -> int:
    return 0 if ((k < 0 and k % 2 != 0) or (k >= 0 and k % 2 == 0)) else 1


def y_odd(k: int) -> int:
    return 0 if k >= 0 else 1


def z_odd(l: int, k: int) -> int:
    return 0 if (l + k) % 2 == 0 else 1


def class_of(l: int, k: int) -> int:
    return 4 * x_odd(k) + 2 * y_odd(k) + z_odd(l, k)


def odd_in(c: int, d: int) -> int:
    return (c >> (2 - d)) & 1


def switch_axis(c: int, d: int) -> int:
    return c ^ (4 >> d)


def parity_of(c: int) -> int:
    """0 for even parity {eee, ooe, oeo, eoo}, 1 for odd parity."""
    return (odd_in(c, 0) + odd_in(c, 1) + odd_in(c, 2)) & 1


def all_moments(N: int):
    return [(l, k) for l in range(N + 1) for k in range(-l, l + 1)]


def class_moments(N: int):
    """Moments of each class, in all_moments order (pn_solver/pn_variable.jl:12)."""
    out = [[] for _ in range(8)]
    for l, k in all_moments(N):
        out[class_of(l, k)].append((l, k))
    return out


def n_eee(N: int) -> int:
    """pn_solver/pn_variable.jl:2 closed form."""
    return int(max(0, (N + 2) // 2 * ((N + 4) // 2) // 2)) if N >= 0 else 0


@dataclass(frozen=True)
class Axis:
    """One axis of the staggered grid: M even points on [lo, hi] (h = (hi-lo)/(M-1))
    and M+1 odd points from lo-h/2 to hi+h/2 including two ghost cells; or a
    singleton axis at position ``lo`` with step 1 (starmap/src/staggered_grid.jl:6-78)."""
    M: int
    lo: float
    hi: float

    @property
    def singleton(self) -> bool:
        return self.M == 1

    @property
    def h(self) -> float:
        return 1.0 if self.singleton else (self.hi - self.lo) / (self.M - 1)

    def size(self, odd: int) -> int:
        if self.singleton:
            return 1
        return self.M + 1 if odd else self.M

    def points(self, odd: int) -> np.ndarray:
        if self.singleton:
            return np.array([self.lo])
        if odd:
            h = self.h
            return np.linspace(self.lo - h / 2, self.hi + h / 2, self.M + 1)
        return np.linspace(self.lo, self.hi, self.M)


@dataclass(frozen=True)
class GridSpec:
    """make_gridspec(M, x, y, z) (starmap/src/staggered_grid.jl:170-174).  Singleton
    axes must be trailing: 1D = x only, 2D = x, y."""
    axes: tuple

    @staticmethod
    def make(M, x, y, z) -> "GridSpec":
        axes = []
        for m, ext in zip(M, (x, y, z)):
            if m == 1:
                p = float(ext if np.isscalar(ext) else ext[0])
                axes.append(Axis(1, p, p))
            else:
                axes.append(Axis(int(m), float(ext[0]), float(ext[1])))
        g = GridSpec(tuple(axes))
        d = g.dims
        if any(not a.singleton for a in g.axes[d:]):
            raise ValueError("non-singleton axes must come first (x, then y, then z)")
        return g

    @property
    def dims(self) -> int:
        return sum(1 for a in self.axes if not a.singleton)

    @property
    def M(self):
        return tuple(a.M for a in self.axes)

    @property
    def h(self):
        return tuple(a.h for a in self.axes)

    @property
    def dvol(self) -> float:
        h = self.h
        return h[0] * h[1] * h[2]

    def class_size(self, c: int):
        return tuple(self.axes[d].size(odd_in(c, d)) for d in range(3))

    def class_points(self, c: int) -> np.ndarray:
        """Coordinates of every point of class c, x fastest: shape (nx*ny*nz, 3)."""
        px, py, pz = (self.axes[d].points(odd_in(c, d)) for d in range(3))
        Z, Y, X = np.meshgrid(pz, py, px, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def class_active(self, c: int) -> bool:
        d = self.dims
        return (d >= 2 or odd_in(c, 1) == 0) and (d >= 3 or odd_in(c, 2) == 0)

    def active_classes(self):
        return [c for c in range(8) if self.class_active(c)]
