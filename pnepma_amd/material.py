"""Material models at the boundary of the hot path (host side).

The solver consumes per-point partial densities rho[point][element] for each of
the eight staggered classes; this module produces them and maps density
gradients back to the material parameters.  Only the models the workloads
need are restated:

* ``MassFractionMaterial`` -- rho_e = w_e / sum_e(w_e / rho_pure_e)
  (pn_solver/pn_parametrizations/pn_material.jl:199-264, non-SC variant) with
  its pullback (same file, component_densities_pullback!).
* ``ConstFractions`` -- the same fractions everywhere
  (pn_parametrizations/pn_parametrization_const.jl:33-39).
* ``CellFractions`` -- piecewise constant fractions on an axis-aligned box grid
  (pn_parametrizations/pn_parametrization_boxed.jl:137-175): a point belongs to
  the box whose lower edge is the largest edge <= x (a point on the upper
  outer edge belongs to the last box); points outside use ``embedding``.
"""
from __future__ import annotations

import numpy as np

from .physics import Element


class ConstFractions:
    def __init__(self, w):
        self.w = np.asarray(w, dtype=np.float64)

    @property
    def n_params(self) -> int:
        return self.w.size

    def params(self) -> np.ndarray:
        return self.w.copy()

    def with_params(self, p) -> "ConstFractions":
        return ConstFractions(np.asarray(p, dtype=np.float64).reshape(self.w.shape))

    def fractions(self, pts: np.ndarray) -> np.ndarray:
        return np.broadcast_to(self.w, (pts.shape[0], self.w.size)).copy()

    def pullback(self, pts: np.ndarray, wbar: np.ndarray) -> np.ndarray:
        return wbar.sum(axis=0)


class CellFractions:
    """Per-box fractions ``vals[ix, iy, iz, e]`` on edges ``x_edges`` (len nx+1) etc."""

    def __init__(self, vals, x_edges, y_edges, z_edges, embedding):
        self.vals = np.asarray(vals, dtype=np.float64)
        self.edges = [np.asarray(e, dtype=np.float64) for e in (x_edges, y_edges, z_edges)]
        self.embedding = np.asarray(embedding, dtype=np.float64)
        assert self.vals.shape[:3] == tuple(len(e) - 1 for e in self.edges)

    @property
    def n_params(self) -> int:
        return self.vals.size

    def params(self) -> np.ndarray:
        return self.vals.ravel().copy()

    def with_params(self, p) -> "CellFractions":
        return CellFractions(np.asarray(p, dtype=np.float64).reshape(self.vals.shape), *self.edges, self.embedding)

    def cells_uniform(self, pts):
        """Cell indices by arithmetic on uniformly spaced edges (for generic points
        such as absorption ray samples, which do not sit on edges)."""
        idx = []
        inside = np.ones(pts.shape[0], dtype=bool)
        for d, e in enumerate(self.edges):
            n = len(e) - 1
            i = np.floor((pts[:, d] - e[0]) * (n / (e[-1] - e[0]))).astype(np.int64)
            inside &= (i >= 0) & (i < n)
            idx.append(np.clip(i, 0, n - 1))
        return idx, inside

    def linear_cells(self, pts):
        """Flat cell index of each generic point, -1 outside the boxes."""
        (ix, iy, iz), inside = self.cells_uniform(pts)
        nx, ny, nz = self.vals.shape[:3]
        lin = (ix * ny + iy) * nz + iz
        return np.where(inside, lin, -1)

    def fractions_fast(self, pts: np.ndarray) -> np.ndarray:
        (ix, iy, iz), inside = self.cells_uniform(pts)
        w = self.vals[ix, iy, iz]
        w[~inside] = self.embedding
        return w

    def _cells(self, pts):
        idx = []
        inside = np.ones(pts.shape[0], dtype=bool)
        for d, e in enumerate(self.edges):
            x = pts[:, d]
            i = np.searchsorted(e, x, side="right") - 1
            i = np.where(x == e[-1], len(e) - 2, i)
            inside &= (i >= 0) & (i <= len(e) - 2)
            idx.append(np.clip(i, 0, len(e) - 2))
        return idx, inside

    def fractions(self, pts: np.ndarray) -> np.ndarray:
        (ix, iy, iz), inside = self._cells(pts)
        w = self.vals[ix, iy, iz]
        w[~inside] = self.embedding
        return w

    def pullback(self, pts: np.ndarray, wbar: np.ndarray) -> np.ndarray:
        (ix, iy, iz), inside = self._cells(pts)
        g = np.zeros_like(self.vals)
        np.add.at(g, (ix[inside], iy[inside], iz[inside]), wbar[inside])
        return g.ravel()


class MassFractionMaterial:
    def __init__(self, elements, param):
        self.elements = list(elements)
        self.param = param
        self.rho_pure = np.array([e.density for e in self.elements])

    @property
    def n_elements(self) -> int:
        return len(self.elements)

    def with_params(self, p) -> "MassFractionMaterial":
        return MassFractionMaterial(self.elements, self.param.with_params(p))

    def densities(self, pts: np.ndarray, fast: bool = False) -> np.ndarray:
        """component_densities! at each point: shape (P, n_e).  fast=True uses the
        parametrization's arithmetic lookup where it has one (generic points only)."""
        w = self.param.fractions_fast(pts) if fast and hasattr(self.param, "fractions_fast") \
            else self.param.fractions(pts)
        R = (w / self.rho_pure).sum(axis=1, keepdims=True)
        return w / R

    def densities_pullback(self, pts: np.ndarray, rho_bar: np.ndarray) -> np.ndarray:
        """Gradient w.r.t. the parametrization's parameters of sum(rho_bar * rho(pts))."""
        w = self.param.fractions(pts)
        R = (w / self.rho_pure).sum(axis=1, keepdims=True)
        R_bar = -(w * rho_bar).sum(axis=1, keepdims=True) / R ** 2
        w_bar = rho_bar / R + R_bar / self.rho_pure
        return self.param.pullback(pts, w_bar)
