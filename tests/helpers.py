"""Shared builders for the parity tests: small seeded problems at the shapes of
the measurement workloads, interior masks, and error measures."""
from __future__ import annotations

import numpy as np

from pnepma_amd import grid as G
from pnepma_amd import workloads
from pnepma_amd.problem import solve_inputs


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def interior_mask(gspec, c):
    """Points of class c that the solver updates (eo_range per axis)."""
    n = gspec.class_size(c)
    ranges = []
    for d in range(3):
        m = np.zeros(n[d], dtype=bool)
        if G.odd_in(c, d):
            m[1:n[d] - 1] = True
        else:
            m[:] = True
        ranges.append(m)
    Z, Y, X = np.meshgrid(ranges[2], ranges[1], ranges[0], indexing="ij")
    return (X & Y & Z).ravel()


def small_1d(N=5, cells=40, n_steps=24, seed=0, E0=15e3, E1=5e3, depth=0.4e-4):
    p, mat, stds, det = workloads.slab1d(N=N, cells=cells, n_steps=n_steps, depth=depth, E0=E0, E1=E1, seed=seed)
    return p, mat, stds, det


def small_2d(N=3, M=(12, 10), n_steps=16, n_beams=2, seed=1, E0=15e3, E1=6e3, extent=(0.3e-4, 0.2e-4)):
    p, mat, stds, det = workloads.raster2d(N=N, M=M, n_steps=n_steps, extent=extent, E0=E0, E1=E1,
                                           n_beams=n_beams, seed=seed)
    return p, mat, stds, det


def random_state(gspec, N, rng, active_only=True):
    cm = G.class_moments(N)
    u = []
    for c in range(8):
        n = int(np.prod(gspec.class_size(c)))
        if (active_only and not gspec.class_active(c)) or len(cm[c]) == 0:
            u.append(np.zeros((n, max(len(cm[c]), 1))))
        else:
            u.append(rng.standard_normal((n, len(cm[c]))))
    return u


def inputs(p, mat, det):
    return solve_inputs(p, mat, det)
