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


def stable_steps(p, mat, cfl=0.5) -> int:
    """Energy steps that put the march at the stability number nu <= cfl."""
    from pnepma_amd import serving
    return serving.cfl_steps(p.N, p.gspec, mat, p.elements, p.E_initial, p.E_cutoff, cfl=cfl)


def cfl_of(p, mat) -> float:
    from pnepma_amd import serving
    return serving.cfl_number(p.N, p.gspec, mat, p.elements, p.E_initial, p.E_cutoff, p.n_steps)


def small_1d(N=5, cells=40, n_steps=None, seed=0, E0=15e3, E1=5e3, depth=0.4e-4, unstable_ok=False):
    """A small Cu/Ni slab.  n_steps None: the step count of nu <= 0.5.  An
    explicit step count must also be stable unless unstable_ok (single-step
    arithmetic tests, which compare one step of a random state)."""
    kw = dict(N=N, cells=cells, depth=depth, E0=E0, E1=E1, seed=seed)
    p, mat, stds, det = workloads.slab1d(n_steps=n_steps or 8, **kw)
    if n_steps is None:
        p, mat, stds, det = workloads.slab1d(n_steps=stable_steps(p, mat), **kw)
    elif not unstable_ok:
        assert cfl_of(p, mat) <= 0.5, f"unstable test configuration: nu = {cfl_of(p, mat):.2f}"
    return p, mat, stds, det


def small_2d(N=3, M=(12, 10), n_steps=None, n_beams=2, seed=1, E0=15e3, E1=6e3, extent=(0.3e-4, 0.2e-4),
             unstable_ok=False):
    """A small Fe/Cr/Ni raster problem; n_steps as small_1d."""
    kw = dict(N=N, M=M, extent=extent, E0=E0, E1=E1, n_beams=n_beams, seed=seed)
    p, mat, stds, det = workloads.raster2d(n_steps=n_steps or 8, **kw)
    if n_steps is None:
        p, mat, stds, det = workloads.raster2d(n_steps=stable_steps(p, mat), **kw)
    elif not unstable_ok:
        assert cfl_of(p, mat) <= 0.5, f"unstable test configuration: nu = {cfl_of(p, mat):.2f}"
    return p, mat, stds, det


def assert_physical(I):
    """Intensities of a stable march: finite, nonzero and of one sign (the
    reference's beam moments are -int Y vMF, so they come out negative)."""
    I = np.asarray(I)
    assert np.all(np.isfinite(I)), I
    assert np.all(I < 0) or np.all(I > 0), I


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
