"""The measurement workloads, built from seeds.

* ``slab1d``   -- 1D forward: P5, slab 0-2 um deep on 200 cells, 15 keV Gaussian
                  beam (sigma 50 nm), 150 energy steps 15 -> 1 keV, Cu/Ni with
                  uniform mass fractions w_Cu ~ U(0.2, 0.8) (seed 0).
* ``raster2d`` -- 2D forward: P7, 128 x 96 grid over 3 x 2 um (depth x lateral),
                  200 energy steps 20 -> 1 keV, Fe/Cr/Ni with per-cell fractions
                  from 6 random rectangular inclusions (seed 1), 32 beams along
                  the lateral axis.
* ``raster2d_grad`` -- the same geometry with a least-squares k-ratio misfit
                  against seeded targets (seed 2); gradient w.r.t. per-cell mass
                  fractions with checkpoints every 16 energy steps.

Choices not fixed by the workload statements (frozen here): beam energy spread
sigma_E = 0.1 keV, von Mises-Fisher direction -x with kappa = 50, lines K-L3,
detector at 40 degrees take-off 1 cm away, E_initial = the stated start energy.
Every workload also has a ``scale`` argument that shrinks the grid and step
count for quick parity tests.
"""
from __future__ import annotations

import math

import numpy as np

from .grid import GridSpec
from .material import CellFractions, ConstFractions, MassFractionMaterial
from .physics import XRay, element
from .problem import ForwardPNProblem

UM = 1e-4  # cm
NM = 1e-7
KAPPA = 50.0
SIGMA_E = 100.0


def detector_40deg(distance=1.0):
    a = math.radians(40.0)
    return np.array([distance * math.sin(a), distance * math.cos(a), 0.0])


def slab1d(N=5, cells=200, n_steps=150, depth=2 * UM, E0=15e3, E1=1e3, seed=0, source_mode="lateral"):
    rng = np.random.default_rng(seed)
    w_cu = rng.uniform(0.2, 0.8)
    els = [element("Cu"), element("Ni")]
    g = GridSpec.make((cells, 1, 1), (-depth, 0.0), 0.0, 0.0)
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(E0, SIGMA_E), beam_positions=[(0.0, 0.0, 0.0)],
                         beam_sigma=50 * NM, beam_direction=((-1.0, 0.0, 0.0), KAPPA), E_initial=E0, E_cutoff=E1,
                         n_steps=n_steps, source_mode=source_mode)
    p.x_rays = [XRay("Cu"), XRay("Ni")]
    mat = MassFractionMaterial(els, ConstFractions([w_cu, 1.0 - w_cu]))
    stds = [MassFractionMaterial(els, ConstFractions([1.0, 0.0])), MassFractionMaterial(els, ConstFractions([0.0, 1.0]))]
    return p, mat, stds, detector_40deg()


def inclusion_fractions(M, n_e, n_incl, rng):
    """Per-cell fractions: U(0,1) base, then n_incl rectangles (corners uniform in
    the cell grid, fractions U(0,1)); later rectangles win; renormalised per cell."""
    vals = np.empty(M + (n_e,))
    vals[...] = rng.uniform(0.0, 1.0, n_e)
    for _ in range(n_incl):
        a = [sorted(rng.integers(0, m + 1, 2)) for m in M]
        f = rng.uniform(0.0, 1.0, n_e)
        sl = tuple(slice(lo, max(hi, lo + 1)) for lo, hi in a)
        vals[sl] = f
    return vals / vals.sum(axis=-1, keepdims=True)


def cell_edges(axis):
    """Box edges centred on the even points of an axis: the odd points."""
    if axis.singleton:
        return np.array([axis.lo - 0.5, axis.lo + 0.5])
    return axis.points(1)


def raster2d(N=7, M=(128, 96), n_steps=200, extent=(3 * UM, 2 * UM), E0=20e3, E1=1e3, n_beams=32, seed=1,
             source_mode="lateral"):
    rng = np.random.default_rng(seed)
    names = ["Fe", "Cr", "Ni"]
    els = [element(s) for s in names]
    g = GridSpec.make((M[0], M[1], 1), (-extent[0], 0.0), (-extent[1] / 2, extent[1] / 2), 0.0)
    ys = np.linspace(-0.45 * extent[1], 0.45 * extent[1], n_beams)
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(E0, SIGMA_E),
                         beam_positions=[(0.0, float(y), 0.0) for y in ys], beam_sigma=50 * NM,
                         beam_direction=((-1.0, 0.0, 0.0), KAPPA), E_initial=E0, E_cutoff=E1, n_steps=n_steps,
                         source_mode=source_mode)
    p.x_rays = [XRay(s) for s in names]
    vals = inclusion_fractions((M[0], M[1], 1), len(els), 6, rng)
    edges = [cell_edges(a) for a in g.axes]
    mat = MassFractionMaterial(els, CellFractions(vals, *edges, embedding=np.full(len(els), 1.0 / len(els))))
    stds = [MassFractionMaterial(els, ConstFractions(np.eye(len(els))[i])) for i in range(len(els))]
    return p, mat, stds, detector_40deg()


def inclusion3d(N=5, M=(128, 128, 96), n_steps=300, extent=(2 * UM, 2 * UM, 1.5 * UM), E0=25e3, E1=1e3,
                n_beams=64, seed=4, source_mode="lateral"):
    """3D P5 Al/Si block with a seeded ellipsoidal Si-rich inclusion, scanned by an
    8 x 8 beam raster in (y, z): the 3D forward+adjoint workload."""
    rng = np.random.default_rng(seed)
    names = ["Al", "Si"]
    els = [element(s) for s in names]
    g = GridSpec.make(M, (-extent[0], 0.0), (-extent[1] / 2, extent[1] / 2), (-extent[2] / 2, extent[2] / 2))
    ys = np.linspace(-0.4 * extent[1], 0.4 * extent[1], 8)
    zs = np.linspace(-0.4 * extent[2], 0.4 * extent[2], 8)
    pos = [(0.0, float(y), float(z)) for y in ys for z in zs][:n_beams]
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(E0, SIGMA_E), beam_positions=pos,
                         beam_sigma=50 * NM, beam_direction=((-1.0, 0.0, 0.0), KAPPA), E_initial=E0, E_cutoff=E1,
                         n_steps=n_steps, source_mode=source_mode)
    p.x_rays = [XRay(s) for s in names]
    centre = [rng.uniform(-0.7, -0.3) * extent[0], rng.uniform(-0.2, 0.2) * extent[1],
              rng.uniform(-0.2, 0.2) * extent[2]]
    semi = [rng.uniform(0.15, 0.3) * e for e in extent]
    X, Y, Z = np.meshgrid(*[a.points(0) for a in g.axes], indexing="ij")
    inside = ((X - centre[0]) / semi[0]) ** 2 + ((Y - centre[1]) / semi[1]) ** 2 + ((Z - centre[2]) / semi[2]) ** 2 < 1
    vals = np.empty(M + (2,))
    vals[...] = (0.9, 0.1)
    vals[inside] = (0.2, 0.8)
    mat = MassFractionMaterial(els, CellFractions(vals, *[cell_edges(a) for a in g.axes], embedding=vals[0, 0, 0]))
    stds = [MassFractionMaterial(els, ConstFractions(np.eye(2)[i])) for i in range(2)]
    return p, mat, stds, detector_40deg()
