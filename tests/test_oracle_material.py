"""Pins of the oracle's material layer (oracle/pnx_ref_material.c), and the
host material code against it.

The reference's material tests compare hand-written pullbacks with finite
differences (pn_parametrizations/pn_material_test.jl:10-153); the same is done
here for the MassFraction + box densities, the absorption factors and the
numbers of atoms, together with closed forms:
  * sum_e rho_e / rho_pure_e = 1 and rho proportional to the box's fractions
    (found by brute force on the edges), the embedding outside the boxes, the
    upper outer edge in the last box;
  * a uniform material attenuates as exp(-(mac . rho) * path length to the
    surface plane), with the path length from geometry;
  * noa = rho / atomic mass.
The absorption pullback (x_ray_intensities.jl:131-153) is the exact derivative
of the 100-sample ray sum with respect to the box values (the samples and dx
do not depend on the material), so finite differences pin it too.
"""
import numpy as np
import pytest

from oracle import oracle
from pnepma_amd import physics, workloads
from pnepma_amd.grid import GridSpec
from pnepma_amd.material import CellFractions, MassFractionMaterial
from pnepma_amd.problem import absorption_factors, absorption_pullback

from .helpers import rel_err


def _material(M=(6, 5, 4), ne=3, seed=0, extent=((-1e-4, 0.0), (-0.5e-4, 0.5e-4), (-0.4e-4, 0.4e-4))):
    rng = np.random.default_rng(seed)
    names = ["Fe", "Cr", "Ni", "Cu"][:ne]
    els = [physics.element(s) for s in names]
    g = GridSpec.make(M, *[e if m > 1 else 0.0 for e, m in zip(extent, M)])
    vals = rng.uniform(0.1, 1.0, M + (ne,))
    mat = MassFractionMaterial(els, CellFractions(vals, *[workloads.cell_edges(a) for a in g.axes],
                                                  embedding=rng.uniform(0.1, 1.0, ne)))
    return g, mat, rng


def _mac(mat, lines):
    return np.array([[physics.mass_absorption(physics.XRay(s), e) for e in mat.elements] for s in lines])


def test_densities_closed_forms():
    g, mat, rng = _material()
    par = mat.param
    lo = [e[0] for e in par.edges]
    hi = [e[-1] for e in par.edges]
    # random points inside the boxes (away from edges), a margin outside, the upper outer corner
    pts = np.stack([rng.uniform(l - 0.2 * (h - l), h + 0.2 * (h - l), 400) for l, h in zip(lo, hi)], axis=1)
    rho = oracle.component_densities(mat, pts)
    assert np.allclose((rho / mat.rho_pure).sum(axis=1), 1.0, rtol=0, atol=1e-14)
    for p, r in zip(pts, rho):
        idx = [np.searchsorted(e, x, side="right") - 1 for e, x in zip(par.edges, p)]
        inside = all(0 <= i < len(e) - 1 for i, e in zip(idx, par.edges))
        w = par.vals[tuple(idx)] if inside else par.embedding
        assert np.allclose(r / r.sum(), w / w.sum(), rtol=1e-14, atol=0)
    corner = np.array([hi[0], hi[1], hi[2]])
    assert oracle.box_index(mat, corner) == int(np.prod(par.vals.shape[:3])) - 1
    assert oracle.box_index(mat, np.array(lo) - 1e-9) == -1


def test_densities_pullback_matches_central_differences():
    g, mat, rng = _material(seed=1)
    pts = np.concatenate([g.class_points(c) for c in range(8)])
    rb = rng.standard_normal((pts.shape[0], len(mat.elements)))
    grad, _ = oracle.component_densities_pullback(mat, pts, rb)
    x = mat.param.params()
    h = 1e-6
    for i in rng.choice(x.size, 12, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (np.sum(rb * oracle.component_densities(mat.with_params(xp), pts)) -
              np.sum(rb * oracle.component_densities(mat.with_params(xm), pts))) / (2 * h)
        assert abs(grad[i] - fd) <= 1e-7 * max(abs(fd), 1e-3 * np.abs(grad).max()), (i, grad[i], fd)


def test_absorption_uniform_closed_form():
    g, mat, rng = _material(seed=2)
    w = np.array([0.2, 0.5, 0.3])
    mat.param.vals[...] = w
    mat.param.embedding[...] = w
    det = workloads.detector_40deg()
    mac = _mac(mat, ["Fe", "Ni"])
    pts = g.class_points(0)
    ac = oracle.absorption(mat, pts, mac, det, 0.0)
    rho = w / (w / mat.rho_pure).sum()
    d = det[None, :] - pts
    path = np.linalg.norm(d, axis=1) * np.maximum(-pts[:, 0] / d[:, 0], 0.0)
    expect = np.exp(-(mac @ rho)[None, :] * path[:, None])
    assert rel_err(ac, expect) < 1e-13


def test_absorption_pullback_matches_central_differences():
    g, mat, rng = _material(seed=3)
    det = workloads.detector_40deg()
    mac = _mac(mat, ["Fe", "Cr", "Ni"])
    pts = g.class_points(0)
    acb = rng.standard_normal((pts.shape[0], mac.shape[0]))
    ac = oracle.absorption(mat, pts, mac, det, 0.0)
    grad, _ = oracle.absorption_pullback(mat, pts, mac, det, 0.0, ac, acb)
    x = mat.param.params()
    h = 1e-6
    for i in rng.choice(x.size, 12, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (np.sum(acb * oracle.absorption(mat.with_params(xp), pts, mac, det, 0.0)) -
              np.sum(acb * oracle.absorption(mat.with_params(xm), pts, mac, det, 0.0))) / (2 * h)
        # the sum over every eee point leaves rounding noise of ~1e-10 |J| / h in fd
        assert abs(grad[i] - fd) <= 1e-5 * max(abs(fd), 1e-2 * np.abs(grad).max()), (i, grad[i], fd)


def test_number_of_atoms():
    g, mat, rng = _material(seed=4)
    am = np.array([e.atom_mass_g for e in mat.elements])
    pts = g.class_points(0)
    noa = oracle.number_of_atoms(mat, pts, am)
    assert rel_err(noa, oracle.component_densities(mat, pts) / am) < 1e-15
    nb = rng.standard_normal(noa.shape)
    g1, _ = oracle.number_of_atoms_pullback(mat, pts, am, nb)
    g2, _ = oracle.component_densities_pullback(mat, pts, nb / am)
    assert rel_err(g1, g2) < 1e-15


@pytest.mark.parametrize("M", [(30, 1, 1), (12, 9, 1), (6, 5, 4)], ids=["1d", "2d", "3d"])
def test_host_material_layer_matches_oracle(M):
    """The package's host material code (densities at every class point, the
    absorption march, and their pullbacks) against the oracle."""
    g, mat, rng = _material(M=M, seed=5)
    det = workloads.detector_40deg()
    lines = [physics.XRay(s) for s in ("Fe", "Ni")]
    mac = _mac(mat, ["Fe", "Ni"])
    for c in range(8):
        pts = g.class_points(c)
        assert rel_err(mat.densities(pts), oracle.component_densities(mat, pts)) < 1e-15
        rb = rng.standard_normal((pts.shape[0], len(mat.elements)))
        assert rel_err(mat.densities_pullback(pts, rb), oracle.component_densities_pullback(mat, pts, rb)[0]) < 1e-14
    ac, aux = absorption_factors(mat, lines, g, det, 0.0)
    ac_o = oracle.absorption(mat, g.class_points(0), mac, det, 0.0)
    assert rel_err(ac, ac_o) < 1e-14
    acb = rng.standard_normal(ac.shape)
    assert rel_err(absorption_pullback(mat, aux, ac, acb),
                   oracle.absorption_pullback(mat, g.class_points(0), mac, det, 0.0, ac_o, acb)[0]) < 1e-13


def test_literal_floor_differs_only_on_interior_edges():
    """The reference finds a point's box as floor((x - lo) / step) in floating
    point (biliner_interp.jl:5-15).  For a grid-aligned box material the x-odd
    class points lie exactly on box edges, and that floor sends some of them
    to the lower box, depending on rounding.  The oracle, the host code and
    the device layer all put a point on an interior edge in the box above it
    (the floor in exact arithmetic); this shows that the two lookups differ
    only at such points."""
    g, mat, rng = _material(M=(30, 9, 1), seed=6)
    edges = mat.param.edges
    differ = on_edge = 0
    for c in range(8):
        for x in g.class_points(c):
            a, b = oracle.box_index(mat, x), oracle.box_index(mat, x, literal_floor=True)
            edge = any(np.any(e[1:-1] == xi) for e, xi in zip(edges, x))
            on_edge += edge
            if a != b:
                differ += 1
                assert edge, x
    print(f"{differ} of {on_edge} points on interior edges change box under the floating-point floor", flush=True)
    assert 0 < differ < on_edge


@pytest.mark.parametrize("M", [(30, 1, 1), (12, 9, 1), (6, 5, 4)], ids=["1d", "2d", "3d"])
def test_cell_box_densities_equal_the_eee_densities(M):
    """On a material with one box per grid cell, every class's densities at the
    points the kernels update (odd axes: indices 1..M-1, even axes 0..M-1) are
    the eee densities at the same index: odd point i lies on the lower edge of
    box i, which the oracle assigns to box i.  The kernels of cell-fraction
    problems read the eee array for every class on this ground."""
    g, mat, _ = _material(M=M, seed=7)
    eee = oracle.component_densities(mat, g.class_points(0)).reshape(M[2], M[1], M[0], -1)
    for c in range(8):
        if not g.class_active(c):
            continue
        n = g.class_size(c)
        rho = oracle.component_densities(mat, g.class_points(c)).reshape(n[2], n[1], n[0], -1)
        lo = [1 if (n[a] == M[a] + 1) else 0 for a in range(3)]  # odd axes skip the ghost at 0
        upd = rho[lo[2]:M[2], lo[1]:M[1], lo[0]:M[0]]
        assert np.array_equal(upd, eee[lo[2]:, lo[1]:, lo[0]:]), c
