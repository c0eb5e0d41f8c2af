"""The per-energy records: the oracle's restatement (oracle/pnx_ref_records.c)
pinned by closed forms, and the library's C++ set-up (pnx_problem_records,
pnx_beam_moments, pnx_beam_profiles) against the oracle and the numpy
formulation.

Pins of the oracle (each a property the reference's formulas fix):
  * at a knot of the stopping-power table the stopping power is the table
    value, and midway between two knots the mean of theirs (a linear
    interpolant, stopping_power.jl:37-44);
  * the forward coefficient of degree 0 is minus the slope of the segment,
    and forward minus adjoint coefficients is minus that slope at every
    degree (pn_problem.jl:157-190);
  * the beam bins of the forward records telescope: sum_k gfac_k dE =
    Phi(E_initial + dE/2) - Phi(E_cutoff + dE/2) (pn_problem.jl:205-211 at
    the energies E - dE/2 of pn_intensities.jl:341-349);
  * the cross sections of the forward records are those at E_initial - k dE,
    of the density records at E_cutoff + (n - i) dE and of the adjoint
    records half a step above (pn_intensities.jl:391-401);
  * the K-shell stand-in is 6.51e-14 * 1.8 / (e E_K^2) at U = e and zero at
    the edge.
"""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle
from pnepma_amd import physics, serving, workloads

from .helpers import rel_err


def _fields(p, rec):
    ne, N = len(p.elements), p.N
    s = rec[..., :ne]
    c = rec[..., ne:ne + ne * (N + 1)].reshape(rec.shape[:-1] + (ne, N + 1))
    g = rec[..., ne + ne * (N + 1)]
    sig = rec[..., ne + ne * (N + 1) + 1:]
    return s, c, g, sig


def test_records_oracle_closed_forms():
    p, _, _, _ = workloads.raster2d(N=5, n_steps=120, n_beams=1)
    kn, _ = physics.stopping_power_table(p.elements[0])
    tables = [physics.stopping_power_table(e)[1] for e in p.elements]
    inside = [k for k in range(len(kn) - 1) if 1.0e3 <= kn[k] and kn[k + 1] <= 20e3]
    for k in inside:
        s, c, _, _ = _fields(p, oracle.record_at(p, kn[k], kn[k]))
        assert np.array_equal(s, [t[k] for t in tables])
        slope = np.array([(t[k + 1] - t[k]) / (kn[k + 1] - kn[k]) for t in tables])
        assert np.allclose(c[:, 0], -slope, rtol=1e-15, atol=0)
        mid = 0.5 * (kn[k] + kn[k + 1])
        s_m, c_m, _, _ = _fields(p, oracle.record_at(p, mid, mid))
        _, c_a, _, _ = _fields(p, oracle.record_at(p, mid, mid, adjoint=True))
        assert np.allclose(s_m, [0.5 * (t[k] + t[k + 1]) for t in tables], rtol=1e-14, atol=0)
        assert np.allclose(c_m - c_a, -slope[:, None], rtol=1e-12, atol=1e-12 * np.abs(c_m).max())
    fwd, adj, dens = oracle.records(p)
    n, dE = p.n_steps, p.dE
    mu, sd = p.beam_energy
    _, _, g, _ = _fields(p, fwd)
    Phi = lambda x: special.ndtr((x - mu) / sd)
    assert abs(g[1:].sum() * dE - (Phi(p.E_initial + dE / 2) - Phi(p.E_cutoff + dE / 2))) < 1e-12
    edges = [x.edge for x in p.x_rays]

    def sigma(E):
        return np.array([6.51e-14 * 1.8 * math.log(E / e) / (E / e * e * e) if E > e else 0.0 for e in edges])
    for k in (0, 1, n // 2, n):
        assert rel_err(_fields(p, fwd[k])[3], sigma(p.E_initial - k * dE)) < 1e-12
    for i in (0, 1, n // 3, n):
        E = p.E_cutoff + (n - i) * dE
        assert rel_err(_fields(p, dens[i])[3], sigma(E)) < 1e-12
        assert rel_err(_fields(p, adj[i])[3], sigma(E + dE / 2)) < 1e-12
    e0 = edges[0]
    r = oracle.record_at(p, 5e3, e0 * math.e)
    assert abs(_fields(p, r)[3][0] - 6.51e-14 * 1.8 / (math.e * e0 * e0)) < 1e-15 * 6.51e-14 / e0 ** 2
    assert _fields(p, oracle.record_at(p, 5e3, e0))[3][0] == 0.0


def _problems():
    out = [workloads.slab1d(n_steps=300)[0], workloads.raster2d(n_steps=400, n_beams=3)[0],
           workloads.inclusion3d(M=(24, 24, 16), n_steps=60, n_beams=3)[0]]
    return out + [s.problem for s in serving.serving_requests(12, seed=3)]


def test_library_records_match_oracle_and_numpy():
    for p in _problems():
        lib_rec, ora, host = p.records(), oracle.records(p), p.records_host()
        for a, b, h in zip(lib_rec, ora, host):
            sc = np.maximum(np.abs(b).max(axis=0), 1e-300)
            assert np.max(np.abs(a - b) / sc) < 1e-15
            assert np.max(np.abs(h - b) / sc) < 1e-13


@pytest.mark.parametrize("kappa", [0.5, 5.0, 50.0, 300.0])
def test_library_beam_moments_and_profiles(kappa):
    """pnx_beam_moments against the numpy Funk-Hecke closed form (itself pinned
    by sphere quadrature in test_oracle_pins), pnx_beam_profiles against the
    numpy pdf, in the lateral and the literal source modes."""
    for N in (1, 3, 5, 7, 9):
        for d in ((-1.0, 0.0, 0.0), (-0.8, 0.3, 0.2), (0.1, -0.6, 0.7)):
            p, _, _, _ = workloads.raster2d(N=N, M=(8, 6), n_steps=10, n_beams=1)
            p.beam_direction = (d, kappa)
            mm, mh = p.beam_moments(), p.beam_moments_host()
            sc = max(np.abs(x).max() for x in mh if x is not None and x.size)
            assert max(np.abs(x - y).max() for x, y in zip(mm, mh) if x is not None and x.size) < 1e-14 * sc
    for p in _problems()[:3]:
        for mode in ("lateral", "literal"):
            p.source_mode = mode
            for b in range(p.n_beams):
                for x, y in zip(p.beam_profiles(b), p.beam_profiles_host(b)):
                    if x is not None:
                        assert rel_err(x, y) < 1e-14


def test_beam_moments_vanishing_by_symmetry_are_exact_zeros():
    """A beam along an axis is symmetric under the reflections of the other two
    axes, so every moment odd under one of them vanishes.  Both the library and
    the numpy formulation evaluate the harmonics from the direction's
    components and give those moments as exact zeros (from the angles they
    would carry residues such as cos(pi/2) ~ 6e-17, which the boundary source
    multiplies by the position pdf)."""
    from pnepma_amd import grid as G
    for N in (1, 3, 5, 7, 9):
        p, _, _, _ = workloads.raster2d(N=N, M=(8, 6), n_steps=10, n_beams=1)
        cm = G.class_moments(N)
        for d, odd_axes in (((-1.0, 0.0, 0.0), (1, 2)), ((0.0, 0.0, 1.0), (0, 1)), ((0.0, -2.0, 0.0), (0, 2))):
            p.beam_direction = (d, 40.0)
            for mm in (p.beam_moments(), p.beam_moments_host()):
                for c in range(8):
                    if mm[c] is None:
                        continue
                    for n, (l, k) in enumerate(cm[c]):
                        # parity of Y_lk under y -> -y: odd for k < 0; under
                        # z -> -z: (-1)^(l + |k|); under x -> -x: odd for
                        # k > 0 with |k| odd or k < 0 with |k| even
                        odd = {1: k < 0, 2: (l + abs(k)) % 2 == 1,
                               0: (k > 0 and abs(k) % 2 == 1) or (k < 0 and abs(k) % 2 == 0)}
                        if any(odd[a] for a in odd_axes):
                            assert mm[c][n] == 0.0, (N, d, c, l, k, mm[c][n])
