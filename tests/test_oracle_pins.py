"""Pins of the CPU oracle by mathematics and by an independent formulation.

The reference ships no golden vectors for this path (no solver tests, no
fixtures), and it is Julia, which cannot run here.  The oracle is pinned by:
  * closed forms and invariants of the PN matrices (class sizes, the
    spherical-harmonic recursion Omega_d Y = A_d Y, symmetry, spectrum =
    Gauss-Legendre nodes, half-range quadrature of the boundary coupling);
  * known answers of the step (pure decay of a uniform state);
  * an independent dense numpy formulation of the step and of the pullback
    loop on tiny grids (tests/numpy_step.py);
  * exact linearity in the beam source and the literal-source underflow.
"""
import math

import numpy as np
import pytest

from pnepma_amd import grid as G
from pnepma_amd import physics
from pnepma_amd.material import ConstFractions, MassFractionMaterial

from . import numpy_step
from .helpers import inputs, interior_mask, random_state, rel_err, small_1d, small_2d


# --------------------------------------------------------------------- moments
@pytest.mark.parametrize("N", range(0, 10))
def test_class_sizes_closed_form(N, oracle_lib):
    cm = G.class_moments(N)
    expect = [G.n_eee(N), G.n_eee(N - 1), G.n_eee(N - 1), G.n_eee(N - 2), G.n_eee(N - 1), G.n_eee(N - 2),
              G.n_eee(N - 2), G.n_eee(N - 3)]
    assert [len(x) for x in cm] == expect
    assert sum(expect) == (N + 1) ** 2
    assert [oracle_lib.lib().ref_class_moments(N, c) for c in range(8)] == expect


def test_active_moment_counts():
    """Active moments per dimension (1D: eee+oee, 2D: z-even classes, 3D: all)."""
    table = {3: (6, 10, 16), 5: (12, 21, 36), 7: (20, 36, 64), 9: (30, 55, 100)}
    for N, counts in table.items():
        cm = G.class_moments(N)
        assert len(cm[0]) + len(cm[4]) == counts[0]
        assert sum(len(cm[c]) for c in (0, 2, 4, 6)) == counts[1]
        assert sum(len(x) for x in cm) == counts[2]


def _full(N, d, blocks):
    allm = G.all_moments(N)
    idx = {m: i for i, m in enumerate(allm)}
    cm = G.class_moments(N)
    A = np.zeros((len(allm), len(allm)))
    for c in range(8):
        f = G.switch_axis(c, d)
        blk = blocks(N, c, d, len(cm[c]), len(cm[f]))
        for i, mi in enumerate(cm[c]):
            for j, mj in enumerate(cm[f]):
                A[idx[mi], idx[mj]] = blk[i, j]
    return A, allm


@pytest.mark.parametrize("N", [1, 3, 5, 7, 9])
def test_advection_recursion_symmetry_spectrum(N, oracle_lib):
    rng = np.random.default_rng(N)
    th, ph = rng.uniform(0, np.pi, 25), rng.uniform(0, 2 * np.pi, 25)
    om = [np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]
    gl = np.polynomial.legendre.leggauss(N + 1)[0]
    for d in range(3):
        A, allm = _full(N, d, oracle_lib.advection_dense)
        Y = np.array([physics.real_sh(l, k, th, ph) for (l, k) in allm])
        rows = [i for i, (l, _) in enumerate(allm) if l < N]
        assert np.abs((om[d][None, :] * Y - A @ Y)[rows]).max() < 1e-13
        assert np.array_equal(A, A.T)
        ev = np.linalg.eigvalsh(A)
        assert abs(np.abs(ev).max() - gl.max()) < 1e-13
    Az, allm = _full(N, 2, oracle_lib.advection_dense)
    k0 = [i for i, (l, k) in enumerate(allm) if k == 0]
    assert np.allclose(np.sort(np.linalg.eigvalsh(Az[np.ix_(k0, k0)])), np.sort(gl), atol=1e-13)


def _halfrange_L(N, c, d, nq=64, nphi=256):
    """2 int_{Omega_d > 0} Y_i Y_j / Omega_d dOmega with mu = Omega_d on (0, 1)."""
    x, w = np.polynomial.legendre.leggauss(nq)
    mu = 0.5 * (x + 1)
    wm = 0.5 * w
    psi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    MU, PS = np.meshgrid(mu, psi, indexing="ij")
    s = np.sqrt(1 - MU ** 2)
    # local frame: Omega_d = mu, the other two components on a circle
    comps = [None, None, None]
    comps[d] = MU
    others = [a for a in range(3) if a != d]
    comps[others[0]] = s * np.cos(PS)
    comps[others[1]] = s * np.sin(PS)
    th = np.arccos(np.clip(comps[2], -1, 1))
    ph = np.arctan2(comps[1], comps[0])
    cm = G.class_moments(N)[c]
    Y = np.array([physics.real_sh(l, k, th, ph) for (l, k) in cm])
    W = (wm[:, None] * (2 * np.pi / nphi)) / MU
    return 2.0 * np.einsum("iab,jab,ab->ij", Y, Y, W)


@pytest.mark.parametrize("N", [3, 5, 7])
def test_boundary_blocks_halfrange(N, oracle_lib):
    import ctypes as C
    L_ = oracle_lib.lib()
    L_.ref_boundary_L.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    cm = G.class_moments(N)
    for d in range(3):
        for c in range(8):
            if not G.odd_in(c, d) or not cm[c]:
                continue
            n = len(cm[c])
            Lo = np.zeros((n, n))
            L_.ref_boundary_L(N, c, d, Lo.ctypes.data_as(C.POINTER(C.c_double)))
            Lq = _halfrange_L(N, c, d)
            assert np.abs(Lo - Lq).max() < 1e-12 * max(1.0, np.abs(Lq).max()), (N, c, d)
            f = G.switch_axis(c, d)
            A = oracle_lib.advection_dense(N, c, d, n, len(cm[f]))
            Bf = oracle_lib.boundary_dense(N, c, d, False, n, len(cm[f]))
            Ba = oracle_lib.boundary_dense(N, c, d, True, n, len(cm[f]))
            assert np.abs(Bf - Lq @ A).max() < 1e-12
            assert np.abs(Ba + Lq.T @ A).max() < 1e-12


# ------------------------------------------------------------------- the step
def test_uniform_state_decays_exactly(oracle_lib):
    """Far from the boundaries a spatially uniform state only decays:
    u' = u exp(-C dE / S) per moment over one full Strang step."""
    p, mat, _, det = small_1d(N=5, cells=30)
    mat = MassFractionMaterial(mat.elements, ConstFractions([0.4, 0.6]))
    si = inputs(p, mat, det)
    cm = G.class_moments(p.N)
    u = [np.tile(np.linspace(0.5, 1.5, max(len(cm[c]), 1)), (int(np.prod(p.gspec.class_size(c))), 1))
         if p.gspec.class_active(c) else np.zeros((int(np.prod(p.gspec.class_size(c))), max(len(cm[c]), 1)))
         for c in range(8)]
    u0 = [x.copy() for x in u]
    fwd, _, _ = p.records()
    rec = fwd[3]
    oracle_lib.step(p, si, u, rec)
    ne = len(p.elements)
    s, cl = rec[:ne], rec[ne:ne + ne * (p.N + 1)].reshape(ne, p.N + 1)
    for c in (0, 4):
        rho = si.rho[c][0]
        S = rho @ s
        ls = np.array([l for l, _ in cm[c]])
        Cm = rho @ cl[:, ls]
        expect = u0[c][0] * np.exp(-Cm * p.dE / S)
        far = u[c][5:-5]
        assert np.abs(far - expect[None, :]).max() < 1e-13 * np.abs(expect).max()


def _dense_vs_oracle(p, mat, det, oracle_lib, adjoint, record, seed):
    si = inputs(p, mat, det)
    A, B = numpy_step.dense_blocks(p.N, adjoint, oracle_lib.advection_dense, oracle_lib.boundary_dense)
    D = numpy_step.DenseProblem(p, si, A, B, adjoint)
    fwd, adj, _ = p.records()
    rec = adj[record] if adjoint else fwd[record]
    D.set_record(rec)
    rng = np.random.default_rng(seed)
    u = random_state(p.gspec, p.N, rng)
    q = None
    Q = None
    if adjoint:
        q = rng.standard_normal(u[0].shape[0])
        Q = D.zeros()
        Q[0][..., 0] = q.reshape(Q[0].shape[:3])
    ud = [u[c].reshape(D.size[c] + (-1,))[..., :len(D.cm[c])].copy() for c in range(8)]
    uo = [x.copy() for x in u]
    D.step(ud, p.dE, Q)
    oracle_lib.step(p, si, uo, rec, adjoint=adjoint, q_eee=q)
    for c in p.gspec.active_classes():
        if not D.cm[c]:
            continue
        m = interior_mask(p.gspec, c)
        e = rel_err(uo[c][m], ud[c].reshape(-1, len(D.cm[c]))[m])
        assert e < 1e-13, (G.CLASS_NAMES[c], e)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_step_dense_numpy_1d(N, oracle_lib):
    p, mat, _, det = small_1d(N=N, cells=12)
    _dense_vs_oracle(p, mat, det, oracle_lib, False, 2, N)


@pytest.mark.parametrize("N,adjoint", [(3, False), (5, False), (3, True), (7, True)])
def test_step_dense_numpy_2d(N, adjoint, oracle_lib):
    p, mat, _, det = small_2d(N=N, M=(6, 5))
    _dense_vs_oracle(p, mat, det, oracle_lib, adjoint, 3, 10 + N)


def test_pullback_dense_numpy(oracle_lib):
    p, mat, _, det = small_2d(N=3, M=(6, 5), n_beams=1, E1=13e3)
    si = inputs(p, mat, det)
    ib = np.array([1.0, -0.4, 0.7])
    I, rb, acb, noab = numpy_step.pullback(p, si, oracle_lib.advection_dense, oracle_lib.boundary_dense, ib)
    Io, rbo, acbo, noabo = oracle_lib.forward_adjoint(p, si, ib)
    assert rel_err(Io, I) < 1e-13
    for c in p.gspec.active_classes():
        assert rel_err(rbo[c], rb[c]) < 1e-12, G.CLASS_NAMES[c]
    assert rel_err(acbo, acb) < 1e-12
    assert rel_err(noabo, noab) < 1e-12


# ------------------------------------------------------------ whole solve
def test_source_linearity_exact(oracle_lib):
    """Intensities are linear in the beam source: scaling it by 2^-20 scales them exactly."""
    p, mat, stds, det = small_1d(N=3, cells=30)
    si = inputs(p, mat, det)
    I1 = oracle_lib.forward(p, si)
    p.beam_sigma_scale = None
    orig = p.beam_profiles_host

    def scaled(b):
        return [None if x is None else x * 2.0 ** -20 for x in orig(b)]

    p.beam_profiles_host = scaled
    I2 = oracle_lib.forward(p, si)
    assert np.array_equal(I2, I1 * 2.0 ** -20)


def test_literal_source_underflows():
    """The reference evaluates the beam pdf at x_min: for the 2 um slab and a
    50 nm beam the factor exp(-800) underflows and every intensity is 0."""
    from pnepma_amd import workloads
    from oracle import oracle
    p, mat, _, det = workloads.slab1d(source_mode="literal", n_steps=20)
    si = inputs(p, mat, det)
    prof = p.beam_profiles(0)
    assert all(np.all(x == 0.0) for x in prof if x is not None)
    assert np.all(oracle.forward(p, si) == 0.0)


def test_cfl_literal_step_count_is_unstable(oracle_lib):
    """150 steps (CFL 0.94) diverge from the converged k-ratios; 300 steps (CFL 0.47)
    agree with 600 steps to 1e-3."""
    from pnepma_amd import workloads

    def k(ns):
        p, mat, stds, det = workloads.slab1d(n_steps=ns)
        I = oracle_lib.forward(p, inputs(p, mat, det))
        Is = np.array([oracle_lib.forward(p, inputs(p, s, det))[i] for i, s in enumerate(stds)])
        return I / Is

    k150, k300, k600 = k(150), k(300), k(600)
    assert np.max(np.abs(k300 - k600) / np.abs(k600)) < 1e-3
    assert np.max(np.abs(k150 - k600) / np.abs(k600)) > 0.1


def test_beam_moments_funk_hecke():
    """Closed-form vMF moments against direct quadrature of -Y_lk vMF over the sphere."""
    N, kappa = 5, 8.0
    mu = np.array([-1.0, 0.0, 0.0])
    moms = G.all_moments(N)
    closed = physics.vmf_beam_moments(moms, mu, kappa)
    x, w = np.polynomial.legendre.leggauss(200)
    nphi = 400
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    TH, PH = np.meshgrid(np.arccos(x), ph, indexing="ij")
    W = w[:, None] * (2 * np.pi / nphi)
    om = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)])
    pdf = kappa / (4 * np.pi * math.sinh(kappa)) * np.exp(kappa * np.tensordot(mu, om, axes=1))
    quad = np.array([-(physics.real_sh(l, k, TH, PH) * pdf * W).sum() for (l, k) in moms])
    assert np.abs(closed - quad).max() < 1e-12


# ----------------------------------------- the reference adjoint against FD
def _adjoint_gap(oracle_lib, M, mult):
    """Signed relative gap (adjoint - FD) / FD between the oracle's reference
    adjoint and central differences (h = 1e-6) of its forward, for the
    directional derivative of J = I_bar . I along a smooth density
    perturbation: a Gaussian bump (width 0.05 um, 0.08 um below the surface,
    weights +1 / -0.7 on Cu / Ni) on every class.  1D P3 Cu/Ni slab 0.4 um
    deep on M cells, 15 -> 5 keV in mult x the stable step count."""
    import copy
    from pnepma_amd import workloads
    from .helpers import stable_steps
    kw = dict(N=3, cells=M, depth=0.4e-4, E0=15e3, E1=5e3)
    p, mat, _, det = workloads.slab1d(n_steps=8, **kw)
    p, mat, _, det = workloads.slab1d(n_steps=stable_steps(p, mat) * mult, **kw)
    si = inputs(p, mat, det)
    ib = np.array([1.0, 0.5])
    _, rb, _, _ = oracle_lib.forward_adjoint(p, si, ib)
    dirs = []
    for c in range(8):
        x = p.gspec.class_points(c)[:, 0]
        bump = np.exp(-((x + 0.08e-4) / 0.05e-4) ** 2)
        dirs.append(si.rho[c] * bump[:, None] * np.array([1.0, -0.7])[None, :])
    ad = sum(float((rb[c] * dirs[c]).sum()) for c in range(8))

    def J(eps):
        s = copy.copy(si)
        s.rho = [r + eps * d for r, d in zip(si.rho, dirs)]
        return float(oracle_lib.forward(p, s) @ ib)
    h = 1e-6
    fd = (J(h) - J(-h)) / (2 * h)
    return (ad - fd) / fd


def _plateau(oracle_lib, M):
    """The gap extrapolated to dE -> 0 at M cells (Richardson, first order in dE)."""
    return 2.0 * _adjoint_gap(oracle_lib, M, 8) - _adjoint_gap(oracle_lib, M, 4)


def test_reference_adjoint_converges_to_central_differences(oracle_lib):
    """The reference adjoint (pn_intensities.jl:353-404) discretises the
    continuous adjoint, so it is not the gradient of the discrete march; what
    fixes it is that the gap to central differences of the forward vanishes at
    first order as the energy step and the cell size shrink.
      * dE: at 40 cells the gap's change per halving of dE halves (first order);
      * h: the gap extrapolated to dE -> 0 halves when the cells are halved.
    A wrong sign of the adjoint boundary blocks, a dropped density term or a
    wrong sign of the adjoint source leaves a gap that does not vanish
    (test_adjoint_refinement_detects_planted_mistakes)."""
    g = [_adjoint_gap(oracle_lib, 40, m) for m in (2, 4, 8, 16)]
    d = np.diff(g)
    r_t = d[1:] / d[:-1]
    P = {M: _plateau(oracle_lib, M) for M in (40, 80, 160)}
    r_h = [P[80] / P[40], P[160] / P[80]]
    print(f"dE refinement at 40 cells: gaps {g}, ratios of changes {r_t}", flush=True)
    print(f"h refinement (dE -> 0): {P}, ratios {r_h}", flush=True)
    assert np.all((r_t > 0.25) & (r_t < 0.65)) and 0.4 < r_t[-1] < 0.6
    assert all(0.4 < r < 0.6 for r in r_h)
    assert abs(P[160]) < 1e-3


@pytest.mark.parametrize("plant", [1, 2, 3], ids=["boundary_sign", "dropped_Cu_term", "source_sign"])
def test_adjoint_refinement_detects_planted_mistakes(oracle_lib, plant):
    """Each mistake planted in the oracle's adjoint leaves a gap to central
    differences that does not halve with the cell size."""
    oracle_lib.set_plant(plant)
    try:
        P40, P80 = _plateau(oracle_lib, 40), _plateau(oracle_lib, 80)
    finally:
        oracle_lib.set_plant(0)
    print(f"plant {plant}: dE -> 0 gap {P40:.3e} at 40 cells, {P80:.3e} at 80", flush=True)
    converges = np.isfinite(P80) and abs(P80) < 1e-3 and 0.4 < P80 / P40 < 0.6
    assert not converges
