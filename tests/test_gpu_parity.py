"""GPU parity: libpnx.so on the device against the CPU oracle on the same inputs.

Tolerances: single steps 1e-12 (relative to the largest value), k-ratios and
intensities 1e-10, gradients (density and generation pullbacks) 1e-8.

Every whole march here runs at a stable step count (nu <= 0.5, see
serving.cfl_number), so the compared intensities are the physical solution
(finite, of one sign).  The single-step tests are arithmetic identities: one
step of a random state, where the stability of a march does not enter.
"""
import numpy as np
import pytest

from pnepma_amd import grid as G
from pnepma_amd.solver import DeviceProblem

from .helpers import assert_physical, cfl_of, inputs, interior_mask, random_state, rel_err, small_1d, small_2d

pytestmark = pytest.mark.gpu


def _step_case(p, mat, det, oracle_lib, adjoint, record, seed=3):
    si = inputs(p, mat, det)
    dp = DeviceProblem(p, si)
    rng = np.random.default_rng(seed)
    u0 = random_state(p.gspec, p.N, rng)
    u_gpu = [x.copy() for x in u0]
    u_ref = [x.copy() for x in u0]
    fwd, adj, dens = p.records()
    ib = np.linspace(0.5, 1.5, len(p.x_rays)) if adjoint else None
    dp.step(u_gpu, record, beam=0, adjoint=adjoint, I_bar=ib)
    rec = adj[record] if adjoint else fwd[record]
    q = None
    if adjoint:
        sig = rec[len(p.elements) * (p.N + 2) + 1:]
        q = -(si.W * sig[None, :] * ib[None, :]).sum(axis=1)
    oracle_lib.step(p, si, u_ref, rec, adjoint=adjoint, q_eee=q)
    dp.close()
    for c in p.gspec.active_classes():
        if len(G.class_moments(p.N)[c]) == 0:
            continue
        m = interior_mask(p.gspec, c)
        e = rel_err(u_gpu[c][m], u_ref[c][m])
        print(f"class {G.CLASS_NAMES[c]} step rel err {e:.3e}", flush=True)
        assert e < 1e-12, (G.CLASS_NAMES[c], e)


@pytest.mark.parametrize("N", [1, 3, 5, 7, 9])
def test_step_forward_1d(N, oracle_lib):
    p, mat, _, det = small_1d(N=N)
    _step_case(p, mat, det, oracle_lib, adjoint=False, record=5)


@pytest.mark.parametrize("N", [3, 5, 7, 9])
def test_step_forward_2d(N, oracle_lib):
    p, mat, _, det = small_2d(N=N)
    _step_case(p, mat, det, oracle_lib, adjoint=False, record=7)


@pytest.mark.parametrize("N", [3, 7])
def test_step_adjoint_2d(N, oracle_lib):
    p, mat, _, det = small_2d(N=N)
    _step_case(p, mat, det, oracle_lib, adjoint=True, record=4)


def test_step_forward_3d(oracle_lib):
    from pnepma_amd.grid import GridSpec
    p, mat, _, det = small_2d(N=5, M=(8, 6))
    p.gspec = GridSpec.make((8, 6, 5), (-0.3e-4, 0.0), (-1e-5, 1e-5), (-0.8e-5, 0.8e-5))
    p.beam_positions = [(0.0, 0.0, 0.0)]
    from pnepma_amd.material import ConstFractions, MassFractionMaterial
    mat = MassFractionMaterial(mat.elements, ConstFractions([0.3, 0.3, 0.4]))
    _step_case(p, mat, det, oracle_lib, adjoint=False, record=3)


def test_forward_intensities_1d(oracle_lib):
    p, mat, stds, det = small_1d(N=5, cells=60)
    si = inputs(p, mat, det)
    dp = DeviceProblem(p, si, adjoint=False)
    I = dp.forward()[0]
    I_ref = oracle_lib.forward(p, si)
    print(f"{p.n_steps} steps, nu {cfl_of(p, mat):.3f}: I gpu", I, "ref", I_ref, flush=True)
    assert_physical(I_ref)
    assert rel_err(I, I_ref) < 1e-10


def test_forward_intensities_2d_beams(oracle_lib):
    p, mat, stds, det = small_2d(N=5, M=(16, 12), n_beams=3)
    si = inputs(p, mat, det)
    dp = DeviceProblem(p, si, adjoint=False)
    I = dp.forward()
    for b in range(p.n_beams):
        I_ref = oracle_lib.forward(p, si, beam=b)
        print(f"{p.n_steps} steps, nu {cfl_of(p, mat):.3f}: beam", b, I[b], I_ref, flush=True)
        assert_physical(I_ref)
        assert rel_err(I[b], I_ref) < 1e-10


@pytest.mark.parametrize("ckpt", [0, 5])
def test_forward_adjoint_2d(oracle_lib, ckpt):
    p, mat, stds, det = small_2d(N=3, M=(10, 8), n_beams=2)
    si = inputs(p, mat, det)
    dp = DeviceProblem(p, si, ckpt_every=ckpt)
    ib = np.array([[1.0, -0.5, 0.25], [0.3, 0.7, -1.1]])
    I, rb, gen = dp.forward_adjoint(ib)
    rb_ref = [np.zeros_like(r) for r in rb]
    acb_ref = np.zeros_like(si.ac)
    for b in range(p.n_beams):
        I_ref, rbr, acb, noab = oracle_lib.forward_adjoint(p, si, ib[b], beam=b)
        assert_physical(I_ref)
        assert rel_err(I[b], I_ref) < 1e-10
        for c in range(8):
            rb_ref[c] += rbr[c]
        acb_ref += acb
    for c in p.gspec.active_classes():
        e = rel_err(rb[c], rb_ref[c])
        print(f"rho_bar {G.CLASS_NAMES[c]} rel err {e:.3e}", flush=True)
        assert e < 1e-8
    acb_gpu = si.noa[:, si.el_idx] * gen
    e = rel_err(acb_gpu, acb_ref)
    print(f"ac_bar rel err {e:.3e}", flush=True)
    assert e < 1e-8


def test_kratios_slab1d_workload(oracle_lib):
    """The 1D slab workload (P5, 200 cells, Cu/Ni) at its CFL-0.5 step count."""
    from pnepma_amd import workloads
    p, mat, stds, det = workloads.slab1d(n_steps=300)
    si = inputs(p, mat, det)
    I = DeviceProblem(p, si, adjoint=False).forward()[0]
    I_ref = oracle_lib.forward(p, si)
    Is = np.array([DeviceProblem(p, inputs(p, s, det), adjoint=False).forward()[0][i] for i, s in enumerate(stds)])
    Is_ref = np.array([oracle_lib.forward(p, inputs(p, s, det))[i] for i, s in enumerate(stds)])
    k, k_ref = I / Is, I_ref / Is_ref
    print("k gpu", k, "k ref", k_ref, flush=True)
    assert np.max(np.abs(k - k_ref) / np.abs(k_ref)) < 1e-10


def test_rev_bench_on_a_fresh_problem():
    """The reverse-pass timing entry point allocates the adjoint buffers itself
    (a problem that has had no adjoint solve has none)."""
    from pnepma_amd import workloads
    from pnepma_amd.problem import solve_inputs
    p, mat, stds, det = workloads.raster2d(M=(40, 30), n_steps=6, n_beams=2)
    dp = DeviceProblem(p, solve_inputs(p, mat, det), ckpt_every=3)
    ms, by = dp.rev_bench(3)
    print(f"rev bench: {ms * 1e3:.1f} us, {by / 1e6:.2f} MB", flush=True)
    assert ms > 0 and by > 0
    # the problem still solves afterwards
    I, rb, gen = dp.forward_adjoint(np.ones((p.n_beams, len(p.x_rays))))
    assert np.all(np.isfinite(I))
    dp.close()
