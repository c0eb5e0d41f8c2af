"""Parity at the full sizes of the measurement workloads.

* 2D P7 128x96, 400 energy steps (the raster workload): intensities of the
  first and last beam and of the three pure-element standards against the
  oracle (1e-10), hence k-ratios to 2e-10.
* The same geometry forward+adjoint for one beam with checkpoints every 16
  steps: density gradients against the oracle's restatement of the reference
  pullback (1e-8).
* Size-independent properties of the 32-beam batch: every intensity finite and
  of one sign, beams symmetric about the axis give mirror intensities only if
  the material is symmetric (not asserted), results independent of the
  checkpoint interval (exactly) and of the batch composition (a beam solved
  alone equals the same beam inside the batch).
"""
import copy

import numpy as np
import pytest

from pnepma_amd import grid as G
from pnepma_amd import workloads
from pnepma_amd.problem import solve_inputs
from pnepma_amd.solver import DeviceProblem

from .helpers import rel_err

pytestmark = pytest.mark.gpu


def _single_beam(p, b):
    q = copy.copy(p)
    q._cache = {}
    q.beam_positions = [p.beam_positions[b]]
    return q


def test_raster2d_intensities_and_kratios(oracle_lib):
    p, mat, stds, det = workloads.raster2d(n_steps=400)
    si = solve_inputs(p, mat, det)
    I = DeviceProblem(p, si, adjoint=False).forward()
    assert np.all(np.isfinite(I))
    for b in (0, p.n_beams - 1):
        I_ref = oracle_lib.forward(p, si, beam=b)
        e = rel_err(I[b], I_ref)
        print(f"beam {b} intensity rel err {e:.3e}", flush=True)
        assert e < 1e-10
    q = _single_beam(p, 0)
    k_gpu, k_ref = [], []
    for i, s in enumerate(stds):
        ss = solve_inputs(q, s, det)
        k_gpu.append(DeviceProblem(q, ss, adjoint=False).forward()[0][i])
        k_ref.append(oracle_lib.forward(q, ss)[i])
    k_gpu = I[0] / np.array(k_gpu)
    k_ref = oracle_lib.forward(p, si, beam=0) / np.array(k_ref)
    print("k-ratios", k_gpu, k_ref, flush=True)
    assert np.max(np.abs(k_gpu - k_ref) / np.abs(k_ref)) < 2e-10


def test_raster2d_gradient_one_beam(oracle_lib):
    p, mat, stds, det = workloads.raster2d(n_steps=400)
    q = _single_beam(p, 5)
    si = solve_inputs(q, mat, det)
    ib = np.array([1.0, 0.7, -0.4])
    I, rb, gen = DeviceProblem(q, si, ckpt_every=16).forward_adjoint(ib[None, :])
    I_ref, rb_ref, acb_ref, _ = oracle_lib.forward_adjoint(q, si, ib)
    assert rel_err(I[0], I_ref) < 1e-10
    for c in q.gspec.active_classes():
        e = rel_err(rb[c], rb_ref[c])
        print(f"rho_bar {G.CLASS_NAMES[c]} rel err {e:.3e}", flush=True)
        assert e < 1e-8
    assert rel_err(si.noa[:, si.el_idx] * gen, acb_ref) < 1e-8


def test_raster2d_batch_properties():
    p, mat, stds, det = workloads.raster2d(n_steps=400)
    si = solve_inputs(p, mat, det)
    ib = np.random.default_rng(2).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    dp16 = DeviceProblem(p, si, ckpt_every=16)
    I16, rb16, g16 = dp16.forward_adjoint(ib)
    dp16.close()
    dp50 = DeviceProblem(p, si, ckpt_every=50)
    I50, rb50, g50 = dp50.forward_adjoint(ib)
    dp50.close()
    # the checkpoint interval only moves the boundaries of the fused odd-pair
    # runs; print how far apart the two results are, then bound it
    e_I = rel_err(I16, I50)
    e_rb = max(rel_err(rb16[c], rb50[c]) for c in range(8) if np.any(rb50[c]))
    e_g = rel_err(g16, g50)
    print(f"ckpt 16 vs 50: I {e_I:.3e} rho_bar {e_rb:.3e} gen {e_g:.3e}", flush=True)
    assert e_I < 1e-13 and e_rb < 1e-12 and e_g < 1e-13
    # a beam alone equals the same beam in the batch
    q = _single_beam(p, 7)
    Iq = DeviceProblem(q, solve_inputs(q, mat, det), adjoint=False).forward()[0]
    assert rel_err(Iq, I16[7]) < 1e-13


def test_raster2d_misfit_gradient_20_cells():
    """The 2D gradient workload: 128x96 P7, 32 beams, 400 energy steps,
    checkpoints every 16, least-squares k-ratio misfit against targets
    k_truth (1 + N(0, 0.01)) at truth + N(0, 0.05) (seed 2).  The gradient with
    respect to the per-cell mass fractions (exact adjoint of the discrete
    march, material pullbacks on the device) against central differences
    (h = 1e-6) of the misfit on 20 random cells the beams reach, to 1e-5.  The reference
    adjoint's gradient is printed beside it: it discretises the continuous
    adjoint, so it differs from the differences by its first-order gap."""
    from pnepma_amd.reconstruct import KRatioMisfit, standard_intensities
    p, truth, stds, det = workloads.raster2d(n_steps=400)
    rng = np.random.default_rng(2)
    I_std = standard_intensities(p, stds, det)
    mf = KRatioMisfit(p, truth, det, I_std, np.zeros((p.n_beams, len(p.x_rays))), ckpt_every=16)
    k_truth = mf.kratios(truth.param.params())
    mf.k_exp = k_truth * (1 + rng.normal(0, 0.01, k_truth.shape))
    ne = len(truth.elements)
    x = np.clip(truth.param.params() + rng.normal(0, 0.05, truth.param.n_params), 0.01, None).reshape(-1, ne)
    x = (x / x.sum(axis=1, keepdims=True)).ravel()
    L, g = mf.loss_and_grad(x)
    mf_ref = KRatioMisfit(p, truth, det, I_std, mf.k_exp, ckpt_every=16, adjoint="reference")
    _, g_ref = mf_ref.loss_and_grad(x)
    mf_ref.close()
    # 20 random cells among those the beams reach (a gradient above 1 % of the
    # largest): deeper cells have gradients below 1e-20, which central
    # differences of a misfit of 1e-3 cannot resolve
    gc = np.abs(g).reshape(-1, ne).max(axis=1)
    cells = rng.choice(np.flatnonzero(gc > 1e-2 * gc.max()), 20, replace=False)
    h = 1e-6
    worst = worst_ref = 0.0
    for cell in cells:
        i = cell * ne + int(rng.integers(ne))
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (mf.loss(xp) - mf.loss(xm)) / (2 * h)
        scale = max(abs(fd), 1e-2 * np.max(np.abs(g)))
        e, e_ref = abs(g[i] - fd) / scale, abs(g_ref[i] - fd) / scale
        print(f"cell {cell} param {i}: exact {g[i]:.6e} reference {g_ref[i]:.6e} fd {fd:.6e} "
              f"rel {e:.2e} / {e_ref:.2e}", flush=True)
        worst, worst_ref = max(worst, e), max(worst_ref, e_ref)
    mf.close()
    print(f"L = {L:.6e}; worst rel err: exact adjoint {worst:.2e}, reference adjoint {worst_ref:.2e}", flush=True)
    assert worst < 1e-5
