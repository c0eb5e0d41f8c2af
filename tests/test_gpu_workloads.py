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
