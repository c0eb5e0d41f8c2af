"""The 3D workload: P5 Al/Si with a seeded ellipsoidal inclusion, scanned by an
8 x 8 beam raster, forward + reference adjoint with checkpoints every 32 steps.

* Reduced size against the oracle: 24 x 24 x 16 cells over the workload's
  2 x 2 x 1.5 um, the whole 25 -> 1 keV range at the stable step count
  (nu <= 0.5, 419 steps), four beams of the raster in one batch.  Intensities
  per beam to 1e-10; density gradients of all eight classes (summed over the
  beams) and the generation sums to 1e-8.
* Full shape (128 x 128 x 96), the first 64 energy steps at the stable step
  length, material and absorption set on the device: size-independent
  properties -- finite intensities of one sign, results independent of the
  checkpoint interval (16 vs 32) to rounding, and a beam solved alone equal to
  the same beam inside the batch.
"""
import copy
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from pnepma_amd import grid as G
from pnepma_amd import serving, workloads
from pnepma_amd.problem import solve_inputs
from pnepma_amd.solver import DeviceProblem

from .helpers import assert_physical, cfl_of, rel_err, stable_steps

pytestmark = pytest.mark.gpu

RASTER_BEAMS = (18, 27, 36, 45)  # a diagonal of the 8 x 8 raster


def _beams(p, idx):
    q = copy.copy(p)
    q._cache = {}
    q.beam_positions = [p.beam_positions[b] for b in idx]
    return q


def test_inclusion3d_reduced_vs_oracle(oracle_lib):
    M = (24, 24, 16)
    p0, mat, _, det = workloads.inclusion3d(M=M, n_steps=8)
    p, mat, _, det = workloads.inclusion3d(M=M, n_steps=stable_steps(p0, mat))
    q = _beams(p, RASTER_BEAMS)
    nu = cfl_of(q, mat)
    print(f"3D P5 {M}: {q.n_steps} steps, nu {nu:.3f}, {q.n_beams} beams, checkpoints every 32", flush=True)
    assert nu <= 0.5
    si = solve_inputs(q, mat, det)
    ib = np.array([[1.0, 0.6], [0.8, -0.4], [0.5, 1.2], [1.1, 0.3]])
    dp = DeviceProblem(q, si, ckpt_every=32)
    I, rb, gen = dp.forward_adjoint(ib)
    dp.close()
    q.records()
    with ThreadPoolExecutor(len(RASTER_BEAMS)) as ex:  # the oracle releases the GIL inside its C calls
        refs = list(ex.map(lambda b: oracle_lib.forward_adjoint(q, si, ib[b], beam=b), range(q.n_beams)))
    rb_ref = [sum(r[1][c] for r in refs) for c in range(8)]
    acb_ref = sum(r[2] for r in refs)
    for b, r in enumerate(refs):
        assert_physical(r[0])
        e = rel_err(I[b], r[0])
        print(f"beam {RASTER_BEAMS[b]}: I {I[b]} oracle {r[0]} rel err {e:.3e}", flush=True)
        assert e < 1e-10
    for c in q.gspec.active_classes():
        e = rel_err(rb[c], rb_ref[c])
        print(f"rho_bar {G.CLASS_NAMES[c]} rel err {e:.3e}", flush=True)
        assert e < 1e-8
    e = rel_err(si.noa[:, si.el_idx] * gen, acb_ref)
    print(f"ac_bar rel err {e:.3e}", flush=True)
    assert e < 1e-8


def _full_shape(n_beams_idx, S=64):
    full, mat, _, det = workloads.inclusion3d(n_steps=16, n_beams=64)
    n_cfl = serving.cfl_steps(full.N, full.gspec, mat, full.elements, full.E_initial, full.E_cutoff)
    E1 = full.E_initial - (full.E_initial - full.E_cutoff) * S / n_cfl
    p, mat, _, det = workloads.inclusion3d(n_steps=S, n_beams=64, E0=full.E_initial, E1=E1)
    return _beams(p, n_beams_idx), mat, det, n_cfl


def _device(q, mat, det, ckpt):
    dp = DeviceProblem(q, None, ckpt_every=ckpt)
    dp.set_cell_fractions(mat, None)
    dp.set_absorption(mat, det)
    return dp


def test_inclusion3d_full_shape_properties():
    q, mat, det, n_cfl = _full_shape((27, 36))
    print(f"3D P5 128x128x96: first {q.n_steps} of {n_cfl} stable steps, nu {cfl_of(q, mat):.3f}", flush=True)
    assert cfl_of(q, mat) <= 0.5
    ib = np.array([[1.0, 0.6], [0.8, -0.4]])
    out = {}
    for ck in (32, 16):
        dp = _device(q, mat, det, ck)
        out[ck] = dp.forward_adjoint(ib)
        dp.close()
    I32, rb32, g32 = out[32]
    I16, rb16, g16 = out[16]
    for b in range(q.n_beams):
        assert_physical(I32[b])
    e_I = rel_err(I16, I32)
    e_rb = max(rel_err(rb16[c], rb32[c]) for c in q.gspec.active_classes())
    e_g = rel_err(g16, g32)
    print(f"ckpt 16 vs 32: I {e_I:.3e} rho_bar {e_rb:.3e} gen {e_g:.3e}", flush=True)
    assert e_I < 1e-13 and e_rb < 1e-12 and e_g < 1e-13
    assert all(np.all(np.isfinite(r)) for r in rb32)
    # the second beam alone equals the same beam in the batch
    one = _beams(q, (1,))
    dp = _device(one, mat, det, 32)
    I1 = dp.forward()[0]
    dp.close()
    e = rel_err(I1, I32[1])
    print(f"beam alone vs in batch: {e:.3e}", flush=True)
    assert e < 1e-13
