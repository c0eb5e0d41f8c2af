"""k-ratio reconstruction driver (scripts/reconstructions_1d.jl, 2d_full_reconstruction_test.jl):
the misfit gradient against central differences, and an L-BFGS run that
recovers per-cell fractions from synthetic k-ratios."""
import numpy as np
import pytest

from pnepma_amd import workloads
from pnepma_amd.grid import GridSpec
from pnepma_amd.material import CellFractions, ConstFractions, MassFractionMaterial
from pnepma_amd.physics import XRay, element
from pnepma_amd.problem import ForwardPNProblem
from pnepma_amd.reconstruct import KRatioMisfit, reconstruct, standard_intensities

pytestmark = pytest.mark.gpu


def _setup(n_beams=6, M=(14, 10, 1), seed=3):
    rng = np.random.default_rng(seed)
    names = ["Fe", "Ni"]
    els = [element(s) for s in names]
    g = GridSpec.make(M, (-1.0e-4, 0.0), (-0.5e-4, 0.5e-4), 0.0)
    ys = np.linspace(-0.4e-4, 0.4e-4, n_beams)
    p = ForwardPNProblem(N=3, gspec=g, elements=els, beam_energy=(15e3, workloads.SIGMA_E),
                         beam_positions=[(0.0, float(y), 0.0) for y in ys], beam_sigma=50 * workloads.NM,
                         beam_direction=((-1.0, 0.0, 0.0), workloads.KAPPA), E_initial=15.5e3, E_cutoff=5e3,
                         n_steps=40, source_mode="lateral")
    p.x_rays = [XRay(s) for s in names]
    edges = [workloads.cell_edges(a) for a in g.axes]
    truth_vals = workloads.inclusion_fractions(M, len(els), 3, rng)
    truth = MassFractionMaterial(els, CellFractions(truth_vals, *edges, embedding=truth_vals.reshape(-1, 2)[0]))
    stds = [MassFractionMaterial(els, ConstFractions(np.eye(len(els))[i])) for i in range(len(els))]
    det = workloads.detector_40deg()
    return p, truth, stds, det, rng


def test_misfit_gradient_matches_central_differences():
    p, truth, stds, det, rng = _setup()
    I_std = standard_intensities(p, stds, det)
    mf = KRatioMisfit(p, truth, det, I_std, np.zeros((p.n_beams, len(p.x_rays))), ckpt_every=8)
    k_truth = mf.kratios(truth.param.params())
    k_exp = k_truth * (1 + rng.normal(0, 0.01, k_truth.shape))
    mf.k_exp = k_exp
    x = np.clip(truth.param.params() + rng.normal(0, 0.05, truth.param.n_params), 0.05, 1.0)
    L, g = mf.loss_and_grad(x)
    h = 1e-6
    worst = 0.0
    for i in rng.choice(x.size, 8, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (mf.loss(xp) - mf.loss(xm)) / (2 * h)
        e = abs(fd - g[i]) / max(abs(fd), 1e-3 * np.max(np.abs(g)))
        print(f"param {i}: adjoint {g[i]:.6e} fd {fd:.6e} rel {e:.2e}", flush=True)
        worst = max(worst, e)
    mf.close()
    assert worst < 1e-5


def test_lbfgs_reduces_the_misfit():
    p, truth, stds, det, rng = _setup()
    I_std = standard_intensities(p, stds, det)
    mf = KRatioMisfit(p, truth, det, I_std, np.zeros((p.n_beams, len(p.x_rays))), ckpt_every=8)
    mf.k_exp = mf.kratios(truth.param.params())        # noise-free targets
    x0 = np.clip(truth.param.params() + rng.normal(0, 0.05, truth.param.n_params), 0.05, 1.0)
    L0 = mf.loss(x0)
    res = reconstruct(mf, x0, max_iter=40)
    print(f"L0 {L0:.3e} -> {res.fun:.3e} in {res.nit} iterations, {mf.evaluations} evaluations", flush=True)
    mf.close()
    assert res.fun < 1e-3 * L0
