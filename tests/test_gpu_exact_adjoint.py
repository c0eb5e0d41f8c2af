"""GPU exact adjoint (pnx_forward_adjoint_exact): the gradient of the discrete
forward march.  Checked against the dense numpy transpose of the march
(tests/numpy_step.py, 1e-8) and against central finite differences of the
forward (1e-5), directional and per entry."""
import copy

import numpy as np
import pytest

from pnepma_amd import grid as G
from pnepma_amd.solver import DeviceProblem

from . import numpy_step
from .helpers import inputs, rel_err, small_1d, small_2d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["1d", "2d"])
@pytest.mark.parametrize("ckpt", [0, 3])
def test_exact_adjoint_vs_dense(oracle_lib, case, ckpt):
    if case == "1d":
        p, mat, _, det = small_1d(N=5, cells=16, E0=15e3, E1=10e3)
    else:
        p, mat, _, det = small_2d(N=3, M=(8, 6), n_beams=1, E1=12e3)
    si = inputs(p, mat, det)
    ib = np.linspace(1.0, -0.5, len(p.x_rays))
    I, rb, gen = numpy_step.discrete_pullback(p, si, oracle_lib.advection_dense, oracle_lib.boundary_dense, ib)
    dp = DeviceProblem(p, si, ckpt_every=ckpt)
    Ig, rbg, geng = dp.forward_adjoint(ib[None, :], exact=True)
    dp.close()
    assert rel_err(Ig[0], I) < 1e-10
    for c in p.gspec.active_classes():
        e = rel_err(rbg[c], rb[c])
        print(f"{case} rho_bar {G.CLASS_NAMES[c]} rel err {e:.3e}", flush=True)
        assert e < 1e-8
    assert rel_err(geng, gen) < 1e-8


def test_exact_adjoint_vs_central_differences(oracle_lib):
    p, mat, _, det = small_2d(N=3, M=(10, 8), n_beams=1, E1=10e3)
    si = inputs(p, mat, det)
    ib = np.array([1.0, -0.6, 0.4])
    dp = DeviceProblem(p, si)
    I, rb, gen = dp.forward_adjoint(ib[None, :], exact=True)
    dp.close()
    rng = np.random.default_rng(4)

    def J(rho):
        s = copy.copy(si)
        s.rho = rho
        return oracle_lib.forward(p, s) @ ib

    # directional derivative along a random perturbation of every density
    dirs = [rng.standard_normal(r.shape) * r for r in si.rho]
    h = 1e-6
    Jp = J([r + h * d for r, d in zip(si.rho, dirs)])
    Jm = J([r - h * d for r, d in zip(si.rho, dirs)])
    fd = (Jp - Jm) / (2 * h)
    ad = sum(float((rb[c] * dirs[c]).sum()) for c in range(8))
    print("directional", ad, fd, flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    # line weights W = ac * noa[el] enter linearly: perturb the absorption factors
    dac = rng.standard_normal(si.ac.shape) * si.ac * 1e-3
    s2 = copy.copy(si)
    s2.ac = si.ac + dac
    fdW = oracle_lib.forward(p, s2) @ ib - oracle_lib.forward(p, si) @ ib
    dW = dac * si.noa[:, si.el_idx]
    assert abs(float((gen * dW).sum()) - fdW) <= 1e-8 * abs(fdW)
