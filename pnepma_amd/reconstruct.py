"""k-ratio reconstruction driver: per-cell mass fractions from measured k-ratios.

Restates the reference scripts' driver (scripts/reconstructions_1d.jl:52-73
and 149-184, scripts/2d_full_reconstruction_test.jl:158-234):

    k_b,r(m) = I_b,r(m) / I_std,b,r       (measure: intensities over standards)
    L(m)     = sum_b,r (k_b,r(m) - k_exp_b,r)^2
    minimise L over the box fractions in [0, 1] with a bounded quasi-Newton method
    (the reference: Optim's Fminbox(BFGS()); here scipy's L-BFGS-B).

Each evaluation runs on the device: the material (densities, absorption factors
and line weights) is set from the fractions, the forward pass gives I, the
seed I_bar = 2 (k - k_exp) / I_std goes into the adjoint sweep, and the
material pullback to the fractions is computed on the device.

The default gradient is the exact gradient of the discrete march
(pnx_forward_adjoint_exact).  The reference's own adjoint (adjoint="reference")
is a continuous adjoint that differs from the derivative of the discrete
objective by 1-30 % on these grids, which misleads the line search.  Fractions
only enter through w / sum(w / rho_pure), so their scale per box is free; the
bounds keep them non-negative.
"""
from __future__ import annotations

import numpy as np

from .material import MassFractionMaterial
from .problem import ForwardPNProblem, solve_inputs
from .solver import Context, DeviceProblem, cell_aligned


def standard_intensities(problem: ForwardPNProblem, standards, detector, ctx: Context | None = None) -> np.ndarray:
    """I_std[b, r]: intensity of line r from the standard of its element, per beam
    (std_ints of the reference scripts).  standards[r] is the standard of line r."""
    out = np.zeros((problem.n_beams, len(problem.x_rays)))
    for r, std in enumerate(standards):
        dp = DeviceProblem(problem, solve_inputs(problem, std, detector), ctx, adjoint=False)
        out[:, r] = dp.forward()[:, r]
        dp.close()
    return out


class KRatioMisfit:
    """L(x) and dL/dx for the box fractions x (flattened like CellFractions.params)."""

    def __init__(self, problem: ForwardPNProblem, material: MassFractionMaterial, detector, I_std, k_exp,
                 ckpt_every: int = 16, adjoint: str = "exact", ctx: Context | None = None):
        if not cell_aligned(material, problem.gspec):
            raise ValueError("the reconstruction needs one box per grid cell")
        if adjoint not in ("exact", "reference"):
            raise ValueError("adjoint must be 'exact' or 'reference'")
        self.problem = problem
        self.template = material
        self.detector = np.asarray(detector, dtype=np.float64)
        self.I_std = np.asarray(I_std, dtype=np.float64)
        self.k_exp = np.asarray(k_exp, dtype=np.float64)
        self.exact = adjoint == "exact"
        self.dp = DeviceProblem(problem, solve_inputs(problem, material, detector), ctx, ckpt_every=ckpt_every)
        self.evaluations = 0

    def _set(self, x):
        m = self.template.with_params(np.asarray(x, dtype=np.float64))
        self.dp.set_cell_fractions(m, None)
        self.dp.set_absorption(m, self.detector)
        return m

    def kratios(self, x) -> np.ndarray:
        self._set(x)
        return self.dp.forward() / self.I_std

    def loss(self, x) -> float:
        k = self.kratios(x)
        return float(np.sum((k - self.k_exp) ** 2))

    def loss_and_grad(self, x):
        self._set(x)
        I = self.dp.forward()
        k = I / self.I_std
        r = k - self.k_exp
        ib = 2.0 * r / self.I_std
        self.dp.forward_adjoint(ib, exact=self.exact, want_rho_bar=False)
        g = self.dp.material_gradient()
        self.evaluations += 1
        return float(np.sum(r ** 2)), g

    def close(self):
        self.dp.close()


def reconstruct(misfit: KRatioMisfit, x0, max_iter: int = 100, tol: float = 1e-12, callback=None):
    """Bounded L-BFGS over the fractions in [0, 1]; returns scipy's OptimizeResult."""
    from scipy.optimize import minimize

    x0 = np.clip(np.asarray(x0, dtype=np.float64), 0.0, 1.0)
    return minimize(misfit.loss_and_grad, x0, jac=True, method="L-BFGS-B", bounds=[(0.0, 1.0)] * x0.size,
                    callback=callback, options={"maxiter": max_iter, "ftol": tol, "gtol": 1e-14})
