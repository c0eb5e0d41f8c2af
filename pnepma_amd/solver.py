"""Host-side mirror of the reference operator API for the k-ratio hot path.

``calc_intensities(problem, material, detector)`` replaces
pn_solver/pn_intensities.jl:325 and ``calc_intensities_pullback`` replaces its
Zygote adjoint (pn_solver/pn_intensities.jl:353-413): it returns the
intensities and a pullback closure mapping the seed I_bar to the gradient with
respect to the material parameters (the ``to_named_tuple(m_)`` slot of the
reference pullback).  Both run on the GPU through libpnx.so; nothing here
computes the solve on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Desc, check, dptr, lib
from .problem import ForwardPNProblem, SolveInputs, absorption_factors, absorption_pullback, solve_inputs


class Context:
    """One device and one HIP stream (pnx_context)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().pnx_context_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pnx_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def make_desc(problem: ForwardPNProblem, ckpt_every: int = 0) -> Desc:
    g = problem.gspec
    d = Desc()
    d.N = problem.N
    for a in range(3):
        d.M[a] = g.axes[a].M
        d.lo[a] = g.axes[a].lo
        d.hi[a] = g.axes[a].hi
    d.n_elements = len(problem.elements)
    d.n_lines = len(problem.x_rays)
    d.n_steps = problem.n_steps
    d.n_beams = problem.n_beams
    d.E_initial = problem.E_initial
    d.E_cutoff = problem.E_cutoff
    d.ckpt_every = ckpt_every
    return d


class DeviceProblem:
    """A batch of beams resident on the device (pnx_problem)."""

    def __init__(self, problem: ForwardPNProblem, inputs: SolveInputs, ctx: Context | None = None,
                 ckpt_every: int = 0, adjoint: bool = True):
        self.problem = problem
        self.ctx = ctx or default_context()
        self.desc = make_desc(problem, ckpt_every)
        h = C.c_void_p()
        check(lib().pnx_problem_create(self.ctx.handle, C.byref(self.desc), C.byref(h)))
        self.handle = h
        fwd, adj, dens = problem.records()
        self._keep = [fwd, adj, dens]
        if adjoint:
            check(lib().pnx_problem_set_records(h, dptr(fwd), dptr(adj), dptr(dens)))
        else:
            check(lib().pnx_problem_set_records(h, dptr(fwd), None, None))
        self.set_material(inputs)
        mms = problem.beam_moments()
        m8 = (_lib._dp * 8)(*[dptr(m) for m in mms])
        profs = [problem.beam_profiles(b) for b in range(problem.n_beams)]
        flat = []
        keep = []
        for b in range(problem.n_beams):
            for c in range(8):
                a = profs[b][c]
                keep.append(a)
                flat.append(dptr(a))
        arr = (_lib._dp * len(flat))(*flat)
        check(lib().pnx_problem_set_beams(h, m8, arr))
        self._keep += [mms, keep]

    def set_material(self, inputs: SolveInputs):
        self.inputs = inputs
        r8 = (_lib._dp * 8)(*[dptr(np.ascontiguousarray(r)) for r in inputs.rho])
        check(lib().pnx_problem_set_material(self.handle, r8, dptr(np.ascontiguousarray(inputs.W))))

    def close(self):
        if getattr(self, "handle", None):
            lib().pnx_problem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(lib().pnx_problem_device_bytes(self.handle))

    def forward(self) -> np.ndarray:
        out = np.zeros((self.problem.n_beams, len(self.problem.x_rays)))
        check(lib().pnx_forward(self.handle, dptr(out)))
        return out

    def forward_adjoint(self, I_bar, exact: bool = False):
        """Returns (I, rho_bar[8] per class (P_c, n_e) summed over beams, gen (P_eee, n_lines)).

        exact=False runs the reference's adjoint; exact=True the exact gradient of the
        discrete forward march."""
        p = self.problem
        ib = np.ascontiguousarray(np.broadcast_to(np.asarray(I_bar, dtype=np.float64),
                                                  (p.n_beams, len(p.x_rays))))
        I = np.zeros((p.n_beams, len(p.x_rays)))
        ne = len(p.elements)
        rb = [np.zeros((int(np.prod(p.gspec.class_size(c))), ne)) for c in range(8)]
        gen = np.zeros((int(np.prod(p.gspec.class_size(0))), len(p.x_rays)))
        r8 = (_lib._dp * 8)(*[dptr(r) for r in rb])
        fn = lib().pnx_forward_adjoint_exact if exact else lib().pnx_forward_adjoint
        check(fn(self.handle, dptr(ib), dptr(I), r8, dptr(gen)))
        return I, rb, gen

    def step(self, u, record: int, beam: int = 0, adjoint: bool = False, I_bar=None):
        """One step_pde! on per-class host arrays u[c] of shape (P_c, n_c) (in place)."""
        u8 = (_lib._dp * 8)(*[dptr(x) if x is not None else None for x in u])
        ib = np.ascontiguousarray(I_bar, dtype=np.float64) if I_bar is not None else None
        check(lib().pnx_step(self.handle, int(adjoint), int(record), int(beam), dptr(ib), u8))

    def forward_timed(self, reps: int = 1, use_graph: bool = True) -> float:
        ms = C.c_double()
        check(lib().pnx_forward_timed(self.handle, reps, int(use_graph), C.byref(ms)))
        return ms.value

    def forward_adjoint_timed(self, reps: int = 1) -> float:
        ms = C.c_double()
        check(lib().pnx_forward_adjoint_timed(self.handle, reps, C.byref(ms)))
        return ms.value

    def half_step_bench(self, parity: int, reps: int):
        ms, by = C.c_double(), C.c_double()
        check(lib().pnx_half_step_bench(self.handle, parity, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


def calc_intensities(problem: ForwardPNProblem, material, detector, ctx: Context | None = None) -> np.ndarray:
    """Intensities of every beam, shape (n_beams, n_lines)."""
    si = solve_inputs(problem, material, detector)
    dp = DeviceProblem(problem, si, ctx, adjoint=False)
    try:
        return dp.forward()
    finally:
        dp.close()


def material_gradient(problem: ForwardPNProblem, material, detector, si: SolveInputs, rho_bar, gen) -> np.ndarray:
    """Material pullbacks of the reference adjoint (pn_intensities.jl:406-410):
    densities at every class point, absorption factors and numbers of atoms."""
    g = problem.gspec
    grad = np.zeros(material.param.n_params)
    for c in range(8):
        if np.any(rho_bar[c]):
            grad += material.densities_pullback(g.class_points(c), rho_bar[c])
    ac_bar = si.noa[:, si.el_idx] * gen
    noa_bar = np.zeros_like(si.noa)
    np.add.at(noa_bar.T, si.el_idx, (si.ac * gen).T)
    ac, aux = absorption_factors(material, problem.x_rays, g, detector, problem.beam_positions[0][0])
    grad += absorption_pullback(material, aux, ac, ac_bar)
    am = np.array([e.atom_mass_g for e in material.elements])
    grad += material.densities_pullback(g.class_points(0), noa_bar / am)
    return grad


def calc_intensities_pullback(problem: ForwardPNProblem, material, detector, ckpt_every: int = 0,
                              ctx: Context | None = None):
    """Forward intensities and a pullback I_bar -> d(I_bar . I)/d(material parameters).

    The reverse sweep needs the seed, so the pullback re-runs the forward
    march with checkpoints (ckpt_every energy steps) on the device."""
    si = solve_inputs(problem, material, detector)
    dp = DeviceProblem(problem, si, ctx, ckpt_every=ckpt_every)
    I = dp.forward()

    def pullback(I_bar):
        _, rb, gen = dp.forward_adjoint(I_bar)
        return material_gradient(problem, material, detector, si, rb, gen)

    pullback.device_problem = dp
    return I, pullback
