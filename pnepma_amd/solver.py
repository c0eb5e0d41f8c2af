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
from .material import CellFractions, MassFractionMaterial
from . import physics
from .problem import (ForwardPNProblem, SolveInputs, absorption_factors, absorption_pullback, el_index,
                      solve_inputs)


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

    def __init__(self, problem: ForwardPNProblem, inputs: SolveInputs | None, ctx: Context | None = None,
                 ckpt_every: int = 0, adjoint: bool = True):
        """inputs None: the material is set later on the device (set_cell_fractions,
        then set_absorption)."""
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
        self.inputs, self.cells, self.has_absorption = None, None, False
        if inputs is not None:
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
        self.cells = None
        self.has_absorption = False
        r8 = (_lib._dp * 8)(*[dptr(np.ascontiguousarray(r)) for r in inputs.rho])
        check(lib().pnx_problem_set_material(self.handle, r8, dptr(np.ascontiguousarray(inputs.W))))

    def set_cell_fractions(self, material: MassFractionMaterial, inputs: SolveInputs | None):
        """Densities of every class computed on the device from the per-cell
        fractions of a grid-aligned box material (see cell_aligned).  The line
        weights come from inputs.W, or, with inputs None, from a following
        set_absorption."""
        if not cell_aligned(material, self.problem.gspec):
            raise ValueError("the material's boxes are not the grid's cells")
        self.inputs = inputs
        self.has_absorption = False
        vals = material.param.vals                        # (nx, ny, nz, n_e)
        self.cells = vals.shape
        w = np.ascontiguousarray(vals.transpose(2, 1, 0, 3).reshape(-1, vals.shape[3]))  # x fastest
        rho_pure = np.ascontiguousarray(material.rho_pure, dtype=np.float64)
        if inputs is not None:
            W = np.ascontiguousarray(inputs.W)
        else:
            g = self.problem.gspec
            W = np.zeros((int(np.prod(g.class_size(0))), len(self.problem.x_rays)))
        check(lib().pnx_problem_set_cell_fractions(self.handle, dptr(w), dptr(rho_pure), dptr(W)))

    def set_absorption(self, material: MassFractionMaterial, detector) -> np.ndarray:
        """Absorption factors and line weights computed on the device
        (after set_cell_fractions); returns the factors (P_eee, n_lines)."""
        p = self.problem
        g = p.gspec
        mac = np.ascontiguousarray([[physics.mass_absorption(x, e) for e in material.elements] for x in p.x_rays],
                                   dtype=np.float64)
        am = np.ascontiguousarray([e.atom_mass_g for e in material.elements], dtype=np.float64)
        el = np.ascontiguousarray(el_index(material.elements, p.x_rays), dtype=np.int32)
        det = np.ascontiguousarray(detector, dtype=np.float64)
        pts = [np.ascontiguousarray(a.points(0), dtype=np.float64) for a in g.axes]
        edges = [np.ascontiguousarray(e, dtype=np.float64) for e in material.param.edges]
        emb = np.ascontiguousarray(material.param.embedding, dtype=np.float64)
        ac = np.zeros((int(np.prod(g.class_size(0))), len(p.x_rays)))
        self._abs_keep = [mac, am, el, det, pts, edges, emb]
        check(lib().pnx_problem_set_absorption(
            self.handle, dptr(mac), dptr(am), el.ctypes.data_as(C.POINTER(C.c_int32)), dptr(det),
            float(p.beam_positions[0][0]), (_lib._dp * 3)(*[dptr(x) for x in pts]),
            (_lib._dp * 3)(*[dptr(x) for x in edges]), dptr(emb), dptr(ac)))
        self.has_absorption = True
        return ac

    def material_gradient(self) -> np.ndarray:
        """The whole material gradient on the device after forward_adjoint,
        flattened like CellFractions.params."""
        nx, ny, nz, ne = self.cells
        g = np.zeros((nx * ny * nz, ne))
        check(lib().pnx_problem_material_gradient(self.handle, dptr(g)))
        return g.reshape(nz, ny, nx, ne).transpose(2, 1, 0, 3).ravel()

    def cell_fraction_gradient(self, extra_eee=None) -> np.ndarray:
        """d(rho_bar . rho)/d(cell fractions) on the device, flattened like
        CellFractions.params (x slowest, element fastest); rho_bar is the last
        forward_adjoint's, plus extra_eee ([eee point][element]) at eee points."""
        nx, ny, nz, ne = self.cells
        g = np.zeros((nx * ny * nz, ne))
        ex = None if extra_eee is None else np.ascontiguousarray(extra_eee, dtype=np.float64)
        check(lib().pnx_problem_cell_fraction_gradient(self.handle, dptr(ex), dptr(g)))
        return g.reshape(nz, ny, nx, ne).transpose(2, 1, 0, 3).ravel()

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

    def reserve_adjoint(self, exact: bool = False):
        """Allocate the adjoint sweep's device buffers now (not in the first forward_adjoint)."""
        check(lib().pnx_problem_reserve_adjoint(self.handle, int(exact)))

    def forward_adjoint(self, I_bar, exact: bool = False, want_rho_bar: bool = True):
        """Returns (I, rho_bar[8] per class (P_c, n_e) summed over beams, gen (P_eee, n_lines)).

        exact=False runs the reference's adjoint; exact=True the exact gradient of the
        discrete forward march.  want_rho_bar=False leaves rho_bar on the device
        (for cell_fraction_gradient) and returns None in its place."""
        p = self.problem
        ib = np.ascontiguousarray(np.broadcast_to(np.asarray(I_bar, dtype=np.float64),
                                                  (p.n_beams, len(p.x_rays))))
        I = np.zeros((p.n_beams, len(p.x_rays)))
        ne = len(p.elements)
        rb = [np.zeros((int(np.prod(p.gspec.class_size(c))), ne)) for c in range(8)] if want_rho_bar else None
        gen = np.zeros((int(np.prod(p.gspec.class_size(0))), len(p.x_rays)))
        r8 = (_lib._dp * 8)(*[dptr(r) for r in rb]) if want_rho_bar else None
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

    def rev_bench(self, reps: int):
        ms, by = C.c_double(), C.c_double()
        check(lib().pnx_rev_bench(self.handle, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


def calc_intensities(problem: ForwardPNProblem, material, detector, ctx: Context | None = None) -> np.ndarray:
    """Intensities of every beam, shape (n_beams, n_lines)."""
    si = solve_inputs(problem, material, detector)
    dp = DeviceProblem(problem, si, ctx, adjoint=False)
    try:
        return dp.forward()
    finally:
        dp.close()


def cell_aligned(material, gspec) -> bool:
    """A MassFraction material with one box per grid cell (boxes centred on the
    even points, edges at the odd points): the case the device material layer
    handles."""
    from .workloads import cell_edges
    if not isinstance(material, MassFractionMaterial) or not isinstance(material.param, CellFractions):
        return False
    par = material.param
    if par.vals.shape[:3] != tuple(a.M for a in gspec.axes):
        return False
    return all(np.array_equal(e, cell_edges(a)) for e, a in zip(par.edges, gspec.axes))


def material_gradient(problem: ForwardPNProblem, material, detector, si: SolveInputs, rho_bar, gen,
                      device: "DeviceProblem | None" = None) -> np.ndarray:
    """Material pullbacks of the reference adjoint (pn_intensities.jl:406-410):
    densities at every class point, absorption factors and numbers of atoms.

    With `device` (a DeviceProblem whose material was set by set_cell_fractions
    and which has run forward_adjoint), the density and number-of-atoms terms
    are pulled back on the device and rho_bar is not needed on the host; if
    the absorption was also set on the device (set_absorption), the whole
    gradient is computed there."""
    if device is not None and getattr(device, "has_absorption", False):
        return device.material_gradient()
    g = problem.gspec
    ac_bar = si.noa[:, si.el_idx] * gen
    noa_bar = np.zeros_like(si.noa)
    np.add.at(noa_bar.T, si.el_idx, (si.ac * gen).T)
    am = np.array([e.atom_mass_g for e in material.elements])
    if device is not None:
        grad = device.cell_fraction_gradient(extra_eee=noa_bar / am)
    else:
        grad = np.zeros(material.param.n_params)
        for c in range(8):
            if np.any(rho_bar[c]):
                grad += material.densities_pullback(g.class_points(c), rho_bar[c])
        grad += material.densities_pullback(g.class_points(0), noa_bar / am)
    ac, aux = absorption_factors(material, problem.x_rays, g, detector, problem.beam_positions[0][0])
    grad += absorption_pullback(material, aux, ac, ac_bar)
    return grad


def calc_intensities_pullback(problem: ForwardPNProblem, material, detector, ckpt_every: int = 0,
                              ctx: Context | None = None):
    """Forward intensities and a pullback I_bar -> d(I_bar . I)/d(material parameters).

    The reverse sweep needs the seed, so the pullback re-runs the forward
    march with checkpoints (ckpt_every energy steps) on the device."""
    si = solve_inputs(problem, material, detector)
    dp = DeviceProblem(problem, si, ctx, ckpt_every=ckpt_every)
    on_device = cell_aligned(material, problem.gspec)
    if on_device:
        dp.set_cell_fractions(material, si)
        dp.set_absorption(material, detector)
    I = dp.forward()

    def pullback(I_bar):
        _, rb, gen = dp.forward_adjoint(I_bar, want_rho_bar=not on_device)
        return material_gradient(problem, material, detector, si, rb, gen, device=dp if on_device else None)

    pullback.device_problem = dp
    return I, pullback
