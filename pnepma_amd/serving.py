"""Continuous-batching serving of forward+gradient requests (host side).

``Server`` wraps the pnx_server C ABI: requests carry their own PN order, grid
and energy range; the library admits them by HBM budget, keeps their state on
the device and advances all admitted requests together, one energy step per
scheduler tick, in persistent kernels grouped by PN order and dimension.

``serving_requests`` draws the mixed request stream: PN order from {3,5,7,9};
either a 1D slab of 100-400 cells or a square 2D grid of 64^2 to 192^2 cells;
beam energy U(10, 25) keV; 2-4 elements from {Al, Si, Cr, Fe, Ni, Cu} with 2-5
piecewise-constant layers (1D) or boxes (2D); energy steps from a CFL-0.5 rule.
Frozen choices: slab depth 1.2x the Kanaya-Okayama range of the mean material
(clamped to 0.4-4 um), lateral extent equal to the depth, E_initial = E_beam +
0.5 keV, E_cutoff = 1 keV, one centred beam, lines K-L3 of every element, seed
U(0.5, 1.5) per line for the pullback, Poisson arrivals.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Request, RequestSpec, Timing, check, dptr, lib
from .grid import GridSpec
from .material import CellFractions, MassFractionMaterial
from .physics import XRay, element
from .problem import ForwardPNProblem, SolveInputs, solve_inputs
from .solver import Context, default_context, make_desc
from .workloads import KAPPA, NM, SIGMA_E, UM, detector_40deg, cell_edges

ELEMENT_POOL = ("Al", "Si", "Cr", "Fe", "Ni", "Cu")
LAMBDA_MAX = {N: float(np.polynomial.legendre.leggauss(N + 1)[0].max()) for N in (1, 3, 5, 7, 9)}


@dataclass
class Spec:
    problem: ForwardPNProblem
    material: MassFractionMaterial
    detector: np.ndarray
    I_bar: np.ndarray


def ko_range_um(E_keV, els, w):
    """Kanaya-Okayama electron range (um) of a mixture (mass-fraction averaged A, Z, density)."""
    A = sum(wi * e.atomic_mass for wi, e in zip(w, els))
    Z = sum(wi * e.Z for wi, e in zip(w, els))
    rho = 1.0 / sum(wi / e.density for wi, e in zip(w, els))
    return 0.0276 * A * E_keV ** 1.67 / (Z ** 0.889 * rho)


def _cfl_x(N, gspec, material, elements, E0, E1):
    """(E0 - E1) lambda_max sqrt(sum 1/h^2) / S_min: the stability number of a
    one-step march over the whole energy range."""
    from .physics import StoppingPower
    h = [a.h for a in gspec.axes if not a.singleton]
    inv = math.sqrt(sum(1.0 / x ** 2 for x in h))
    par = material.param
    w = par.vals.reshape(-1, len(elements)) if hasattr(par, "vals") else np.asarray(par.w).reshape(1, -1)
    rho = w / (w / material.rho_pure).sum(axis=1, keepdims=True)
    sp = np.array([StoppingPower(e)(E0) for e in elements])
    smin = float((rho @ sp).min())
    return (E0 - E1) * LAMBDA_MAX[N] * inv / smin


def cfl_number(N, gspec, material, elements, E0, E1, n_steps):
    """nu = dE lambda_max sqrt(sum 1/h^2) / S_min of an energy march: the explicit
    stability number that the reference prints but never enforces
    (pn_solver/pn_intensities.jl:2-8).  S_min is the smallest rho . sp over the
    material at the highest energy.  The march is stable for nu below about 1;
    the parity tests and workloads here run at nu <= 0.5."""
    return _cfl_x(N, gspec, material, elements, E0, E1) / n_steps


def cfl_steps(N, gspec, material, elements, E0, E1, cfl=0.5, floor=8):
    """Smallest step count (at least `floor`) with nu <= cfl (see cfl_number)."""
    return max(floor, int(math.ceil(_cfl_x(N, gspec, material, elements, E0, E1) / cfl)))


def draw_spec(rng: np.random.Generator, sizes_1d=(100, 400), sizes_2d=(64, 192), max_steps=None) -> Spec:
    N = int(rng.choice([3, 5, 7, 9]))
    n_e = int(rng.integers(2, 5))
    syms = list(rng.choice(ELEMENT_POOL, n_e, replace=False))
    els = [element(s) for s in syms]
    E_beam = float(rng.uniform(10e3, 25e3))
    is2d = bool(rng.random() < 0.5)
    n_parts = int(rng.integers(2, 6))
    fr = rng.uniform(0.05, 1.0, (n_parts, n_e))
    fr /= fr.sum(axis=1, keepdims=True)
    depth = float(np.clip(1.2 * ko_range_um(E_beam / 1e3, els, fr.mean(axis=0)), 0.4, 4.0)) * UM
    if is2d:
        side = int(rng.integers(sizes_2d[0], sizes_2d[1] + 1))
        g = GridSpec.make((side, side, 1), (-depth, 0.0), (-depth / 2, depth / 2), 0.0)
    else:
        side = int(rng.integers(sizes_1d[0], sizes_1d[1] + 1))
        g = GridSpec.make((side, 1, 1), (-depth, 0.0), 0.0, 0.0)
    Mx, My = g.axes[0].M, g.axes[1].M
    vals = np.empty((Mx, My, 1, n_e))
    if is2d:
        vals[...] = fr[0]
        for b in range(1, n_parts):  # random boxes, later ones win
            x0, x1 = sorted(rng.integers(0, Mx + 1, 2))
            y0, y1 = sorted(rng.integers(0, My + 1, 2))
            vals[x0:max(x1, x0 + 1), y0:max(y1, y0 + 1), 0] = fr[b]
    else:
        cuts = np.sort(rng.integers(1, Mx, n_parts - 1))
        edges = np.concatenate([[0], cuts, [Mx]])
        for b in range(n_parts):
            vals[edges[b]:edges[b + 1], 0, 0] = fr[b]
    mat = MassFractionMaterial(els, CellFractions(vals, *[cell_edges(a) for a in g.axes], embedding=fr[0]))
    E0, E1 = E_beam + 500.0, 1000.0
    n_steps = cfl_steps(N, g, mat, els, E0, E1)
    if max_steps is not None and n_steps > max_steps:  # small test streams: shorter energy range
        E1 = E0 - (E0 - E1) * max_steps / n_steps
        n_steps = max_steps
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(E_beam, SIGMA_E),
                         beam_positions=[(0.0, 0.0, 0.0)], beam_sigma=50 * NM,
                         beam_direction=((-1.0, 0.0, 0.0), KAPPA), E_initial=E0, E_cutoff=E1, n_steps=n_steps)
    p.x_rays = [XRay(s) for s in syms]
    return Spec(p, mat, detector_40deg(), rng.uniform(0.5, 1.5, n_e))


def serving_requests(n: int, seed: int = 3, **kw):
    rng = np.random.default_rng(seed)
    return [draw_spec(rng, **kw) for _ in range(n)]


class _Prepared:
    """Host arrays of one request, kept alive while the library refers to them."""

    def __init__(self, spec: Spec, si: SolveInputs, want_rho_bar: bool):
        p = spec.problem
        self.spec = spec
        self.si = si
        self.fwd, self.adj, self.dens = p.records()
        self.mms = p.beam_moments()
        self.ppos = p.beam_profiles(0)
        self.rho = [np.ascontiguousarray(r) for r in si.rho]
        self.W = np.ascontiguousarray(si.W)
        self.I_bar = np.ascontiguousarray(spec.I_bar, dtype=np.float64)
        self.I_out = np.zeros(len(p.x_rays))
        self.rho_bar, self.gen = [None] * 8, None
        if want_rho_bar:
            self.want_gradients()

    def want_gradients(self):
        """Host arrays that receive the request's density gradients and
        generation sums at completion (before the request is submitted)."""
        p = self.spec.problem
        ne = len(p.elements)
        self.rho_bar = [np.zeros((int(np.prod(p.gspec.class_size(c))), ne)) for c in range(8)]
        self.gen = np.zeros((int(np.prod(p.gspec.class_size(0))), len(p.x_rays)))


def prepare(spec: Spec, want_rho_bar: bool = False) -> _Prepared:
    return _Prepared(spec, solve_inputs(spec.problem, spec.material, spec.detector), want_rho_bar)


class SpecRequest:
    """A request as the library takes it with pnx_server_submit_spec: geometry,
    element tables, cell fractions and absorption inputs (no host-side
    densities, absorption or records: the library builds those), plus the
    output arrays.  Only array views and small tables are made here."""

    def __init__(self, spec: Spec, want_gradients: bool = False, want_material_gradient: bool = True):
        from .physics import mass_absorption
        from .problem import el_index
        p, mat = spec.problem, spec.material
        par = mat.param
        if not hasattr(par, "vals") or par.vals.shape[:3] != tuple(a.M for a in p.gspec.axes):
            raise ValueError("a spec request needs one box per grid cell")
        self.spec = spec
        ne = len(mat.elements)
        self.ph, self._keep = p.physics_tables()
        self.cell_w = np.ascontiguousarray(par.vals.transpose(2, 1, 0, 3).reshape(-1, ne))
        self.rho_pure = np.ascontiguousarray(mat.rho_pure, dtype=np.float64)
        self.embedding = np.ascontiguousarray(par.embedding, dtype=np.float64)
        self.mac = np.ascontiguousarray([[mass_absorption(x, e) for e in mat.elements] for x in p.x_rays],
                                        dtype=np.float64)
        self.atom_mass = np.ascontiguousarray([e.atom_mass_g for e in mat.elements], dtype=np.float64)
        self.line_element = np.ascontiguousarray(el_index(mat.elements, p.x_rays), dtype=np.int32)
        self.I_bar = np.ascontiguousarray(spec.I_bar, dtype=np.float64)
        self.I_out = np.zeros(len(p.x_rays))
        self.rho_bar, self.gen = [None] * 8, None
        if want_gradients:
            self.rho_bar = [np.zeros((int(np.prod(p.gspec.class_size(c))), ne)) for c in range(8)]
            self.gen = np.zeros((int(np.prod(p.gspec.class_size(0))), len(p.x_rays)))
        self.grad = np.zeros(self.cell_w.shape) if want_material_gradient else None
        r = RequestSpec()
        r.desc = make_desc(p, 0)
        r.physics = C.pointer(self.ph)
        r.cell_w, r.rho_pure, r.embedding = dptr(self.cell_w), dptr(self.rho_pure), dptr(self.embedding)
        r.mac, r.atom_mass = dptr(self.mac), dptr(self.atom_mass)
        r.line_element = self.line_element.ctypes.data_as(C.POINTER(C.c_int32))
        for a in range(3):
            r.detector[a] = float(spec.detector[a])
            r.beam_position[a] = float(p.beam_positions[0][a])
        direction, kappa = p.beam_direction
        d = np.asarray(direction, dtype=np.float64)
        for a in range(3):
            r.beam_direction[a] = float(d[a])
        r.beam_kappa = float(kappa)
        r.beam_sigma = float(p.beam_sigma)
        r.literal_source = int(p.source_mode == "literal")
        r.I_bar, r.I_out = dptr(self.I_bar), dptr(self.I_out)
        for c in range(8):
            r.rho_bar[c] = dptr(self.rho_bar[c])
        r.gen_out, r.grad_out = dptr(self.gen), dptr(self.grad)
        self.req = r

    def material_gradient(self) -> np.ndarray:
        """The served material gradient, flattened like CellFractions.params."""
        nx, ny, nz = (a.M for a in self.spec.problem.gspec.axes)
        ne = self.cell_w.shape[1]
        return self.grad.reshape(nz, ny, nx, ne).transpose(2, 1, 0, 3).ravel()


class Server:
    def __init__(self, hbm_budget_bytes: int, max_active: int = 4096, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().pnx_server_create(self.ctx.handle, int(hbm_budget_bytes), int(max_active), C.byref(h)))
        self.handle = h
        self._live = {}

    def clock(self) -> float:
        return lib().pnx_server_clock(self.handle)

    def set_tape_limit(self, nbytes: int):
        """Admit requests taped (every forward state kept, no recompute) while the
        arena's bytes in use stay within nbytes; 0 disables taping."""
        check(lib().pnx_server_set_tape_limit(self.handle, int(nbytes)))

    def submit(self, prep: _Prepared, arrival_s: float = 0.0) -> int:
        p = prep.spec.problem
        r = Request()
        r.desc = make_desc(p, 0)
        r.fwd, r.adj, r.dens = dptr(prep.fwd), dptr(prep.adj), dptr(prep.dens)
        for c in range(8):
            r.rho[c] = dptr(prep.rho[c])
            r.mms[c] = dptr(prep.mms[c])
            r.ppos[c] = dptr(prep.ppos[c])
            r.rho_bar[c] = dptr(prep.rho_bar[c])
        r.W = dptr(prep.W)
        r.I_bar = dptr(prep.I_bar)
        r.arrival_s = float(arrival_s)
        r.I_out = dptr(prep.I_out)
        r.gen_out = dptr(prep.gen)
        t = C.c_int64()
        check(lib().pnx_server_submit(self.handle, C.byref(r), C.byref(t)))
        self._live[t.value] = prep
        return t.value

    def submit_spec(self, req: SpecRequest, arrival_s: float = 0.0) -> int:
        """Submit a request given as cell fractions; the library prepares it in C++."""
        req.req.arrival_s = float(arrival_s)
        t = C.c_int64()
        check(lib().pnx_server_submit_spec(self.handle, C.byref(req.req), C.byref(t)))
        self._live[t.value] = req
        return t.value

    def wait(self, ticket: int) -> Timing:
        """Blocks until the request is complete and releases its ticket."""
        tm = Timing()
        check(lib().pnx_server_wait(self.handle, ticket, C.byref(tm)))
        self._live.pop(ticket, None)
        return tm

    def cancel(self, ticket: int):
        """Cancels a pending or admitted request; its outputs are not written and
        its Timing.status becomes PNX_ERR_CANCELLED."""
        check(lib().pnx_server_cancel(self.handle, ticket))

    def poll(self, ticket: int):
        d, tm = C.c_int32(), Timing()
        check(lib().pnx_server_poll(self.handle, ticket, C.byref(d), C.byref(tm)))
        return bool(d.value), tm

    def drain(self):
        check(lib().pnx_server_drain(self.handle))

    def stats(self):
        v = [C.c_int64() for _ in range(4)]
        check(lib().pnx_server_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("ticks", "completed", "peak_active", "peak_bytes"), (x.value for x in v)))

    def close(self):
        if getattr(self, "handle", None):
            lib().pnx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_device_bytes(problem: ForwardPNProblem) -> int:
    d = make_desc(problem, 0)
    return int(check(lib().pnx_request_device_bytes(C.byref(d))))
