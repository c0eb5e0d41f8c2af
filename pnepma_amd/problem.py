"""Problem set-up on the host: the PN-CSD model of one EPMA measurement geometry.

``ForwardPNProblem`` mirrors the reference constructor of the same name
(pn_solver/pn_problem.jl:72-106) with one extension: it carries a list of beam
positions, because beams that share geometry, material and energy grid are
solved as one batch on the device.

Everything the device needs per energy is precomputed here as an energy-point
record of n_e + n_e*(N+1) + 1 + n_lines doubles:

    s[e]        stopping power sp_e(E)                       (eV cm^2 / g)
    c[e][l]     forward: -sp_e'(E) - (tc_e[l] - tc_e[0]); adjoint: -(tc_e[l] - tc_e[0])
                (pn_solver/pn_problem.jl:157-190)
    gfac        (Phi(E + dE) - Phi(E)) / dE of the beam energy distribution
                (pn_solver/pn_problem.jl:205-211)
    sigma[r]    emission cross section of line r at the record's sigma energy

Three record tables are built, following the energy bookkeeping of
pn_solver/pn_intensities.jl:10-30 and 325-413 exactly (including the
floating-point recurrences of the energies):

    fwd[k], k = 1..n : coefficients at E_{k-1} - dE/2, sigma at E_k; fwd[0]: sigma at E_0
    adj[i], i = 1..n : adjoint coefficients and sigma at Ê_i + dE/2
    dens[i], i = 0..n: forward coefficients, gfac and sigma at Ê_i

where E_k is the forward march energy and Ê_i the reverse-sweep energy of step i.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import physics
from .grid import GridSpec, class_moments, odd_in
from .material import MassFractionMaterial


@dataclass
class ForwardPNProblem:
    N: int
    gspec: GridSpec
    elements: list                  # physics.Element, in material order
    beam_energy: tuple              # (mean, std) in eV
    beam_positions: list            # [(x, y, z)] beam centres; x must be the surface
    beam_sigma: float               # std of the isotropic Gaussian beam profile, cm
    beam_direction: tuple           # ((dx, dy, dz), kappa) von Mises-Fisher
    E_initial: float
    E_cutoff: float
    n_steps: int
    source_mode: str = "lateral"    # "literal" or "lateral" (see beam_profiles)
    _cache: dict = field(default_factory=dict, repr=False)

    # ------------------------------------------------------------ energies
    @property
    def dE(self) -> float:
        return (self.E_initial - self.E_cutoff) / self.n_steps

    def forward_energies(self):
        """(coefficient energy, sigma energy) per forward step k = 1..n."""
        E, dE, out = self.E_initial, self.dE, []
        for _ in range(self.n_steps):
            t = self.E_initial - E          # E_to_t (forward)
            Ec = E - dE / 2
            t = t + dE
            E = self.E_initial - t          # t_to_E (forward)
            out.append((Ec, E))
        return out

    def reverse_energies(self):
        """Ê_i for i = n..0 (returned indexed by i)."""
        E, dE = self.E_cutoff, self.dE
        out = [0.0] * (self.n_steps + 1)
        for i in range(self.n_steps, 0, -1):
            out[i] = E
            t = E - self.E_cutoff           # E_to_t (adjoint)
            t = t + dE
            E = self.E_cutoff + t           # t_to_E (adjoint)
        out[0] = E
        return out

    # --------------------------------------------------------- coefficient data
    @property
    def record_len(self) -> int:
        ne = len(self.elements)
        return ne + ne * (self.N + 1) + 1 + len(self.x_rays)

    x_rays: list = None  # set by the caller (list of physics.XRay)

    def _coef(self, E, adjoint: bool):
        E = np.asarray(E, dtype=np.float64)
        ne, N = len(self.elements), self.N
        s = np.empty(E.shape + (ne,))
        c = np.empty(E.shape + (ne, N + 1))
        for e, el in enumerate(self.elements):
            sp = physics.StoppingPower(el)
            s[..., e] = sp(E)
            tc = physics.transport_coefficients(el, E.ravel(), N).reshape(E.shape + (N + 1,))
            dtc = tc - tc[..., :1]
            if adjoint:
                c[..., e, :] = 0.0 - dtc
            else:
                c[..., e, :] = -sp.derivative(E)[..., None] - dtc
        return s, c

    def _gfac(self, E):
        mu, sd = self.beam_energy
        dE = self.dE
        E = np.asarray(E, dtype=np.float64)
        return (physics.normal_cdf(E + dE, mu, sd) - physics.normal_cdf(E, mu, sd)) / dE

    def _sigma(self, E):
        E = np.asarray(E, dtype=np.float64)
        return np.stack([physics.ionisation_cross_section(E, x) for x in self.x_rays], axis=-1)

    def _records(self, Ecoef, Esig, adjoint):
        s, c = self._coef(Ecoef, adjoint)
        g = self._gfac(Ecoef) if not adjoint else np.zeros(np.shape(Ecoef))
        sig = self._sigma(Esig)
        n = len(Ecoef)
        return np.concatenate([s.reshape(n, -1), c.reshape(n, -1), g.reshape(n, 1), sig.reshape(n, -1)], axis=1)

    def physics_tables(self):
        """The pnx_physics tables of this problem's elements and lines (and the
        arrays they point into)."""
        from ._lib import Physics, dptr
        ph = Physics()
        knots = None
        keep = []
        for e, el in enumerate(self.elements):
            kn, sp = physics.stopping_power_table(el)
            tE, tc = physics.transport_table(el)
            knots = np.ascontiguousarray(kn, dtype=np.float64) if knots is None else knots
            sp = np.ascontiguousarray(sp, dtype=np.float64)
            tc = np.ascontiguousarray(tc, dtype=np.float64)
            keep += [sp, tc]
            ph.sp[e] = dptr(sp)
            ph.tc[e] = dptr(tc)
            ph.n_tc = tc.shape[0]
            ph.tc_stride = tc.shape[1]
            tE = np.ascontiguousarray(tE, dtype=np.float64)
            keep.append(tE)
            ph.tc_E = dptr(tE)
        keep.append(knots)
        ph.n_knots = knots.size
        ph.sp_E = dptr(knots)
        for r, x in enumerate(self.x_rays):
            ph.line_edge[r] = x.edge
        ph.beam_E_mean, ph.beam_E_sd = self.beam_energy
        return ph, keep

    def records(self):
        """(fwd, adj, dens) record tables, each (n_steps + 1, record_len), C-contiguous,
        built by the library's C++ set-up (pnx_problem_records)."""
        if "records" in self._cache:
            return self._cache["records"]
        import ctypes as C
        from ._lib import check, dptr, lib
        from .solver import make_desc
        ph, keep = self.physics_tables()
        out = tuple(np.zeros((self.n_steps + 1, self.record_len)) for _ in range(3))
        check(lib().pnx_problem_records(C.byref(make_desc(self)), C.byref(ph), *[dptr(a) for a in out]))
        self._cache["records"] = out
        return out

    def records_host(self):
        """The same record tables computed in numpy (kept as the package's
        cross-check of the C++ set-up)."""
        fe = self.forward_energies()
        Ec = np.array([self.E_initial] + [a for a, _ in fe])
        Es = np.array([self.E_initial] + [b for _, b in fe])
        fwd = self._records(Ec, Es, adjoint=False)
        Er = np.array(self.reverse_energies())
        adj = self._records(Er + self.dE / 2, Er + self.dE / 2, adjoint=True)
        dens = self._records(Er, Er, adjoint=False)
        return tuple(np.ascontiguousarray(a) for a in (fwd, adj, dens))

    # ----------------------------------------------------------------- beam
    def beam_moments(self):
        """mms per x-odd class (pn_solver/pn_problem.jl:192-203), None for x-even
        classes, from the library's C++ set-up (pnx_beam_moments)."""
        import ctypes as C
        from ._lib import _dp, check, dptr, lib
        cm = class_moments(self.N)
        direction, kappa = self.beam_direction
        out = [np.zeros(len(cm[c])) if odd_in(c, 0) else None for c in range(8)]
        d = np.ascontiguousarray(direction, dtype=np.float64)
        check(lib().pnx_beam_moments(self.N, dptr(d), float(kappa), (_dp * 8)(*[dptr(x) for x in out])))
        return out

    def beam_moments_host(self):
        """The same moments from the numpy Funk-Hecke closed form."""
        cm = class_moments(self.N)
        direction, kappa = self.beam_direction
        return [physics.vmf_beam_moments(cm[c], direction, kappa) if odd_in(c, 0) else None for c in range(8)]

    def beam_profiles(self, beam: int):
        """Position pdf on the (y, z) points of each x-odd class (y fastest), from
        the library's C++ set-up (pnx_beam_profiles); see beam_profiles_host."""
        import ctypes as C
        from ._lib import _dp, check, dptr, lib
        from .solver import make_desc
        g = self.gspec
        out = [np.zeros(g.class_size(c)[1] * g.class_size(c)[2]) if odd_in(c, 0) else None for c in range(8)]
        xb = np.ascontiguousarray(self.beam_positions[beam], dtype=np.float64)
        check(lib().pnx_beam_profiles(C.byref(make_desc(self)), dptr(xb), float(self.beam_sigma),
                                      int(self.source_mode == "literal"), (_dp * 8)(*[dptr(x) for x in out])))
        return out

    def beam_profiles_host(self, beam: int):
        """Position pdf on the (y, z) points of each x-odd class (y fastest).

        The reference evaluates the 3D isotropic normal at x = x_min, the first
        even x point (pn_solver/pn_problem.jl:197-200).  In "literal" mode that
        is reproduced; the factor exp(-(x_min - x_b)^2 / (2 sigma^2)) is the same
        for every point and underflows to zero for deep slabs (x_min/sigma > 38.6).
        "lateral" mode drops that global factor; k-ratios are unchanged because
        sample and standard share the geometry.
        """
        g = self.gspec
        xb = np.asarray(self.beam_positions[beam], dtype=np.float64)
        sd = self.beam_sigma
        norm = (2.0 * math.pi * sd * sd) ** -1.5
        out = [None] * 8
        x0 = g.axes[0].points(0)[0]
        for c in range(8):
            if not odd_in(c, 0):
                continue
            py = g.axes[1].points(odd_in(c, 1))
            pz = g.axes[2].points(odd_in(c, 2))
            Z, Y = np.meshgrid(pz, py, indexing="ij")
            r2 = (Y - xb[1]) ** 2 + (Z - xb[2]) ** 2
            if self.source_mode == "literal":
                r2 = r2 + (x0 - xb[0]) ** 2
            out[c] = np.ascontiguousarray((norm * np.exp(-r2 / (2.0 * sd * sd))).ravel())
        return out

    @property
    def n_beams(self) -> int:
        return len(self.beam_positions)


@dataclass
class SolveInputs:
    """Material-dependent inputs of one solve batch (all beams share them)."""
    rho: list        # 8 arrays (P_c, n_e), class points x fastest
    noa: np.ndarray  # (P_eee, n_e) atoms / cm^3
    ac: np.ndarray   # (P_eee, n_lines) absorption factors
    el_idx: np.ndarray  # (n_lines,) int32
    W: np.ndarray    # (P_eee, n_lines) ac * noa[el_idx]


def el_index(elements, x_rays):
    syms = [e.symbol for e in elements]
    return np.array([syms.index(x.symbol) for x in x_rays], dtype=np.int32)


def absorption_factors(material: MassFractionMaterial, x_rays, gspec: GridSpec, detector, beam_x0: float):
    """compute_absorption_coefficients (pn_solver/x_ray_intensities.jl:100-128): a
    100-point march from each eee point towards the detector, cut where the ray
    crosses the plane x = beam_x0."""
    pts = gspec.class_points(0)
    macs = np.array([[physics.mass_absorption(x, e) for e in material.elements] for x in x_rays])  # (r, e)
    xd = np.asarray(detector, dtype=np.float64)
    d = xd[None, :] - pts
    lam = np.maximum((beam_x0 - pts[:, 0]) / d[:, 0], 0.0)
    dx = np.linalg.norm(d, axis=1) * lam / 100.0
    acc = np.zeros((pts.shape[0], len(x_rays)))
    ts = np.linspace(0.0, 1.0, 100)
    param = material.param
    if hasattr(param, "linear_cells"):
        # per-box attenuation sum_e mac[r][e] rho_e, looked up per ray sample
        w = np.concatenate([param.vals.reshape(-1, param.vals.shape[-1]), param.embedding[None, :]])
        rho = w / (w / material.rho_pure).sum(axis=1, keepdims=True)
        mu = rho @ macs.T
        for t in ts:
            q = pts + (t * lam)[:, None] * d
            acc -= mu[param.linear_cells(q)]  # index -1 selects the embedding row
    else:
        for t in ts:
            q = pts + (t * lam)[:, None] * d
            acc -= material.densities(q) @ macs.T
    return np.exp(acc * dx[:, None]), (pts, d, lam, dx, macs)


def absorption_pullback(material: MassFractionMaterial, aux, ac, ac_bar):
    """compute_absorption_coefficient_pullback (pn_solver/x_ray_intensities.jl:131-153),
    including its approximation (the march is re-traversed with the factor fixed)."""
    pts, d, lam, dx, macs = aux
    a_ = -ac * dx[:, None] * ac_bar                       # (P, r)
    rho_bar = a_ @ macs                                   # (P, e)
    g = np.zeros(material.param.n_params)
    for t in np.linspace(0.0, 1.0, 100):
        q = pts + (t * lam)[:, None] * d
        g += material.densities_pullback(q, rho_bar)
    return g


def solve_inputs(problem: ForwardPNProblem, material: MassFractionMaterial, detector) -> SolveInputs:
    g = problem.gspec
    rho = [np.ascontiguousarray(material.densities(g.class_points(c))) for c in range(8)]
    am = np.array([e.atom_mass_g for e in material.elements])
    noa = np.ascontiguousarray(rho[0] / am)
    ac, _ = absorption_factors(material, problem.x_rays, g, detector, problem.beam_positions[0][0])
    ei = el_index(material.elements, problem.x_rays)
    W = np.ascontiguousarray(ac * noa[:, ei])
    return SolveInputs(rho, noa, np.ascontiguousarray(ac), ei, W)
