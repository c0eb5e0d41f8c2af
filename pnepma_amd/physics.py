"""Physical input data for the PN-CSD model (host side).

The solver kernels only ever see per-energy tables; this module produces them.

* Stopping power: ICRU-77 total stopping power, linearly interpolated in eV and
  divided by the element density, with the derivative of the linear
  interpolant (pn_solver/stopping_power.jl:37-44).  The rows used are in
  ``data/icru77_tot_sp.json``.
* Transport coefficients: the reference's ``transport_coefficient.jl`` is not
  part of the published sources (pn_solver/pn_solver.jl:43 includes it, the
  file is absent), so this package supplies its own model: Legendre moments of
  a screened-Rutherford elastic cross section per gram.  Values are NOT the
  reference's (parity unpinned); the solver only consumes the differences
  tc[l] - tc[0] (pn_solver/pn_problem.jl:157-190).
* Ionisation cross section: NeXLCore's Bote2009 tables are not available
  offline; a Bethe-type K-shell formula (Powell's constants) stands in
  (parity unpinned).  As in the reference, no fluorescence yield is applied
  (pn_solver/x_ray_intensities.jl:32-37).
* Mass absorption coefficients: FFAST is not available offline; an
  illustrative table for the K-L3 lines of the six elements used stands in
  (parity unpinned).
* Element densities and atomic masses: the PeriodicTable values the reference
  reads through NeXLCore (pn_solver/pn_parametrizations/pn_material.jl:213,
  pn_solver/x_ray_intensities.jl:45).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass

import numpy as np
from scipy import special

ATOMIC_MASS_UNIT_G = 1.66053906660e-24  # g per unified atomic mass unit
E2_EV_CM = 1.43996448e-7                # e^2 / (4 pi eps0) in eV cm
AVOGADRO = 6.02214076e23


@dataclass(frozen=True)
class Element:
    symbol: str
    Z: int
    density: float       # g / cm^3
    atomic_mass: float   # u
    k_edge: float        # eV
    k_l3: float          # eV, K-L3 (K alpha 1) line energy

    @property
    def atom_mass_g(self) -> float:
        return self.atomic_mass * ATOMIC_MASS_UNIT_G


ELEMENTS = {
    "Al": Element("Al", 13, 2.7, 26.9815385, 1559.6, 1486.7),
    "Si": Element("Si", 14, 2.329, 28.085, 1838.9, 1739.98),
    "Cr": Element("Cr", 24, 7.19, 51.9961, 5989.2, 5414.72),
    "Fe": Element("Fe", 26, 7.874, 55.845, 7112.0, 6403.84),
    "Ni": Element("Ni", 28, 8.908, 58.6934, 8332.8, 7478.15),
    "Cu": Element("Cu", 29, 8.96, 63.546, 8978.9, 8047.78),
}


def element(sym: str) -> Element:
    return ELEMENTS[sym]


@dataclass(frozen=True)
class XRay:
    """A characteristic line: ``XRay("Cu")`` is Cu K-L3."""
    symbol: str

    @property
    def element(self) -> Element:
        return ELEMENTS[self.symbol]

    @property
    def energy(self) -> float:
        return self.element.k_l3

    @property
    def edge(self) -> float:
        return self.element.k_edge


# ---------------------------------------------------------------- stopping power
_ICRU = None


def _icru():
    global _ICRU
    if _ICRU is None:
        with open(os.path.join(os.path.dirname(__file__), "data", "icru77_tot_sp.json")) as f:
            _ICRU = json.load(f)
    return _ICRU


class StoppingPower:
    """sp(E) in eV cm^2 / g and its derivative (pn_solver/stopping_power.jl:37-44).

    Linear interpolation between the ICRU-77 knots (energies converted to eV,
    stopping powers from MeV/cm to eV/cm, divided by the density).  The
    derivative is the slope of the segment that contains E; at a knot the
    segment to the right is used (the half-open [knot_i, knot_i+1) lookup of a
    gridded linear interpolant), and the last knot uses the last segment.
    """

    def __init__(self, el: Element):
        d = _icru()
        self.knots = np.asarray(d["energy_MeV"], dtype=np.float64) * 1e6
        self.values = np.asarray(d["tot_sp_MeV_per_cm"][str(el.Z)], dtype=np.float64) * 1e6 / el.density

    def _segment(self, E):
        E = np.asarray(E, dtype=np.float64)
        if np.any(E < self.knots[0]) or np.any(E > self.knots[-1]):
            raise ValueError("energy outside the ICRU-77 table")
        i = np.searchsorted(self.knots, E, side="right") - 1
        return np.clip(i, 0, len(self.knots) - 2)

    def __call__(self, E):
        i = self._segment(E)
        x0, x1 = self.knots[i], self.knots[i + 1]
        y0, y1 = self.values[i], self.values[i + 1]
        t = (np.asarray(E) - x0) / (x1 - x0)
        return y0 + t * (y1 - y0)

    def derivative(self, E):
        i = self._segment(E)
        return (self.values[i + 1] - self.values[i]) / (self.knots[i + 1] - self.knots[i])


# ------------------------------------------------------- transport coefficients
_GL_NODES, _GL_WEIGHTS = np.polynomial.legendre.leggauss(4000)


_TC_GRID = np.geomspace(50.0, 1.0e5, 4097)
_TC_CACHE = {}


def transport_coefficients(el: Element, E, N: int) -> np.ndarray:
    """tc[l](E), l = 0..N, in cm^2/g: Legendre moments 2 pi int P_l(mu) dsigma/dOmega dmu
    of a screened-Rutherford cross section (screening 3.4 Z^(2/3) / E[eV]) per gram.

    Stand-in for the reference's missing transport_coefficient.jl.  The moments
    are tabulated once per element on 4097 log-spaced energies (50 eV - 100 keV)
    and interpolated linearly in log E.  Returns an array of shape E.shape + (N + 1,).
    """
    if N > 9:
        return _transport_coefficients_direct(el, E, N)
    tab = _TC_CACHE.get(el.symbol)
    if tab is None:
        tab = _transport_coefficients_direct(el, _TC_GRID, 9)
        _TC_CACHE[el.symbol] = tab
    E = np.asarray(E, dtype=np.float64)
    x = np.log(np.clip(E, _TC_GRID[0], _TC_GRID[-1]))
    lg = np.log(_TC_GRID)
    i = np.clip(np.searchsorted(lg, x, side="right") - 1, 0, len(lg) - 2)
    t = ((x - lg[i]) / (lg[i + 1] - lg[i]))[..., None]
    return (1.0 - t) * tab[i, :N + 1] + t * tab[i + 1, :N + 1]


def transport_table(el: Element) -> tuple:
    """(energies, tc[energy][l] for l = 0..9): the tabulation that
    transport_coefficients interpolates, as the C++ record builder takes it."""
    transport_coefficients(el, 1000.0, 9)
    return _TC_GRID, _TC_CACHE[el.symbol]


def stopping_power_table(el: Element) -> tuple:
    """(knot energies in eV, sp in eV cm^2 / g): the interpolation table of StoppingPower."""
    sp = StoppingPower(el)
    return sp.knots, sp.values


def _transport_coefficients_direct(el: Element, E, N: int) -> np.ndarray:
    E = np.atleast_1d(np.asarray(E, dtype=np.float64))
    eta = 3.4 * el.Z ** (2.0 / 3.0) / E
    pref = el.Z * (el.Z + 1) * E2_EV_CM ** 2 / (4.0 * E ** 2)
    mu = _GL_NODES
    dcs = pref[:, None] / (1.0 + 2.0 * eta[:, None] - mu[None, :]) ** 2
    P = np.stack([special.eval_legendre(l, mu) for l in range(N + 1)], axis=0)
    mom = 2.0 * math.pi * (dcs * _GL_WEIGHTS[None, :]) @ P.T
    return mom * (AVOGADRO / el.atomic_mass)


# ------------------------------------------------------------- cross sections
def ionisation_cross_section(E, xray: XRay) -> np.ndarray:
    """K-shell ionisation cross section in cm^2: Bethe form
    6.51e-14 n_K b_K ln(U) / (U E_K^2) with n_K = 2, b_K = 0.9, U = E / E_K.

    Stand-in for NeXLCore.ionizationcrosssection(..., Bote2009); zero below the edge.
    """
    E = np.asarray(E, dtype=np.float64)
    U = E / xray.edge
    with np.errstate(divide="ignore", invalid="ignore"):
        q = 6.51e-14 * 2.0 * 0.9 * np.log(np.maximum(U, 1e-300)) / (U * xray.edge ** 2)
    return np.where(U > 1.0, q, 0.0)


# illustrative mass absorption coefficients (cm^2/g) of the K-L3 line of the
# emitter (outer key) in the absorber (inner key)
_MAC = {
    "Al": {"Al": 386.0, "Si": 500.0, "Cr": 2500.0, "Fe": 3300.0, "Ni": 4200.0, "Cu": 4800.0},
    "Si": {"Al": 3300.0, "Si": 330.0, "Cr": 1700.0, "Fe": 2200.0, "Ni": 2800.0, "Cu": 3300.0},
    "Cr": {"Al": 110.0, "Si": 130.0, "Cr": 90.0, "Fe": 110.0, "Ni": 135.0, "Cu": 150.0},
    "Fe": {"Al": 70.0, "Si": 80.0, "Cr": 480.0, "Fe": 70.0, "Ni": 85.0, "Cu": 95.0},
    "Ni": {"Al": 45.0, "Si": 55.0, "Cr": 320.0, "Fe": 380.0, "Ni": 60.0, "Cu": 65.0},
    "Cu": {"Al": 49.0, "Si": 60.0, "Cr": 260.0, "Fe": 310.0, "Ni": 50.0, "Cu": 52.0},
}


def mass_absorption(xray: XRay, absorber: Element) -> float:
    return _MAC[xray.symbol][absorber.symbol]


# ------------------------------------------------------------------ beam model
def normal_cdf(x, mu, sigma):
    return special.ndtr((np.asarray(x, dtype=np.float64) - mu) / sigma)


def real_sh(l: int, k: int, theta, phi):
    """Real spherical harmonic Y_lk without Condon-Shortley phase:
    k > 0: sqrt2 N P_l^k cos(k phi); k = 0: N P_l; k < 0: sqrt2 N P_l^|k| sin(|k| phi),
    N = sqrt((2l+1)/(4 pi) (l-|k|)!/(l+|k|)!).  This is the convention under which
    the advection matrices of pn_solver/pn_matrices.jl:24-152 satisfy
    Omega_d Y = A_d Y (checked in the tests)."""
    ak = abs(k)
    x = np.cos(theta)
    # scipy's lpmv includes the Condon-Shortley phase (-1)^m; remove it
    P = special.lpmv(ak, l, x) * (-1.0) ** ak
    N = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - ak) / math.factorial(l + ak))
    if k > 0:
        return math.sqrt(2.0) * N * P * np.cos(ak * phi)
    if k < 0:
        return math.sqrt(2.0) * N * P * np.sin(ak * phi)
    return N * P


def vmf_beam_moments(moments, direction, kappa: float) -> np.ndarray:
    """-int Y_lk(Omega) vMF(Omega; direction, kappa) dOmega for each moment (l, k).

    Closed form by the Funk-Hecke theorem: -Y_lk(mu_hat) i_l(kappa) / i_0(kappa),
    with i_l the modified spherical Bessel function of the first kind.  The
    reference evaluates the same integral with adaptive cubature
    (pn_solver/pn_matrices.jl:331-338).

    Y_lk(mu_hat) is real_sh evaluated from the components of mu_hat instead of
    its angles: with x = mu_z, P_l^m(x) = sin^m(theta) Q_l^m(x) (Q_m^m = (2m-1)!!
    and the three-term recurrence in l), and sin^m(theta) (cos, sin)(m phi) =
    (Re, Im)(mu_x + i mu_y)^m.  A component that is exactly zero then makes the
    moments that vanish by symmetry exact zeros.
    """
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    x = float(d[2])
    L = max((l for l, _ in moments), default=0)
    Q = np.zeros((L + 1, L + 1))
    for m in range(L + 1):
        Q[m, m] = math.prod(2.0 * q - 1.0 for q in range(1, m + 1))
        if m + 1 <= L:
            Q[m + 1, m] = x * (2.0 * m + 1.0) * Q[m, m]
        for l in range(m + 2, L + 1):
            Q[l, m] = ((2.0 * l - 1.0) * x * Q[l - 1, m] - (l + m - 1.0) * Q[l - 2, m]) / (l - m)
    z = [complex(1.0, 0.0)]
    w = complex(float(d[0]), float(d[1]))
    for m in range(1, L + 1):
        z.append(z[-1] * w)
    i0 = special.spherical_in(0, kappa)
    out = np.empty(len(moments))
    for n, (l, k) in enumerate(moments):
        ak = abs(k)
        Nlk = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - ak) / math.factorial(l + ak))
        Y = Nlk * Q[l, ak]
        if k > 0:
            Y *= math.sqrt(2.0) * z[ak].real
        elif k < 0:
            Y *= math.sqrt(2.0) * z[ak].imag
        out[n] = -Y * special.spherical_in(l, kappa) / i0
    return out
