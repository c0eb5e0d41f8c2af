"""ctypes front end of the CPU oracle (libpnx_oracle.so) -- TEST INFRASTRUCTURE.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use
this module; the product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpnx_oracle.so")

_dp = C.POINTER(C.c_double)


class Desc(C.Structure):
    _fields_ = [("N", C.c_int), ("M", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("n_e", C.c_int), ("n_lines", C.c_int), ("n_steps", C.c_int),
                ("E_initial", C.c_double), ("E_cutoff", C.c_double)]


class Inputs(C.Structure):
    _fields_ = [("rho", _dp * 8), ("ac", _dp), ("noa", _dp), ("el_idx", C.POINTER(C.c_int)),
                ("mms", _dp * 8), ("ppos", _dp * 8)]


SOURCES = ("pnx_ref.c", "pnx_ref_material.c", "pnx_ref_records.c")


def build(force: bool = False) -> str:
    newest = max(os.path.getmtime(os.path.join(HERE, f)) for f in SOURCES + ("pnx_ref.h",))
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB


_lib = None
_path = LIB


def use_native(out_dir: str) -> str:
    """Builds a copy of the oracle for the host it runs on (-O3 -march=native)
    into out_dir and loads that copy instead of the committed build (which
    targets x86-64-v3 so that it runs on any host it is copied to).  Used by the
    benchmark's CPU baseline."""
    global _lib, _path
    out = os.path.join(out_dir, "libpnx_oracle_native.so")
    subprocess.check_call(["gcc", "-O3", "-march=native", "-std=c99", "-fPIC", "-shared", "-o", out] +
                          [os.path.join(HERE, f) for f in SOURCES] + ["-lm"])
    _lib, _path = None, out
    lib()
    return out


def lib():
    global _lib
    if _lib is None:
        if _path == LIB and not os.path.exists(LIB):
            build()
        L = C.CDLL(_path)
        L.ref_forward.argtypes = [C.POINTER(Desc), C.POINTER(Inputs), _dp, _dp]
        L.ref_forward_adjoint.argtypes = [C.POINTER(Desc), C.POINTER(Inputs), _dp, _dp, _dp, _dp, _dp,
                                          _dp * 8, _dp, _dp]
        L.ref_step.argtypes = [C.POINTER(Desc), C.c_int, _dp, _dp * 8, _dp * 8, _dp * 8, _dp, _dp * 8]
        L.ref_advection_dense.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.ref_boundary_dense.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.ref_class_moments.argtypes = [C.c_int, C.c_int]
        L.ref_class_points.argtypes = [C.POINTER(Desc), C.c_int]
        L.ref_record_len.argtypes = [C.POINTER(Desc)]
        L.ref_set_plant.argtypes = [C.c_int]
        L.ref_set_plant.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def desc_of(problem) -> Desc:
    g = problem.gspec
    d = Desc()
    d.N = problem.N
    for a in range(3):
        d.M[a] = g.axes[a].M
        d.lo[a] = g.axes[a].lo
        d.hi[a] = g.axes[a].hi
    d.n_e = len(problem.elements)
    d.n_lines = len(problem.x_rays)
    d.n_steps = problem.n_steps
    d.E_initial = problem.E_initial
    d.E_cutoff = problem.E_cutoff
    return d


class _Keep:
    """Holds the numpy arrays whose pointers an Inputs struct refers to."""


def inputs_of(problem, si, beam: int):
    keep = _Keep()
    keep.rho = [np.ascontiguousarray(r, dtype=np.float64) for r in si.rho]
    keep.ac = np.ascontiguousarray(si.ac, dtype=np.float64)
    keep.noa = np.ascontiguousarray(si.noa, dtype=np.float64)
    keep.el = np.ascontiguousarray(si.el_idx, dtype=np.int32)
    keep.mms = problem.beam_moments_host()
    keep.ppos = problem.beam_profiles_host(beam)
    inp = Inputs()
    for c in range(8):
        inp.rho[c] = _ptr(keep.rho[c])
        inp.mms[c] = _ptr(keep.mms[c])
        inp.ppos[c] = _ptr(keep.ppos[c])
    inp.ac = _ptr(keep.ac)
    inp.noa = _ptr(keep.noa)
    inp.el_idx = keep.el.ctypes.data_as(C.POINTER(C.c_int))
    return inp, keep


def _records(problem):
    """The oracle's own records of the problem (cached on it).  Together with the
    numpy beam moments and profiles this keeps the oracle independent of
    libpnx.so, so it runs in a process that has not loaded the library."""
    key = "oracle_records"
    if key not in problem._cache:
        problem._cache[key] = records(problem)
    return problem._cache[key]


def forward(problem, si, beam: int = 0) -> np.ndarray:
    d = desc_of(problem)
    inp, keep = inputs_of(problem, si, beam)
    fwd, _, _ = _records(problem)
    out = np.zeros(len(problem.x_rays))
    rc = lib().ref_forward(C.byref(d), C.byref(inp), _ptr(fwd), _ptr(out))
    if rc:
        raise RuntimeError("ref_forward failed")
    return out


def forward_adjoint(problem, si, I_bar, beam: int = 0):
    """Returns (I, rho_bar[8] as (P_c, n_e), ac_bar (P, r), noa_bar (P, e))."""
    d = desc_of(problem)
    inp, keep = inputs_of(problem, si, beam)
    fwd, adj, dens = _records(problem)
    ne = len(problem.elements)
    I = np.zeros(len(problem.x_rays))
    rb = [np.zeros((lib().ref_class_points(C.byref(d), c), ne)) for c in range(8)]
    acb = np.zeros_like(si.ac)
    noab = np.zeros_like(si.noa)
    ib = np.ascontiguousarray(I_bar, dtype=np.float64)
    arr = (_dp * 8)(*[_ptr(r) for r in rb])
    rc = lib().ref_forward_adjoint(C.byref(d), C.byref(inp), _ptr(fwd), _ptr(adj), _ptr(dens), _ptr(ib), _ptr(I), arr,
                                   _ptr(acb), _ptr(noab))
    if rc:
        raise RuntimeError("ref_forward_adjoint failed")
    return I, rb, acb, noab


def step(problem, si, u, rec, adjoint=False, q_eee=None, beam: int = 0):
    """One step_pde! on the list of 8 per-class arrays u (modified in place)."""
    d = desc_of(problem)
    rho = (_dp * 8)(*[_ptr(np.ascontiguousarray(r)) for r in si.rho])
    mms = problem.beam_moments_host()
    ppos = problem.beam_profiles_host(beam)
    m8 = (_dp * 8)(*[_ptr(m) for m in mms])
    p8 = (_dp * 8)(*[_ptr(p) for p in ppos])
    u8 = (_dp * 8)(*[_ptr(x) for x in u])
    q = np.ascontiguousarray(q_eee) if q_eee is not None else None
    rc = lib().ref_step(C.byref(d), int(adjoint), _ptr(np.ascontiguousarray(rec)), rho, m8, p8, _ptr(q), u8)
    if rc:
        raise RuntimeError("ref_step failed")


def set_plant(k: int):
    """Plants a known mistake in the oracle's adjoint (see pnx_ref.h); 0 removes it."""
    lib().ref_set_plant(int(k))


def advection_dense(N, c, d, nc, nf):
    out = np.zeros((nc, nf))
    lib().ref_advection_dense(N, c, d, _ptr(out))
    return out


def boundary_dense(N, c, d, adjoint, nc, nf):
    out = np.zeros((nc, nf))
    if lib().ref_boundary_dense(N, c, d, int(adjoint), _ptr(out)):
        raise ValueError("class not odd in axis")
    return out


# ------------------------------------------------------------------ material
class RefBoxes(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("edges", _dp * 3),
                ("n_e", C.c_int), ("vals", _dp), ("embedding", _dp), ("rho_pure", _dp)]


def _bind_material(L):
    B = C.POINTER(RefBoxes)
    L.ref_box_index.argtypes = [B, _dp]
    L.ref_box_index.restype = C.c_long
    L.ref_box_index_floor.argtypes = [B, _dp]
    L.ref_box_index_floor.restype = C.c_long
    L.ref_component_densities.argtypes = [B, _dp, C.c_long, _dp]
    L.ref_component_densities_pullback.argtypes = [B, _dp, C.c_long, _dp, _dp, _dp]
    L.ref_absorption.argtypes = [B, _dp, C.c_long, _dp, C.c_int, _dp, C.c_double, _dp]
    L.ref_absorption_pullback.argtypes = [B, _dp, C.c_long, _dp, C.c_int, _dp, C.c_double, _dp, _dp, _dp, _dp]
    L.ref_number_of_atoms.argtypes = [B, _dp, C.c_long, _dp, _dp]
    L.ref_number_of_atoms_pullback.argtypes = [B, _dp, C.c_long, _dp, _dp, _dp, _dp]


class Boxes:
    """The oracle's view of a MassFraction material: a box parametrization with
    uniform slices (CellFractions), or the embedding everywhere (ConstFractions)."""

    def __init__(self, material):
        par = material.param
        self.ne = len(material.elements)
        self.s = RefBoxes()
        if hasattr(par, "vals"):
            self.shape = par.vals.shape[:3]
            self.vals = np.ascontiguousarray(par.vals.transpose(2, 1, 0, 3).reshape(-1, self.ne))
            self.emb = np.ascontiguousarray(par.embedding, dtype=np.float64)
            self.edges = [np.ascontiguousarray(e, dtype=np.float64) for e in par.edges]
            for a in range(3):
                self.s.n[a] = self.shape[a]
                self.s.lo[a] = float(par.edges[a][0])
                self.s.hi[a] = float(par.edges[a][-1])
                self.s.edges[a] = _ptr(self.edges[a])
        else:
            self.shape = (0, 0, 0)
            self.vals = np.zeros((1, self.ne))
            self.emb = np.ascontiguousarray(par.w, dtype=np.float64)
        self.rho_pure = np.ascontiguousarray(material.rho_pure, dtype=np.float64)
        self.s.n_e = self.ne
        self.s.vals = _ptr(self.vals)
        self.s.embedding = _ptr(self.emb)
        self.s.rho_pure = _ptr(self.rho_pure)

    def ref(self):
        L = lib()
        if not hasattr(L, "_material_bound"):
            _bind_material(L)
            L._material_bound = True
        return C.byref(self.s)

    def params_grad(self, vals_bar):
        """[box x fastest][element] -> the CellFractions.params order (x slowest)."""
        nx, ny, nz = self.shape
        return vals_bar.reshape(nz, ny, nx, self.ne).transpose(2, 1, 0, 3).ravel()


def _pts(pts):
    return np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)


def component_densities(material, pts):
    b, pts = Boxes(material), _pts(pts)
    out = np.zeros((pts.shape[0], b.ne))
    lib().ref_component_densities(b.ref(), _ptr(pts), pts.shape[0], _ptr(out))
    return out


def box_index(material, x, literal_floor: bool = False):
    """Box of point x (-1: embedding); literal_floor: the reference's floating-point floor."""
    b = Boxes(material)
    ref = b.ref()
    f = lib().ref_box_index_floor if literal_floor else lib().ref_box_index
    return int(f(ref, _ptr(np.ascontiguousarray(x, dtype=np.float64))))


def component_densities_pullback(material, pts, rho_bar):
    """(gradient in CellFractions.params order, embedding gradient)."""
    b, pts = Boxes(material), _pts(pts)
    vb, eb = np.zeros_like(b.vals), np.zeros(b.ne)
    lib().ref_component_densities_pullback(b.ref(), _ptr(pts), pts.shape[0],
                                           _ptr(np.ascontiguousarray(rho_bar, dtype=np.float64)), _ptr(vb), _ptr(eb))
    return b.params_grad(vb), eb


def absorption(material, pts, mac, detector, x_b0):
    """ac[p][r] from every point; mac[r][e]."""
    b, pts = Boxes(material), _pts(pts)
    mac = np.ascontiguousarray(mac, dtype=np.float64)
    out = np.zeros((pts.shape[0], mac.shape[0]))
    lib().ref_absorption(b.ref(), _ptr(pts), pts.shape[0], _ptr(mac), mac.shape[0],
                         _ptr(np.ascontiguousarray(detector, dtype=np.float64)), float(x_b0), _ptr(out))
    return out


def absorption_pullback(material, pts, mac, detector, x_b0, ac, ac_bar):
    b, pts = Boxes(material), _pts(pts)
    mac = np.ascontiguousarray(mac, dtype=np.float64)
    vb, eb = np.zeros_like(b.vals), np.zeros(b.ne)
    lib().ref_absorption_pullback(b.ref(), _ptr(pts), pts.shape[0], _ptr(mac), mac.shape[0],
                                  _ptr(np.ascontiguousarray(detector, dtype=np.float64)), float(x_b0),
                                  _ptr(np.ascontiguousarray(ac)), _ptr(np.ascontiguousarray(ac_bar)), _ptr(vb),
                                  _ptr(eb))
    return b.params_grad(vb), eb


def number_of_atoms(material, pts, atom_mass):
    b, pts = Boxes(material), _pts(pts)
    out = np.zeros((pts.shape[0], b.ne))
    lib().ref_number_of_atoms(b.ref(), _ptr(pts), pts.shape[0], _ptr(np.ascontiguousarray(atom_mass)), _ptr(out))
    return out


def number_of_atoms_pullback(material, pts, atom_mass, noa_bar):
    b, pts = Boxes(material), _pts(pts)
    vb, eb = np.zeros_like(b.vals), np.zeros(b.ne)
    lib().ref_number_of_atoms_pullback(b.ref(), _ptr(pts), pts.shape[0], _ptr(np.ascontiguousarray(atom_mass)),
                                       _ptr(np.ascontiguousarray(noa_bar, dtype=np.float64)), _ptr(vb), _ptr(eb))
    return b.params_grad(vb), eb


def material_gradient(problem, material, detector, mac, rho_bar, ac_bar, noa_bar):
    """The material pullbacks of the reference adjoint (pn_intensities.jl:406-410)
    from the oracle's rho_bar (per class), ac_bar and noa_bar: the densities of
    every class, the absorption factors and the numbers of atoms.  Returns the
    gradient in CellFractions.params order."""
    g = problem.gspec
    am = np.array([e.atom_mass_g for e in material.elements])
    eee = g.class_points(0)
    grad = np.zeros(material.param.n_params)
    for c in range(8):
        if rho_bar[c].size and np.any(rho_bar[c]):
            grad += component_densities_pullback(material, g.class_points(c), rho_bar[c])[0]
    ac = absorption(material, eee, mac, detector, problem.beam_positions[0][0])
    grad += absorption_pullback(material, eee, mac, detector, problem.beam_positions[0][0], ac, ac_bar)[0]
    grad += number_of_atoms_pullback(material, eee, am, noa_bar)[0]
    return grad


# ------------------------------------------------------------------- records
class RefTables(C.Structure):
    _fields_ = [("n_knots", C.c_int), ("sp_E", _dp), ("sp", _dp * 8), ("n_tc", C.c_int), ("tc_stride", C.c_int),
                ("tc_E", _dp), ("tc", _dp * 8), ("line_edge", C.c_double * 8), ("beam_E_mean", C.c_double),
                ("beam_E_sd", C.c_double)]


def _tables(problem):
    """The oracle's element tables from the same data the package uses: the
    ICRU-77 rows over the density and the transport-coefficient tabulation."""
    from pnepma_amd import physics
    T = RefTables()
    keep = []
    for e, el in enumerate(problem.elements):
        kn, sp = physics.stopping_power_table(el)
        tE, tc = physics.transport_table(el)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (kn, sp, tE, tc)]
        keep += arrs
        T.sp_E, T.sp[e], T.tc_E, T.tc[e] = [_ptr(a) for a in arrs]
        T.n_knots, T.n_tc, T.tc_stride = arrs[0].size, arrs[3].shape[0], arrs[3].shape[1]
    for r, x in enumerate(problem.x_rays):
        T.line_edge[r] = x.edge
    T.beam_E_mean, T.beam_E_sd = problem.beam_energy
    return T, keep


def _bind_records(L):
    L.ref_records.argtypes = [C.POINTER(Desc), C.POINTER(RefTables), _dp, _dp, _dp]
    L.ref_record_at.argtypes = [C.POINTER(Desc), C.POINTER(RefTables), C.c_double, C.c_double, C.c_int, _dp]


def records(problem):
    """(fwd, adj, dens) as the solver takes them, from the oracle."""
    L = lib()
    _bind_records(L)
    T, keep = _tables(problem)
    d = desc_of(problem)
    out = tuple(np.zeros((problem.n_steps + 1, L.ref_record_len(C.byref(d)))) for _ in range(3))
    if L.ref_records(C.byref(d), C.byref(T), *[_ptr(a) for a in out]):
        raise ValueError("an energy lies outside the stopping-power table")
    return out


def record_at(problem, Ec, Es, adjoint=False):
    L = lib()
    _bind_records(L)
    T, keep = _tables(problem)
    d = desc_of(problem)
    out = np.zeros(L.ref_record_len(C.byref(d)))
    if L.ref_record_at(C.byref(d), C.byref(T), float(Ec), float(Es), int(adjoint), _ptr(out)):
        raise ValueError("energy outside the stopping-power table")
    return out
