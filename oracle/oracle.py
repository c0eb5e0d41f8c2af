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


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "pnx_ref.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
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
    subprocess.check_call(["gcc", "-O3", "-march=native", "-std=c99", "-fPIC", "-shared", "-o", out,
                           os.path.join(HERE, "pnx_ref.c"), "-lm"])
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
    keep.mms = problem.beam_moments()
    keep.ppos = problem.beam_profiles(beam)
    inp = Inputs()
    for c in range(8):
        inp.rho[c] = _ptr(keep.rho[c])
        inp.mms[c] = _ptr(keep.mms[c])
        inp.ppos[c] = _ptr(keep.ppos[c])
    inp.ac = _ptr(keep.ac)
    inp.noa = _ptr(keep.noa)
    inp.el_idx = keep.el.ctypes.data_as(C.POINTER(C.c_int))
    return inp, keep


def forward(problem, si, beam: int = 0) -> np.ndarray:
    d = desc_of(problem)
    inp, keep = inputs_of(problem, si, beam)
    fwd, _, _ = problem.records()
    out = np.zeros(len(problem.x_rays))
    rc = lib().ref_forward(C.byref(d), C.byref(inp), _ptr(fwd), _ptr(out))
    if rc:
        raise RuntimeError("ref_forward failed")
    return out


def forward_adjoint(problem, si, I_bar, beam: int = 0):
    """Returns (I, rho_bar[8] as (P_c, n_e), ac_bar (P, r), noa_bar (P, e))."""
    d = desc_of(problem)
    inp, keep = inputs_of(problem, si, beam)
    fwd, adj, dens = problem.records()
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
    mms = problem.beam_moments()
    ppos = problem.beam_profiles(beam)
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
