"""ctypes binding of libpnx.so (the C ABI declared in include/pnx.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C pnepma_amd/csrc``).  Loading fails loudly if it is missing: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpnx.so")

_dp = C.POINTER(C.c_double)


class Desc(C.Structure):
    _fields_ = [("N", C.c_int32), ("M", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("n_elements", C.c_int32), ("n_lines", C.c_int32), ("n_steps", C.c_int32), ("n_beams", C.c_int32),
                ("E_initial", C.c_double), ("E_cutoff", C.c_double), ("ckpt_every", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("desc", Desc), ("fwd", _dp), ("adj", _dp), ("dens", _dp), ("rho", _dp * 8), ("W", _dp),
                ("mms", _dp * 8), ("ppos", _dp * 8), ("I_bar", _dp), ("arrival_s", C.c_double), ("I_out", _dp),
                ("rho_bar", _dp * 8), ("gen_out", _dp)]


class Physics(C.Structure):
    _fields_ = [("n_knots", C.c_int32), ("sp_E", _dp), ("sp", _dp * 8), ("n_tc", C.c_int32), ("tc_stride", C.c_int32),
                ("tc_E", _dp), ("tc", _dp * 8), ("line_edge", C.c_double * 8), ("beam_E_mean", C.c_double),
                ("beam_E_sd", C.c_double)]


class RequestSpec(C.Structure):
    _fields_ = [("desc", Desc), ("physics", C.POINTER(Physics)), ("cell_w", _dp), ("rho_pure", _dp),
                ("embedding", _dp), ("mac", _dp), ("atom_mass", _dp), ("line_element", C.POINTER(C.c_int32)),
                ("detector", C.c_double * 3), ("beam_position", C.c_double * 3), ("beam_sigma", C.c_double),
                ("beam_direction", C.c_double * 3), ("beam_kappa", C.c_double), ("literal_source", C.c_int32),
                ("I_bar", _dp), ("arrival_s", C.c_double), ("I_out", _dp), ("rho_bar", _dp * 8), ("gen_out", _dp),
                ("grad_out", _dp)]


class Timing(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double), ("status", C.c_int32),
                ("ticks", C.c_int32), ("taped", C.c_int32)]


# name -> (restype, argtypes); every symbol include/pnx.h declares
SIGNATURES = {
    "pnx_last_error": (C.c_char_p, []),
    "pnx_version": (C.c_int, []),
    "pnx_context_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pnx_context_destroy": (C.c_int, [C.c_void_p]),
    "pnx_context_synchronize": (C.c_int, [C.c_void_p]),
    "pnx_problem_create": (C.c_int, [C.c_void_p, C.POINTER(Desc), C.POINTER(C.c_void_p)]),
    "pnx_problem_destroy": (C.c_int, [C.c_void_p]),
    "pnx_problem_device_bytes": (C.c_int64, [C.c_void_p]),
    "pnx_problem_set_records": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pnx_problem_records": (C.c_int, [C.POINTER(Desc), C.POINTER(Physics), _dp, _dp, _dp]),
    "pnx_beam_moments": (C.c_int, [C.c_int, _dp, C.c_double, _dp * 8]),
    "pnx_beam_profiles": (C.c_int, [C.POINTER(Desc), _dp, C.c_double, C.c_int, _dp * 8]),
    "pnx_problem_set_material": (C.c_int, [C.c_void_p, _dp * 8, _dp]),
    "pnx_problem_set_beams": (C.c_int, [C.c_void_p, _dp * 8, C.POINTER(_dp)]),
    "pnx_problem_set_cell_fractions": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pnx_problem_cell_fraction_gradient": (C.c_int, [C.c_void_p, _dp, _dp]),
    "pnx_problem_set_absorption": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32), _dp, C.c_double,
                                             _dp * 3, _dp * 3, _dp, _dp]),
    "pnx_problem_material_gradient": (C.c_int, [C.c_void_p, _dp]),
    "pnx_march_sequence": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_int32]),
    "pnx_forward": (C.c_int, [C.c_void_p, _dp]),
    "pnx_forward_noalloc": (C.c_int, [C.c_void_p, _dp]),
    "pnx_problem_reserve_adjoint": (C.c_int, [C.c_void_p, C.c_int]),
    "pnx_forward_adjoint": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(_dp), _dp]),
    "pnx_forward_adjoint_exact": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(_dp), _dp]),
    "pnx_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp * 8]),
    "pnx_forward_timed": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "pnx_forward_adjoint_timed": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "pnx_half_step_bench": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pnx_rev_bench": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pnx_server_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "pnx_server_destroy": (C.c_int, [C.c_void_p]),
    "pnx_server_clock": (C.c_double, [C.c_void_p]),
    "pnx_server_set_tape_limit": (C.c_int, [C.c_void_p, C.c_int64]),
    "pnx_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "pnx_server_submit_spec": (C.c_int, [C.c_void_p, C.POINTER(RequestSpec), C.POINTER(C.c_int64)]),
    "pnx_server_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(Timing)]),
    "pnx_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Timing)]),
    "pnx_server_cancel": (C.c_int, [C.c_void_p, C.c_int64]),
    "pnx_server_drain": (C.c_int, [C.c_void_p]),
    "pnx_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
    "pnx_request_device_bytes": (C.c_int64, [C.POINTER(Desc)]),
    "pnx_coefficients": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int]),
    "pnx_advection_block": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp]),
    "pnx_boundary_block": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
}

_lib = None


PNX_ERR_CANCELLED = -6


class PnxError(RuntimeError):
    pass


def lib():
    """Load libpnx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PnxError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C pnepma_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise PnxError(f"pnx error {rc}: {lib().pnx_last_error().decode()}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None
