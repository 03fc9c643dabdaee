"""ctypes binding of libsimpeg_dc.so (the C ABI declared in include/simpeg_dc.h).

torch is imported before the library is opened so that the HIP runtime torch ships
is the one the library binds to (both carry the soname libamdhip64.so.7).  The
library is required: there is no CPU fallback, and a missing or unloadable library
raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime first)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimpeg_dc.so")

SDC_OK = 0
SDC_ERR_ARG = 1
SDC_ERR_HIP = 2
SDC_ERR_NOCONV = 3
SDC_ERR_STATE = 4
SDC_ERR_ALLOC = 5

SDC_BC_ROBIN = 0
SDC_BC_NEUMANN = 1
SDC_F32 = 0
SDC_F64 = 1
SDC_MODE_MIXED = 0
SDC_MODE_FP64 = 1
SDC_MODE_MIXED_IR = 2

_dp = C.POINTER(C.c_double)


class MeshDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("hx", _dp), ("hy", _dp), ("hz", _dp),
        ("bc_type", C.c_int32),
        ("beta_xmin", _dp), ("beta_xmax", _dp),
        ("beta_ymin", _dp), ("beta_ymax", _dp),
        ("beta_zmin", _dp), ("beta_zmax", _dp),
        ("coarse_max", C.c_int32),
        ("device", C.c_int32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("nu_pre", C.c_int32), ("nu_post", C.c_int32),
        ("cheb_lmax", C.c_double), ("cheb_ratio", C.c_double),
        ("rhs_per_thread", C.c_int32), ("check_every", C.c_int32),
        ("reliable_delta", C.c_double),
        ("xfer_rhs_per_thread", C.c_int32),
        ("xcd_remap", C.c_int32),
        ("vcycle_bf16", C.c_int32),
        ("xfer_separable", C.c_int32),
        ("wcycle_level", C.c_int32),
        ("fused_tail", C.c_int32),
        ("refine_rhs_per_thread", C.c_int32),
        ("graph", C.c_int32),
        ("interp", C.c_int32),
        ("wcycle_smooth", C.c_int32),
        ("nu_coarse", C.c_int32),
        ("deriv_robin", C.c_int32),
        ("col_loop", C.c_int32),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("tol", C.c_double), ("inner_tol", C.c_double),
        ("max_outer", C.c_int32), ("max_iter", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("outer", C.c_int32), ("iters", C.c_int32),
        ("max_rel_res", C.c_double), ("converged", C.c_int32), ("fp64_iters", C.c_int32),
    ]


# name -> (restype, argtypes); every name here is declared in include/simpeg_dc.h
_vp = C.c_void_p
_SIGS = {
    "sdc_version": (C.c_char_p, []),
    "sdc_last_error": (C.c_char_p, []),
    "sdc_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(_vp)]),
    "sdc_destroy": (C.c_int, [_vp]),
    "sdc_ld": (C.c_int64, [_vp]),
    "sdc_num_nodes": (C.c_int64, [_vp]),
    "sdc_num_cells": (C.c_int64, [_vp]),
    "sdc_num_levels": (C.c_int32, [_vp]),
    "sdc_default_params": (C.c_int, [C.POINTER(Params)]),
    "sdc_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "sdc_set_sigma": (C.c_int, [_vp, _vp, _vp]),
    "sdc_set_sigma_sets": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "sdc_set_coefficients": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sdc_apply": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "sdc_vcycle": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "sdc_solve": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(SolveOpts), C.POINTER(SolveStats), _vp]),
    "sdc_solve_sets": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.POINTER(SolveOpts),
                                 C.POINTER(SolveStats), _vp]),
    "sdc_jvec_rhs": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "sdc_jtvec_contract": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int64, _vp]),
    "sdc_jtvec_contract_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    "sdc_jtj_diag_accumulate": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "sdc_gather": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sdc_scatter": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "sdc_profile": (C.c_int, [_vp, C.c_uint32]),
    "sdc_profile_read": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
}

PROF_TAGS = {"smooth": 1, "apply": 2, "residual": 3, "refine": 4, "jtvec": 5, "restrict": 6,
             "prolong": 7, "update": 8, "pair": 9, "smooth0": 10, "update_p": 11, "fold": 12}

EXPORTED = tuple(_SIGS)

_lib = None
_lock = threading.Lock()


class SolverNotConverged(RuntimeError):
    """The GPU solver stopped before every column reached its tolerance."""


def load():
    """Open libsimpeg_dc.so (built by ``python -m simpeg_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m simpeg_amd.build` "
                "(the HIP library is required; there is no CPU fallback)"
            )
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == SDC_OK:
        return
    msg = load().sdc_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == SDC_ERR_ARG:
        raise ValueError(text)
    if rc == SDC_ERR_NOCONV:
        raise SolverNotConverged(text)
    raise RuntimeError(f"libsimpeg_dc error {rc}: {text}")


def ptr(t) -> int:
    """Device (or host) address of a contiguous torch tensor."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream_ptr(device=None):
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU is visible to torch (the nodal DC solver runs on the GPU)")
    return torch.cuda.current_stream(device).cuda_stream
