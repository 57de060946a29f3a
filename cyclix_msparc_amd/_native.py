"""ctypes binding of the gfx950 C-ABI library (include/cyclix_msparc.h).

The library is built in-tree by ``cyclix_msparc_amd.build``; there is no
fallback: if it is missing or does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libcyclix_msparc.so"

CYX_OK, CYX_ERR_ARG, CYX_ERR_HIP, CYX_ERR_NOCONV, CYX_ERR_NOMEM, CYX_ERR_SOLVER, CYX_ERR_STATE = 0, -1, -2, -3, -4, -5, -6
CYX_F64, CYX_C128, CYX_F32, CYX_C64 = 0, 1, 2, 3
DTYPE_NAMES = {"f64": CYX_F64, "c128": CYX_C128, "f32": CYX_F32, "c64": CYX_C64}

# every symbol declared in include/cyclix_msparc.h
EXPORTS = (
    "cyx_last_error", "cyx_version", "cyx_ctx_create", "cyx_ctx_destroy", "cyx_ctx_set_stream",
    "cyx_ctx_synchronize", "cyx_ctx_set_timing", "cyx_ctx_reset_timing", "cyx_ctx_read_timing", "cyx_ham_create", "cyx_ham_set_veff", "cyx_ham_destroy", "cyx_ham_weights", "cyx_ham_rotation_counts",
    "cyx_ham_set_tuning",
    "cyx_h_apply", "cyx_chebyshev_filter", "cyx_chebyshev_filter_noalloc", "cyx_eigsolver_workspace_bytes",
    "cyx_eigsolver_pass",
    "cyx_rayleigh_ritz", "cyx_ritz_residuals", "cyx_weighted_gram", "cyx_lanczos_bounds", "cyx_density_accumulate", "cyx_lda_pw", "cyx_energy_terms",
    "cyx_density_response", "cyx_lda_pw_fxc",
    "cyx_poisson_create", "cyx_poisson_destroy", "cyx_aar",
    "cyx_poisson_solve", "cyx_scf_terms", "cyx_mixer_create", "cyx_mixer_destroy", "cyx_mixer_step",
    "cyx_sternheimer_apply", "cyx_sternheimer_solve", "cyx_sternheimer_plan_create", "cyx_sternheimer_plan_iterate",
    "cyx_sternheimer_plan_status", "cyx_sternheimer_plan_destroy", "cyx_sternheimer_rhs",
    "cyx_unet_create", "cyx_unet_destroy", "cyx_unet_num_params", "cyx_unet_set_params", "cyx_unet_get_params",
    "cyx_unet_forward", "cyx_unet_gradient", "cyx_unet_train_step", "cyx_unet_input",
    "cyx_poisson_set_preconditioner", "cyx_poisson_apply", "cyx_poisson_set_spectral", "cyx_poisson_spectral_apply",
    "cyx_filter_plan_create", "cyx_filter_plan_run", "cyx_filter_plan_tuning", "cyx_filter_plan_destroy",
    "cyx_filter_plan_attach_sternheimer",
)


class CyxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cyx status {code}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nt", C.c_int32), ("nz", C.c_int32), ("fdn", C.c_int32),
                ("cell_typ", C.c_int32), ("reserved", C.c_int32),
                ("xin", C.c_double), ("dr", C.c_double), ("dt", C.c_double), ("dz", C.c_double),
                ("twist", C.c_double)]


class ProjectorDesc(C.Structure):
    _fields_ = [("n_atom", C.c_int32), ("n_image", C.c_int32),
                ("angnum", C.POINTER(C.c_int32)), ("gamma", C.POINTER(C.c_double)),
                ("image_atom", C.POINTER(C.c_int32)), ("image_nnodes", C.POINTER(C.c_int32)),
                ("image_disp", C.POINTER(C.c_double)), ("rc_pos", C.POINTER(C.c_int32)),
                ("chi", C.POINTER(C.c_double))]


class Tuning(C.Structure):
    _fields_ = [("tr", C.c_int32), ("zb", C.c_int32), ("flags", C.c_int32), ("reserved", C.c_int32)]


class AarInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double),
                ("precond", C.c_int32), ("pinv_eig", C.c_int32)]


class SternInfo(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("n_converged", C.c_int32), ("max_relres", C.c_double)]


_lib = None


def load():
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m cyclix_msparc_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    pd = C.POINTER(C.c_double)
    sig = {
        "cyx_last_error": (C.c_char_p, []),
        "cyx_version": (C.c_char_p, []),
        "cyx_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "cyx_ctx_destroy": (i32, [vp]),
        "cyx_ctx_set_stream": (i32, [vp, vp]),
        "cyx_ctx_synchronize": (i32, [vp]),
        "cyx_ctx_set_timing": (i32, [vp, i32]),
        "cyx_ctx_reset_timing": (i32, [vp]),
        "cyx_ctx_read_timing": (i32, [vp, C.c_char_p, pd, C.POINTER(C.c_int64)]),
        "cyx_ham_create": (i32, [vp, C.POINTER(GridDesc), pd, pd, pd, C.POINTER(ProjectorDesc), C.POINTER(vp)]),
        "cyx_ham_set_veff": (i32, [vp, vp, i32]),
        "cyx_ham_destroy": (i32, [vp]),
        "cyx_ham_weights": (vp, [vp]),
        "cyx_ham_rotation_counts": (i32, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "cyx_ham_set_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "cyx_h_apply": (i32, [vp, i32, vp, vp, i32, i64]),
        "cyx_chebyshev_filter": (i32, [vp, i32, vp, vp, i32, i64, i32, dbl, dbl, dbl]),
        "cyx_chebyshev_filter_noalloc": (i32, [vp, i32, vp, vp, i32, i64, i32, dbl, dbl, dbl]),
        "cyx_eigsolver_workspace_bytes": (i64, [vp, i32, i32]),
        "cyx_eigsolver_pass": (i32, [vp, i32, vp, vp, i32, i32, dbl, dbl, dbl, pd]),
        "cyx_rayleigh_ritz": (i32, [vp, i32, vp, vp, i32, pd]),
        "cyx_ritz_residuals": (i32, [vp, i32, vp, vp, i32, pd, pd]),
        "cyx_lanczos_bounds": (i32, [vp, pd, i32, dbl, pd, pd]),
        "cyx_density_accumulate": (i32, [vp, i32, vp, i32, pd, dbl, vp]),
        "cyx_lda_pw": (i32, [vp, i64, vp, vp, vp, vp, dbl]),
        "cyx_density_response": (i32, [vp, i32, vp, vp, i32, pd, pd, dbl, vp]),
        "cyx_lda_pw_fxc": (i32, [vp, i64, vp, vp, vp, vp, vp, dbl]),
        "cyx_energy_terms": (i32, [vp, vp, vp, vp, vp, dbl, pd]),
        "cyx_poisson_create": (i32, [vp, C.POINTER(GridDesc), C.POINTER(vp)]),
        "cyx_poisson_destroy": (i32, [vp]),
        "cyx_aar": (i32, [vp, vp, vp, i32, dbl, i32, dbl, dbl, i32, i32, C.POINTER(AarInfo)]),
        "cyx_poisson_solve": (i32, [vp, vp, vp, vp, i32, dbl, i32, C.POINTER(AarInfo)]),
        "cyx_scf_terms": (i32, [vp, vp, vp, vp, pd]),
        "cyx_mixer_create": (i32, [vp, i64, i32, dbl, dbl, i32, vp, C.POINTER(vp)]),
        "cyx_mixer_destroy": (i32, [vp]),
        "cyx_mixer_step": (i32, [vp, vp, vp, vp, i32]),
        "cyx_unet_create": (i32, [vp, C.POINTER(GridDesc), C.POINTER(vp)]),
        "cyx_unet_destroy": (i32, [vp]),
        "cyx_unet_num_params": (i64, [vp]),
        "cyx_unet_set_params": (i32, [vp, C.POINTER(C.c_float), i32]),
        "cyx_unet_get_params": (i32, [vp, C.POINTER(C.c_float)]),
        "cyx_unet_forward": (i32, [vp, vp, vp, i32]),
        "cyx_unet_gradient": (i32, [vp, vp, vp, i32, C.POINTER(C.c_float), pd]),
        "cyx_unet_train_step": (i32, [vp, vp, vp, i32, C.c_float, pd]),
        "cyx_unet_input": (i32, [vp, vp, vp, vp, i32, vp]),
        "cyx_poisson_set_preconditioner": (i32, [vp, vp]),
        "cyx_poisson_apply": (i32, [vp, vp, vp]),
        "cyx_poisson_set_spectral": (i32, [vp, i32]),
        "cyx_poisson_spectral_apply": (i32, [vp, vp, vp, i32, i64]),
        "cyx_filter_plan_create": (i32, [vp, i32, C.POINTER(vp), i32, C.POINTER(vp), C.POINTER(vp),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int64), i32, i32, pd, pd, pd,
                                         C.POINTER(vp)]),
        "cyx_filter_plan_run": (i32, [vp, pd, pd, pd, C.POINTER(C.c_int32)]),
        "cyx_filter_plan_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "cyx_filter_plan_destroy": (i32, [vp]),
        "cyx_sternheimer_apply": (i32, [vp, i32, vp, vp, i32, i64, pd, vp, i64, i32]),
        "cyx_sternheimer_solve": (i32, [vp, i32, vp, vp, i32, i64, pd, vp, i64, i32, i32, dbl, i32, i32,
                                        C.POINTER(SternInfo), C.POINTER(C.c_int32), pd]),
        "cyx_sternheimer_plan_create": (i32, [vp, i32, vp, vp, i32, i64, pd, vp, i64, i32, i32, dbl, i32,
                                              C.POINTER(vp)]),
        "cyx_sternheimer_plan_iterate": (i32, [vp, i32]),
        "cyx_sternheimer_plan_status": (i32, [vp, C.POINTER(C.c_int32), C.POINTER(SternInfo),
                                              C.POINTER(C.c_int32), pd]),
        "cyx_sternheimer_plan_destroy": (i32, [vp]),
        "cyx_sternheimer_rhs": (i32, [vp, i32, vp, i64, i32, pd, vp, i64, pd]),
        "cyx_filter_plan_attach_sternheimer": (i32, [vp, vp, i32]),
        "cyx_weighted_gram": (i32, [vp, i32, vp, i64, i32, vp, i64, i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code, allow=()):
    if code != CYX_OK and code not in allow:
        raise CyxError(code, load().cyx_last_error().decode())
    return code


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))
