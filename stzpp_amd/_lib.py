"""ctypes loader for the in-tree HIP library (stzpp_amd/libstzmg.so).

The library is the product path: there is no CPU fallback.  Loading fails
loudly when the shared object is missing.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libstzmg.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class StzmgDesc(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("o", ctypes.c_int),
        ("prd", ctypes.c_int * 3),
        ("coarsest_total", ctypes.c_long),
        ("v_down", ctypes.c_int), ("v_up", ctypes.c_int),
        ("bottom_sweeps", ctypes.c_int), ("sweep_growth", ctypes.c_int),
        ("op_mode", ctypes.c_int),
    ]


class StzmgSolveOpts(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("extra_cycles", ctypes.c_int),
        ("ref_norm_schedule", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("max_cycles", ctypes.c_int),
        ("inner_cycles", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("restart", ctypes.c_int),
        ("inner_rtol", ctypes.c_double),
    ]


class StzmgSolveStats(ctypes.Structure):
    _fields_ = [
        ("cycles", ctypes.c_int), ("outer", ctypes.c_int), ("norms", ctypes.c_int),
        ("res0", ctypes.c_double), ("res", ctypes.c_double), ("ms", ctypes.c_double),
    ]


class StzParams(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("prd", ctypes.c_int * 3)] + [
        (k, ctypes.c_double) for k in (
            "ax", "bx", "ay", "by", "az", "bz", "mu", "lamb", "sy", "tau0", "c0", "eps0",
            "delta", "omega", "bath_temp", "chi_inf", "ez", "diff_l", "U", "eta")] + [
        ("transform", ctypes.c_int), ("Ut", ctypes.c_double), ("rho", ctypes.c_double),
        ("kappa", ctypes.c_double), ("tmult", ctypes.c_double)]


class StzAdjointOpts(ctypes.Structure):
    _fields_ = [("checkpoint_every", ctypes.c_int), ("rtol", ctypes.c_double)]


class StzGrfParams(ctypes.Structure):
    _fields_ = [("gauss_fac", ctypes.c_double), ("conv_l", ctypes.c_double), ("chi_avg", ctypes.c_double),
                ("sigma", ctypes.c_double), ("TZ", ctypes.c_double), ("sim_case", ctypes.c_int),
                ("seed", ctypes.c_ulonglong)]


class StzStepStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("vcycles", ctypes.c_int), ("yielded_cells", ctypes.c_int),
                ("solve_ms", ctypes.c_double), ("step_ms", ctypes.c_double),
                ("last_res", ctypes.c_double)]


# Every symbol declared in include/*.h with (argtypes, restype).
SIGNATURES = {
    "stz_params_default": ([ctypes.POINTER(StzParams)] + [ctypes.c_int] * 3, ctypes.c_int),
    "stz_create": ([ctypes.POINTER(StzParams), ctypes.POINTER(_vp)], ctypes.c_int),
    "stz_destroy": ([_vp], ctypes.c_int),
    "stz_last_error": ([_vp], ctypes.c_char_p),
    "stz_set_state": ([_vp, ctypes.c_double, _dp, _dp, _dp], ctypes.c_int),
    "stz_get_state": ([_vp, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "stz_step_qs": ([_vp, ctypes.c_double, ctypes.c_int, _vp, ctypes.POINTER(StzStepStats)], ctypes.c_int),
    "stz_multigrid": ([_vp], _vp),
    "stz_adjoint_opts_default": ([ctypes.POINTER(StzAdjointOpts)], None),
    "stz_adjoint": ([_vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(StzAdjointOpts), _dp, _dp, _dp],
                    ctypes.c_int),
    "stz_transform": ([ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp],
                      ctypes.c_int),
    "stz_grf_params_default": ([ctypes.POINTER(StzGrfParams)], None),
    "stz_grf_geometry": ([_ip, _ip, ctypes.POINTER(StzGrfParams), _ip, ctypes.POINTER(ctypes.c_longlong), _dp],
                         ctypes.c_int),
    "stz_grf_cache_name": ([_ip, _ip, ctypes.POINTER(StzGrfParams), ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "stz_grf_field": ([_ip, _ip, ctypes.POINTER(StzGrfParams), _dp, _dp], ctypes.c_int),
    "stz_grf_field_device": ([_ip, _ip, ctypes.POINTER(StzGrfParams), _vp, _vp, _vp], ctypes.c_int),
    "stz_grf_time_kernels": ([_ip, _ip, ctypes.POINTER(StzGrfParams), ctypes.c_int, _dp], ctypes.c_int),
    "stz_grf_last_error": ([], ctypes.c_char_p),
    "stz_set_transform": ([_vp, _dp, _dp], ctypes.c_int),
    "stz_set_random_initial_conditions": ([_vp, ctypes.POINTER(StzGrfParams), _dp, ctypes.c_char_p], ctypes.c_int),
    "stz_cube_size": ([_vp, ctypes.c_int], ctypes.c_longlong),
    "stz_output_time_kernel": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stz_output_cube": ([_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_longlong], ctypes.c_int),
    "stz_write_cube": ([_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "stz_write_files": ([_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "stz_net_force": ([_vp, _dp], ctypes.c_int),
    "stz_write_traction": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "stz_proj_mag": ([_vp], ctypes.c_double),
    "stz_set_output": ([_vp, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "stz_direct_dt": ([_vp], ctypes.c_double),
    "stz_init_ref": ([_vp], ctypes.c_int),
    "stz_set_ref_map": ([_vp, _dp], ctypes.c_int),
    "stz_get_ref_map": ([_vp, _dp], ctypes.c_int),
    "stz_step_direct": ([_vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(StzStepStats)], ctypes.c_int),
    "stz_direct_time_kernels": ([_vp, ctypes.c_double, ctypes.c_int, _dp], ctypes.c_int),
    "stz_l2_comparison": ([_vp, _vp, _dp], ctypes.c_int),
    "stzmg_desc_default": ([ctypes.POINTER(StzmgDesc)] + [ctypes.c_int] * 6, ctypes.c_int),
    "stzmg_create": ([ctypes.POINTER(StzmgDesc), ctypes.POINTER(_vp)], ctypes.c_int),
    "stzmg_destroy": ([_vp], ctypes.c_int),
    "stzmg_last_error": ([_vp], ctypes.c_char_p),
    "stzmg_nlevels": ([_vp], ctypes.c_int),
    "stzmg_level_dims": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
    "stzmg_set_operator_const": ([_vp, _dp, _dp], ctypes.c_int),
    "stzmg_set_operator_planes": ([_vp, _dp], ctypes.c_int),
    "stzmg_set_operator_points": ([_vp, _dp], ctypes.c_int),
    "stzmg_get_level_operator": ([_vp, ctypes.c_int, _dp, _lp], ctypes.c_int),
    "stzmg_set_level_fields": ([_vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "stzmg_get_level_fields": ([_vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "stzmg_set_fields_device": ([_vp, _vp, _vp], ctypes.c_int),
    "stzmg_get_x_device": ([_vp, _vp], ctypes.c_int),
    "stzmg_gauss_seidel": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "stzmg_restriction": ([_vp, ctypes.c_int], ctypes.c_int),
    "stzmg_interpolation": ([_vp, ctypes.c_int], ctypes.c_int),
    "stzmg_residual": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "stzmg_norm": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stzmg_norm_noalloc": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stzmg_vcycle": ([_vp, ctypes.c_int], ctypes.c_int),
    "stzmg_vcycle_fp32": ([_vp, ctypes.c_int], ctypes.c_int),
    "stzmg_gauss_seidel_fp32": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "stzmg_residual_fp32": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "stzmg_solve": ([_vp, ctypes.POINTER(StzmgSolveOpts), ctypes.POINTER(StzmgSolveStats)], ctypes.c_int),
    "stzmg_solve_opts_default": ([ctypes.POINTER(StzmgSolveOpts)], None),
    "stzmg_stream": ([_vp], _vp),
    "stzmg_set_option": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "stzmg_time_kernel": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stzmg_shear_stencil": ([ctypes.c_double] * 7 + [_dp, _dp], ctypes.c_int),
    "stzmg_trans_stencil": ([_dp] + [ctypes.c_double] * 6 + [_dp], ctypes.c_int),
}

_LIB = None


def load():
    """Load libstzmg.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library not built: {LIB_PATH}; run python -c 'import __graft_entry__ as g; g.build()'")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None
