"""ctypes binding of libubg.so (include/ubg_c.h).

The library is the product: every compute call goes through it.  There is no
fallback -- if the shared library is missing, importing this module's `lib()`
raises, and on a host without a GPU the handle constructors report
UBG_E_NO_DEVICE.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libubg.so")

# status codes
OK = 0
E_INVALID, E_DOMAIN_TOO_SMALL, E_FLAG_VALUE, E_HIP, E_OOM, E_NONFINITE, E_BUDGET, \
    E_NOT_CONVERGED, E_STATE, E_NO_DEVICE = range(1, 11)
ERROR_NAMES = {
    1: "UBG_E_INVALID", 2: "UBG_E_DOMAIN_TOO_SMALL", 3: "UBG_E_FLAG_VALUE", 4: "UBG_E_HIP",
    5: "UBG_E_OOM", 6: "UBG_E_NONFINITE", 7: "UBG_E_BUDGET", 8: "UBG_E_NOT_CONVERGED",
    9: "UBG_E_STATE", 10: "UBG_E_NO_DEVICE",
}

# field ids
(FIELD_VX, FIELD_VX_BACK, FIELD_VY, FIELD_VY_BACK, FIELD_VX_ACCUM, FIELD_VY_ACCUM,
 FIELD_VX_CURRENT, FIELD_VY_CURRENT, FIELD_P, FIELD_F, FIELD_FLAG, FIELD_R) = range(12)
FIELD_NAMES = ["vx", "vx_back", "vy", "vy_back", "vx_accum", "vy_accum", "vx_current",
               "vy_current", "p", "f", "flag", "r"]
# stage ids
(STAGE_ACCUM, STAGE_DIFFUSE, STAGE_ADVECT, STAGE_SET_VBCS, STAGE_PROJECT, STAGE_SET_PBC,
 STAGE_SAVE_CURRENT, STAGE_DIVERGENCE, STAGE_CORRECT_VELOCITY, STAGE_MG_SOLVE) = range(10)
# MG level operations
MGOP_RBGS, MGOP_ZERO_GRADIENT, MGOP_RESIDUAL, MGOP_RESTRICT_RESIDUAL, MGOP_PROLONG_CORRECT = range(5)
LEVEL_P, LEVEL_F, LEVEL_R = range(3)
TIMER_SMOOTH_L0, TIMER_L0_PRE, TIMER_L0_POST, TIMER_ADVECT = range(4)
FLAG_NEAREST, FLAG_LINEAR, FLAG_NORMAL = range(3)


# ubg_item (components.hpp:6-42): CoItem + CoKinematics
ITEM_DTYPE = np.dtype([("size", "<f4", 2), ("pos", "<f4", 2), ("rotation", "<f4"), ("mass", "<f4"),
                       ("vel", "<f4", 2), ("force", "<f4", 2), ("ang_vel", "<f4"),
                       ("ang_force", "<f4"), ("bump_count", "<i4")])


class UbgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nu_pre", C.c_int), ("nu_post", C.c_int), ("nu_coarse", C.c_int),
                ("vcycles_per_step", C.c_int), ("alpha", C.c_float), ("use_graph", C.c_int),
                ("save_current", C.c_int)]


IR_HISTORY = 64


class IRParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("vcycles_per_iter", C.c_int), ("max_iter", C.c_int),
                ("ref_norm", C.c_double)]


class IRResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("vcycles", C.c_int), ("converged", C.c_int),
                ("r0_norm", C.c_double), ("rel_residual", C.c_double),
                ("history", C.c_double * IR_HISTORY)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_w", C.c_int), ("tile_h", C.c_int), ("blocks", C.c_int),
                ("pass_rows", C.c_int), ("sweep_ms", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("tol", C.c_double),
                ("vcycles_per_iter", C.c_int), ("max_iter", C.c_int),
                ("steps_per_turn", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("state", C.c_int), ("timesteps_done", C.c_int),
                ("iterations", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_long), ("completed", C.c_long), ("failed", C.c_long),
                ("solves", C.c_long), ("steps", C.c_long), ("active", C.c_int),
                ("queued", C.c_int), ("peak_active", C.c_int), ("bytes_in_use", C.c_size_t),
                ("peak_bytes", C.c_size_t), ("busy_s", C.c_double), ("setup_s", C.c_double),
                ("wait_s", C.c_double), ("retire_s", C.c_double), ("admit_s", C.c_double)]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

# name -> (restype, argtypes); the list is also the set of symbols the CPU
# tests check the library exports.
SIGNATURES = {
    "ubg_version": (C.c_int, []),
    "ubg_last_error": (C.c_char_p, []),
    "ubg_params_default": (None, [C.POINTER(Params)]),
    "ubg_device_count": (C.c_int, [_ip]),
    "ubg_mg_level_sizes": (C.c_int, [C.c_int, C.c_int, _ip, _ip, C.c_int]),
    "ubg_mg_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "ubg_mg_destroy": (C.c_int, [_vp]),
    "ubg_mg_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "ubg_mg_num_levels": (C.c_int, [_vp, _ip]),
    "ubg_mg_update_fields": (C.c_int, [_vp, _fp]),
    "ubg_mg_set_level_field": (C.c_int, [_vp, C.c_int, C.c_int, _fp]),
    "ubg_mg_get_level_field": (C.c_int, [_vp, C.c_int, C.c_int, _fp]),
    "ubg_mg_get_level_flag": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_mg_level_op": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, C.c_float]),
    "ubg_mg_solve": (C.c_int, [_vp, C.c_float, C.c_int, C.c_int]),
    "ubg_mg_solve_host": (C.c_int, [_vp, _fp, _fp, _fp, C.c_float, C.c_int]),
    "ubg_mg_rbgs": (C.c_int, [_vp, C.c_float, C.c_float, C.c_int]),
    "ubg_mg_gs": (C.c_int, [_vp, C.c_float, C.c_float, C.c_int]),
    "ubg_mg_solve_device_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_float,
                                              C.c_int, C.c_int, C.POINTER(Params), _vp]),
    "ubg_mg_residual": (C.c_int, [_vp, C.c_float, _fp, _dp]),
    "ubg_mg_sync": (C.c_int, [_vp]),
    "ubg_ir_params_default": (None, [C.POINTER(IRParams)]),
    "ubg_mg_ir_set_rhs": (C.c_int, [_vp, _fp]),
    "ubg_mg_ir_set_guess": (C.c_int, [_vp, _dp]),
    "ubg_mg_ir_solve": (C.c_int, [_vp, C.c_float, C.POINTER(IRParams), C.POINTER(IRResult)]),
    "ubg_mg_ir_get_solution": (C.c_int, [_vp, _dp]),
    "ubg_mg_ir_device_ptrs": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_fp)]),
    "ubg_batch_bytes": (C.c_int, [C.c_int, _ip, _ip, C.POINTER(C.c_size_t)]),
    "ubg_batch_create": (C.c_int, [C.c_int, _ip, _ip, C.POINTER(Params), C.POINTER(_vp)]),
    "ubg_batch_destroy": (C.c_int, [_vp]),
    "ubg_batch_set_domain": (C.c_int, [_vp, C.c_int, _fp, _fp, C.c_float]),
    "ubg_batch_set_rhs": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_batch_set_guess": (C.c_int, [_vp, C.c_int, _dp]),
    "ubg_batch_solve": (C.c_int, [_vp, C.POINTER(IRParams), _ip]),
    "ubg_batch_result": (C.c_int, [_vp, C.c_int, C.POINTER(IRResult)]),
    "ubg_batch_get_solution": (C.c_int, [_vp, C.c_int, _dp]),
    "ubg_batch_autotune": (C.c_int, [_vp, C.c_char_p, C.POINTER(BatchTuning)]),
    "ubg_batch_set_tuning": (C.c_int, [_vp, C.POINTER(BatchTuning)]),
    "ubg_batch_get_tuning": (C.c_int, [_vp, C.POINTER(BatchTuning)]),
    "ubg_batch_set_projection_domain": (C.c_int, [_vp, C.c_int, _fp, _fp, C.c_float]),
    "ubg_batch_set_pressure": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_batch_get_pressure": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_batch_vcycles": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ubg_batch_create_for_sims": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "ubg_sims_step": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _fp]),
    "ubg_flag_components": (C.c_int, [_fp, C.c_int, C.c_int, _ip, _ip]),
    "ubg_remove_component_means": (C.c_int, [_fp, C.c_int, C.c_int, _fp, _ip]),
    "ubg_server_params_default": (None, [C.POINTER(ServerParams)]),
    "ubg_server_create": (C.c_int, [C.POINTER(ServerParams), C.POINTER(Params), C.POINTER(_vp)]),
    "ubg_server_destroy": (C.c_int, [_vp]),
    "ubg_server_add_geometry": (C.c_int, [_vp, C.c_int, C.c_int, _fp, C.c_float, _ip]),
    "ubg_server_submit": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_fp), _dp, _dp,
                                    C.POINTER(C.c_uint64)]),
    "ubg_server_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, C.POINTER(RequestInfo)]),
    "ubg_server_drain": (C.c_int, [_vp]),
    "ubg_server_stats_get": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "ubg_server_clock": (C.c_double, [_vp]),
    "ubg_sim_create": (C.c_int, [_fp, C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(_vp)]),
    "ubg_sim_destroy": (C.c_int, [_vp]),
    "ubg_sim_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "ubg_sim_get_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "ubg_sim_dims": (C.c_int, [_vp, _ip, _ip, _fp, _fp, _fp]),
    "ubg_sim_field_shape": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "ubg_sim_set_field": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_sim_get_field": (C.c_int, [_vp, C.c_int, _fp]),
    "ubg_sim_field_device_ptr": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "ubg_sim_step": (C.c_int, [_vp, C.c_float]),
    "ubg_sim_run": (C.c_int, [_vp, C.c_float, C.c_int]),
    "ubg_sim_stage": (C.c_int, [_vp, C.c_int, C.c_float]),
    "ubg_sim_set_grids": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float]),
    "ubg_sim_set_bcs": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ubg_sim_get_bcs": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "ubg_sim_add_sink": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float]),
    "ubg_sim_set_sinks": (C.c_int, [_vp, _fp, C.c_int]),
    "ubg_sim_get_sinks": (C.c_int, [_vp, _fp, C.c_int, _ip]),
    "ubg_sim_get_mg": (C.c_int, [_vp, C.POINTER(_vp)]),
    "ubg_sim_bilinear_vel": (C.c_int, [_vp, _fp, _fp, C.c_int]),
    "ubg_sim_create_blank": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(_vp)]),
    "ubg_sim_sample_flag": (C.c_int, [_vp, C.c_int, _fp, _fp, C.c_int]),
    "ubg_sim_get_dt": (C.c_int, [_vp, _fp]),
    "ubg_sim_center_p": (C.c_int, [_vp]),
    "ubg_sim_collocate": (C.c_int, [_vp, _fp, _fp, C.c_int]),
    "ubg_sim_shift_map": (C.c_int, [_vp, _fp]),
    "ubg_sim_advect_items": (C.c_int, [_vp, _vp, C.c_int, C.c_float, C.c_int]),
    "ubg_terrain_flags": (C.c_int, [C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int, _fp, _fp]),
    "ubg_sim_set_tracers": (C.c_int, [_vp, _fp, _fp, C.c_int]),
    "ubg_sim_get_tracers": (C.c_int, [_vp, _fp, _fp, C.c_int]),
    "ubg_sim_advect_tracers": (C.c_int, [_vp, C.c_float, C.c_uint32]),
    "ubg_sim_sync": (C.c_int, [_vp]),
    "ubg_sim_timer_enable": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ubg_sim_timer_read": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_long)]),
}

_lib = None


def lib():
    """Load libubg.so (built in-tree by __graft_entry__.build()); no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libubg.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        msg = lib().ubg_last_error()
        raise UbgError(rc, msg.decode() if msg else "")
    return rc


def fptr(a):
    """float* of a C-contiguous float32 numpy array (None for None)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, "need C-contiguous float32"
    return a.ctypes.data_as(_fp)


def device_count():
    n = C.c_int(0)
    check(lib().ubg_device_count(C.byref(n)))
    return n.value


def default_params():
    p = Params()
    lib().ubg_params_default(C.byref(p))
    return p
