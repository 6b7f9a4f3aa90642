"""ctypes binding of the C ABI declared in include/stz_amd.h.

The shared library ``stz_amd/libstz_amd.so`` is built in-tree by
``stz_amd/build.py``.  There is no fallback: if the library is missing or no
device is visible, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstz_amd.so")

# status codes
OK, ERR_ARG, ERR_HIP, ERR_NOT_CONVERGED, ERR_NAN, ERR_NO_DEVICE, ERR_STATE = range(7)
PREC_FP64, PREC_MIXED = 0, 1
NFIELDS = 10
FIELD_NAMES = ("u", "v", "p", "q", "s", "tau", "chi", "X", "Y", "ddev")
F_CU = 10  # the c_field::cu slot: output plane of the diagnostics
FIELD_INDEX = dict({k: i for i, k in enumerate(FIELD_NAMES)}, cu=F_CU)

_MODEL_DOUBLES = ("TZ", "c0", "tau0", "kappa", "Delta", "Omegaeps0", "chi_inf",
                  "theta", "rho", "t_icpt", "t_grad")
_SIM_DOUBLES = ("ax", "bx", "ay", "by", "mu", "K", "viscosity", "chi_diff",
                "chi_len", "t_scale", "tmult", "adapt_fac", "u_bdry")


class ModelParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("t_model", C.c_int32)] + \
        [(k, C.c_double) for k in _MODEL_DOUBLES]


class SimParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32)] + \
        [(k, C.c_double) for k in _SIM_DOUBLES] + [("stz", ModelParams)]


class MGConfig(C.Structure):
    _fields_ = [("nu1", C.c_int32), ("nu2", C.c_int32), ("bottom_sweeps", C.c_int32),
                ("extra_cycles", C.c_int32), ("predict_mult", C.c_int32),
                ("predict_decay", C.c_int32), ("predict_init", C.c_int32),
                ("predict_max_iters", C.c_int32), ("tol_mult", C.c_double),
                ("max_refine", C.c_int32), ("kernels", C.c_int32), ("direct_nodes", C.c_int32),
                ("predict_unit", C.c_int32), ("reserved", C.c_int32 * 4)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("cycles", C.c_int32), ("total_cycles", C.c_int32),
                ("refinements", C.c_int32), ("residual", C.c_double), ("tolerance", C.c_double)]


class StzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("stz_amd error %d: %s" % (code, msg))
        self.code = code


def make_params(d):
    p = SimParams()
    p.m, p.n = int(d["m"]), int(d["n"])
    for k in _SIM_DOUBLES:
        setattr(p, k, float(d[k]))
    mp = d["stz"]
    p.stz.kind = int(mp.get("kind", 0))
    p.stz.t_model = int(mp.get("t_model", -1))
    for k in _MODEL_DOUBLES:
        setattr(p.stz, k, float(mp[k]))
    return p


_BD_DOUBLES = ("ax", "bx", "ay", "by", "mu", "K", "viscosity", "chi_len",
               "filter_stress", "t_scale", "tmult", "adapt_fac")


class BdParams(C.Structure):
    """stz_bd_params: bd_sim constructor arguments (bd_sim.hh:147-151)."""
    _fields_ = [("m", C.c_int32), ("n", C.c_int32)] + \
        [(k, C.c_double) for k in _BD_DOUBLES] + [("stz", ModelParams)]


def make_bd_params(d):
    p = BdParams()
    p.m, p.n = int(d["m"]), int(d["n"])
    for k in _BD_DOUBLES:
        setattr(p, k, float(d[k]))
    mp = d["stz"]
    p.stz.kind = int(mp.get("kind", 0))
    p.stz.t_model = int(mp.get("t_model", -1))
    for k in _MODEL_DOUBLES:
        setattr(p.stz, k, float(mp[k]))
    return p


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "stz_last_error": [],
    "stz_device_count": [C.POINTER(C.c_int)],
    "stz_sim_create": [C.POINTER(SimParams), C.c_int, C.POINTER(C.c_void_p)],
    "stz_sim_destroy": [C.c_void_p],
    "stz_sim_set_stream": [C.c_void_p, C.c_void_p],
    "stz_sim_synchronize": [C.c_void_p],
    "stz_sim_upload_field": [C.c_void_p, C.c_int, _dp],
    "stz_sim_download_field": [C.c_void_p, C.c_int, _dp],
    "stz_sim_field_device_ptr": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "stz_sim_upload_vel": [C.c_void_p, _dp],
    "stz_sim_download_vel": [C.c_void_p, _dp],
    "stz_sim_download_src": [C.c_void_p, _dp],
    "stz_sim_vel_device_ptr": [C.c_void_p, C.POINTER(C.c_void_p)],
    "stz_sim_set_time": [C.c_void_p, C.c_double],
    "stz_sim_get_time": [C.c_void_p, C.POINTER(C.c_double)],
    "stz_sim_set_chi_inf": [C.c_void_p, C.c_double],
    "stz_sim_init_fields": [C.c_void_p, C.c_int, C.c_double, C.c_double],
    "stz_sim_step_quasistatic": [C.c_void_p, C.c_double, C.c_int],
    "stz_sim_advection_step": [C.c_void_p, C.c_double],
    "stz_sim_projection_step": [C.c_void_p, C.c_double],
    "stz_sim_set_boundaries": [C.c_void_p],
    "stz_sim_get_mg_config": [C.c_void_p, C.POINTER(MGConfig)],
    "stz_sim_set_mg_config": [C.c_void_p, C.POINTER(MGConfig)],
    "stz_sim_last_solve": [C.c_void_p, C.POINTER(SolveInfo)],
    "stz_sim_avg_iters": [C.c_void_p, C.POINTER(C.c_double)],
    "stz_sim_mg_hierarchy": [C.c_void_p, C.c_int, C.POINTER(C.c_int), _ip],
    "stz_mg_setup": [C.c_void_p, C.c_double],
    "stz_mg_solve": [C.c_void_p, C.POINTER(SolveInfo)],
    "stz_mg_v_cycle": [C.c_void_p],
    "stz_mg_l2_error": [C.c_void_p, C.POINTER(C.c_double)],
    "stz_mg_upload_src": [C.c_void_p, _dp],
    "stz_mg_smooth": [C.c_void_p, C.c_int, C.c_int],
    "stz_sim_step_forward": [C.c_void_p, C.c_double, C.c_int],
    "stz_sim_tractions": [C.c_void_p, _dp],
    "stz_sim_qs_measure": [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "stz_sim_compute_strain": [C.c_void_p, C.c_int],
    "stz_sim_l2_comparison": [C.c_void_p, C.c_void_p, _dp],
    "stz_sim_initialize_random": [C.c_void_p, C.c_double, C.c_double, C.c_double],
    "stz_sim_initialize_chi_bicubic": [C.c_void_p, C.c_int, C.c_int, _dp, C.c_double, C.c_double],
    "stz_sim_record_decisions": [C.c_void_p, C.c_int],
    "stz_sim_download_decisions": [C.c_void_p, _up],
    "stz_sim_replay_decisions": [C.c_void_p, C.c_int, _up],
    "stz_sim_get_predictor": [C.c_void_p, C.POINTER(C.c_int32)],
    "stz_sim_set_predictor": [C.c_void_p, C.c_int32],
    "stz_bd_create": [C.POINTER(BdParams), C.POINTER(C.c_void_p)],
    "stz_bd_destroy": [C.c_void_p],
    "stz_bd_set_stream": [C.c_void_p, C.c_void_p],
    "stz_bd_synchronize": [C.c_void_p],
    "stz_bd_set_level_set": [C.c_void_p, _dp],
    "stz_bd_upload_level_set": [C.c_void_p, _dp, _ip],
    "stz_bd_init_bar": [C.c_void_p, C.c_int],
    "stz_bd_upload_cells": [C.c_void_p, C.c_int, _dp],
    "stz_bd_download_cells": [C.c_void_p, C.c_int, _dp],
    "stz_bd_upload_nodes": [C.c_void_p, C.c_int, _dp],
    "stz_bd_download_nodes": [C.c_void_p, C.c_int, _dp],
    "stz_bd_download_vel": [C.c_void_p, _dp],
    "stz_bd_download_src": [C.c_void_p, _dp],
    "stz_bd_download_status": [C.c_void_p, _ip, _ip],
    "stz_bd_set_time": [C.c_void_p, C.c_double],
    "stz_bd_get_time": [C.c_void_p, C.POINTER(C.c_double)],
    "stz_bd_advection_step": [C.c_void_p, C.c_double],
    "stz_bd_projection_step": [C.c_void_p, C.c_double],
    "stz_bd_post_process": [C.c_void_p, C.c_int],
    "stz_bd_step_quasistatic": [C.c_void_p, C.c_double, C.c_int],
    "stz_bd_last_solve": [C.c_void_p, C.POINTER(SolveInfo)],
    "stz_bd_interface_velocity": [C.c_void_p, _dp],
    "stz_bd_step_stats": [C.c_void_p, _ip],
    "stz_bd_remarch": [C.c_void_p],
    "stz_bd_set_solver": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
    "stz_bd_set_precision": [C.c_void_p, C.c_int],
    "stz_bd_operator": [C.c_void_p, C.c_double, _ip, _dp, _dp],
    "stz_bd_set_body_force": [C.c_void_p, C.c_void_p],
    "stz_bd_set_edge_velocity": [C.c_void_p, C.c_double, C.c_double],
    "stz_ensemble_create": [C.POINTER(SimParams), C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    "stz_ensemble_destroy": [C.c_void_p],
    "stz_ensemble_size": [C.c_void_p, C.POINTER(C.c_int)],
    "stz_ensemble_member": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "stz_ensemble_set_mg_config": [C.c_void_p, C.POINTER(MGConfig)],
    "stz_ensemble_use_graphs": [C.c_void_p, C.c_int],
    "stz_ensemble_step_quasistatic": [C.c_void_p, C.c_double, C.c_int],
    "stz_ensemble_stats": [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")],
    "stz_ensemble_tile_choice": [C.c_void_p, _ip],
    "stz_sim_set_model": [C.c_void_p, C.POINTER(ModelParams)],
    "stz_adjoint_create": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "stz_adjoint_destroy": [C.c_void_p],
    "stz_adjoint_forward": [C.c_void_p, C.c_double, C.c_int],
    "stz_adjoint_backward": [C.c_void_p, _dp, _dp, _dp],
    "stz_adjoint_set_solve_accuracy": [C.c_void_p, C.c_double],
    "stz_adjoint_stats": [C.c_void_p, _dp],
    "stz_model_dplastic": [C.POINTER(ModelParams), C.c_double, C.c_double, C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "stz_sim_advection_launch": [C.c_void_p, C.c_double, C.c_int],
    "stz_mg_fmg_pre": [C.c_void_p],
    "stz_mg_clear_z": [C.c_void_p],
    "stz_mg_vcycle_launch": [C.c_void_p, C.c_int, C.c_int, C.c_int],
    "stz_qs_l2_error_noalloc": [C.POINTER(SimParams), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p],
}
_RESTYPE = {"stz_last_error": C.c_char_p}


def load():
    """Load and bind the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("stz_amd: %s is not built (run python stz_amd/build.py)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    return tuple(_SIGS)


def check(code):
    if code != OK:
        raise StzError(code, load().stz_last_error().decode(errors="replace"))
    return code


def device_count():
    n = C.c_int(0)
    load().stz_device_count(C.byref(n))
    return n.value
