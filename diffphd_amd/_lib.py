"""ctypes binding of libphd.so (the C ABI declared in include/phd.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphd.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_P = ctypes.c_void_p
_cp = ctypes.c_char_p

# (name, restype, argtypes) for every entry point of include/phd.h
SIGNATURES = [
    ("phd_ctx_create", _P, [ctypes.c_int]),
    ("phd_ctx_destroy", None, [_P]),
    ("phd_last_error", _cp, [_P]),
    ("phd_init_mesh", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_double]),
    ("phd_add_pd_energy", ctypes.c_int, [_P, _cp, _dp, ctypes.c_int, _ip, ctypes.c_int]),
    ("phd_set_dirichlet", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
    ("phd_add_state_force", ctypes.c_int, [_P, _cp, _dp, ctypes.c_int]),
    ("phd_set_frictional_boundary", ctypes.c_int, [_P, _cp, _dp, ctypes.c_int, _ip, ctypes.c_int]),
    ("phd_set_press_plane", ctypes.c_int, [_P, ctypes.c_double]),
    ("phd_factor_build", ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int]),
    ("phd_apply_ainv", ctypes.c_int, [_P, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    ("phd_apply_ainv_noalloc", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("phd_step_forward", ctypes.c_int, [_P, _dp, _dp, _dp, ctypes.c_double, ctypes.POINTER(_cp), _dp, ctypes.c_int,
                                        _dp, _dp, _ip, _ip]),
    ("phd_step_backward", ctypes.c_int, [_P, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _ip, ctypes.c_int, _dp, _dp,
                                         ctypes.POINTER(_cp), _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("phd_step_forward_baseline", ctypes.c_int, [_P, _cp, _dp, _dp, _dp, ctypes.c_double, ctypes.POINTER(_cp), _dp,
                                                 ctypes.c_int, _dp, _dp, _ip, ctypes.c_int, _ip, _ip, _ip]),
    ("phd_step_backward_baseline", ctypes.c_int, [_P, _dp, _dp, _dp, ctypes.c_double, _dp, _ip, ctypes.c_int, _dp, _dp,
                                                  ctypes.POINTER(_cp), _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip]),
    ("phd_step_forward_device", ctypes.c_int, [_P, _P, _P, _P, ctypes.c_double, ctypes.POINTER(_cp), _dp, ctypes.c_int,
                                               _P, _P, _ip, _ip]),
    ("phd_step_backward_device", ctypes.c_int, [_P, _P, _P, _P, ctypes.c_double, _P, _P, _ip, ctypes.c_int, _P, _P,
                                                ctypes.POINTER(_cp), _dp, ctypes.c_int, _P, _P, _P, _dp, _dp, _dp]),
    ("phd_dofs", ctypes.c_int, [_P]),
    ("phd_num_elements", ctypes.c_int, [_P]),
    ("phd_num_pd_element_energies", ctypes.c_int, [_P]),
    ("phd_add_actuation", ctypes.c_int, [_P, ctypes.c_double, _dp, _ip, ctypes.c_int]),
    ("phd_set_actuation", ctypes.c_int, [_P, _dp, ctypes.c_int]),
    ("phd_act_dofs", ctypes.c_int, [_P]),
    ("phd_num_state_force_parameters", ctypes.c_int, [_P]),
    ("phd_num_pd_muscle_energies", ctypes.c_int, [_P]),
    ("phd_actuation_gradients", ctypes.c_int, [_P, _dp, ctypes.c_int, _dp, ctypes.c_int]),
    ("phd_s_memory_bytes", ctypes.c_longlong, [_P]),
    ("phd_s_nnz", ctypes.c_longlong, [_P]),
    ("phd_cached_n", ctypes.c_int, [_P]),
    ("phd_frame_stats", ctypes.c_int, [_P, _ip, _ip, _dp, ctypes.c_int]),
    ("phd_reset_frame_stats", None, [_P]),
    ("phd_reset_contact_history", ctypes.c_int, [_P]),
    ("phd_last_backward_iters", ctypes.c_int, [_P]),
    ("phd_last_backward_stats", ctypes.c_int, [_P, _dp, ctypes.c_int]),
    ("phd_factor_info", ctypes.c_int, [_P, _dp, ctypes.c_int]),
    ("phd_autotune", ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]),
    ("phd_set_launch_shapes", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("phd_timing_enable", ctypes.c_int, [_P, ctypes.c_int]),
    ("phd_timing_read", ctypes.c_int, [_P, _dp, ctypes.c_int]),
    ("phd_phase_times", ctypes.c_int, [_P, _dp, ctypes.c_int, ctypes.c_int]),
    ("phd_last_contact_multipliers", ctypes.c_int, [_P, _dp, _dp, _dp, ctypes.c_int]),
    ("phd_dense_spd_solve", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    ("phd_mlp_num_params", ctypes.c_int, []),
    ("phd_mlp_init", ctypes.c_int, [_P, ctypes.c_int]),
    ("phd_mlp_set_normalization", ctypes.c_int, [_P, _fp, _fp, ctypes.c_float]),
    ("phd_mlp_get_params", ctypes.c_int, [_P, _fp]),
    ("phd_mlp_set_params", ctypes.c_int, [_P, _fp]),
    ("phd_mlp_forward", ctypes.c_int, [_P, ctypes.c_int, _fp, _fp]),
    ("phd_mlp_loss_grad", ctypes.c_int, [_P, ctypes.c_int, _fp, _fp, _fp, _dp]),
    ("phd_mlp_train", ctypes.c_int, [_P, ctypes.c_int, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, _fp]),
    ("phd_mlp_features", ctypes.c_int, [_P, _dp, _dp, _dp, ctypes.c_double, _fp]),
    ("phd_warm_start_stats", ctypes.c_int, [_P, _ip, _ip, ctypes.c_int]),
    ("phd_mlp_mfma_probe", ctypes.c_int, [_fp, _fp, _fp]),
    ("phd_copy_probe", ctypes.c_int, [_P, ctypes.c_longlong, ctypes.c_int, _dp]),
    ("phd_host_factor_check", ctypes.c_int, [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]),
]


def lib():
    """Loads libphd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libphd.so not found at {LIB_PATH}; build it with `python -m diffphd_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def fptr(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def iptr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def options(opt):
    keys = list(opt.keys())
    karr = (_cp * max(1, len(keys)))(*[k.encode() for k in keys])
    vals, vp = dptr([float(opt[k]) for k in keys] or [0.0])
    return karr, vals, vp, len(keys)
