"""ctypes binding of lbm_amd/liblbm_hip.so (declared in include/lbm_hip.h).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly if the shared library has not been built (``python -m lbm_amd.build``).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

LBM_OK = 0
LBM_EINVAL = -1
LBM_EHIP = -2
LBM_ENOMEM = -3
LBM_ENONFINITE = -4
LBM_ENODEV = -5

NUM_VARIANTS = 10
BATCH_NUM_VARIANTS = 8

PRECISION_F32 = 0
PRECISION_F16S = 1
PRECISION_F16E = 2  # binary16 of f - feq(1, (u_in, 0)): centred on the free stream
BOUNDARY_REFERENCE = 0
BOUNDARY_ZERO_GRADIENT = 1


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class lbm_config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32),
        ("ny", ctypes.c_uint32),
        ("omega", ctypes.c_float),
        ("u_init", ctypes.c_float),
        ("precision", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("row_pad", ctypes.c_uint32),
    ]


class lbm_domain(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32),
        ("ny", ctypes.c_uint32),
        ("omega", ctypes.c_float),
        ("u_in", ctypes.c_float),
    ]


_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_u32, _u64, _i32, _f32, _sz = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)
_bp = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> (restype, argtypes); every symbol here is declared in include/lbm_hip.h
SIGNATURES = {
    "lbm_config_default": (_i32, [ctypes.POINTER(lbm_config), _u32, _u32]),
    "lbm_create": (_i32, [ctypes.POINTER(lbm_config), _PP]),
    "lbm_destroy": (_i32, [_P]),
    "lbm_last_error": (ctypes.c_char_p, [_P]),
    "lbm_set_omega": (_i32, [_P, _f32]),
    "lbm_set_viscosity": (_i32, [_P, _f32]),
    "lbm_get_omega": (_i32, [_P, _fp]),
    "lbm_reset_equilibrium": (_i32, [_P, _f32, _f32, _f32]),
    "lbm_single_cell": (_i32, [_P, _i32]),
    "lbm_set_pulse": (_i32, [_P, _i32, _u32, _u32, _f32]),
    "lbm_reset_barrier": (_i32, [_P]),
    "lbm_paint": (_i32, [_P, _up, _sz]),
    "lbm_set_mask": (_i32, [_P, _bp]),
    "lbm_get_mask": (_i32, [_P, _bp]),
    "lbm_step": (_i32, [_P, _u32]),
    "lbm_sync": (_i32, [_P]),
    "lbm_step_timed": (_i32, [_P, _u32, _fp]),
    "lbm_set_variant": (_i32, [_P, _i32]),
    "lbm_get_variant": (_i32, [_P, ctypes.POINTER(ctypes.c_int)]),
    "lbm_autotune": (_i32, [_P, _u32, ctypes.POINTER(ctypes.c_int), _fp]),
    "lbm_autotune_cached": (_i32, [_P, ctypes.c_char_p, _u32, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    "lbm_step_count": (_i32, [_P, _u64p]),
    "lbm_get_stream": (_i32, [_P, _PP]),
    "lbm_read_moments": (_i32, [_P, _fp, _fp, _fp, _i32]),
    "lbm_read_fields": (_i32, [_P, _fp, _fp, _fp, _fp, _i32]),
    "lbm_summary": (_i32, [_P, _i32, _fp]),
    "lbm_render": (_i32, [_P, _i32, _i32, _fp]),
    "lbm_conservation": (_i32, [_P, _dp]),
    "lbm_conservation_static": (_i32, [_P, _dp]),
    "lbm_read_f": (_i32, [_P, _fp]),
    "lbm_write_f": (_i32, [_P, _fp]),
    "lbm_state_size": (_i32, [_P, ctypes.POINTER(_sz)]),
    "lbm_save_state": (_i32, [_P, _P, _sz]),
    "lbm_load_state": (_i32, [_P, _P, _sz]),
    "lbm_batch_create": (_i32, [ctypes.POINTER(lbm_domain), _u32, _i32, _i32, _i32, _PP]),
    "lbm_batch_destroy": (_i32, [_P]),
    "lbm_batch_last_error": (ctypes.c_char_p, [_P]),
    "lbm_batch_reset": (_i32, [_P]),
    "lbm_batch_set_mask": (_i32, [_P, _u32, _bp]),
    "lbm_batch_step": (_i32, [_P, _u32]),
    "lbm_batch_step_timed": (_i32, [_P, _u32, _fp]),
    "lbm_batch_sync": (_i32, [_P]),
    "lbm_batch_step_count": (_i32, [_P, _u64p]),
    "lbm_batch_info": (_i32, [_P, _u64p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "lbm_batch_set_graph": (_i32, [_P, _i32]),
    "lbm_batch_set_variant": (_i32, [_P, _i32]),
    "lbm_batch_get_variant": (_i32, [_P, ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_autotune": (_i32, [_P, _u32, ctypes.POINTER(ctypes.c_int), _fp]),
    "lbm_batch_autotune_cached": (_i32, [_P, ctypes.c_char_p, _u32, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_read_f": (_i32, [_P, _u32, _fp]),
    "lbm_batch_read_moments": (_i32, [_P, _u32, _fp, _fp, _fp]),
    "lbm_gen_batch": (_i32, [_u64, _u32, ctypes.POINTER(lbm_domain)]),
    "lbm_gen_shapes": (_i32, [_bp, _u32, _u32, _u64, _u32, _u32]),
    "lbm_unet_last_error": (ctypes.c_char_p, []),
    "lbm_unet_create": (_i32, [_i32, _i32, _i32, _u64, _i32, _PP]),
    "lbm_unet_destroy": (_i32, [_P]),
    "lbm_unet_num_params": (_i32, [_P, ctypes.POINTER(_sz)]),
    "lbm_unet_get_params": (_i32, [_P, _fp]),
    "lbm_unet_set_params": (_i32, [_P, _fp]),
    "lbm_unet_forward": (_i32, [_P, _fp, _fp]),
    "lbm_unet_loss_grad": (_i32, [_P, _fp, _fp, _fp, _fp, _fp]),
    "lbm_unet_train_step": (_i32, [_P, _fp, _fp, _fp, _f32, _fp]),
    "lbm_unet_time_forward": (_i32, [_P, _i32, _fp]),
    "lbm_unet_set_graph": (_i32, [_P, _i32]),
    "lbm_surrogate_jump": (_i32, [_P, _P, _u32, _f32, _f32, _f32, _f32]),
    "lbm_snapshot": (_i32, [_P]),
    "lbm_restore": (_i32, [_P]),
    "lbm_velocity_mark": (_i32, [_P]),
    "lbm_velocity_change": (_i32, [_P, _dp]),
    "lbm_velocity_mark_slot": (_i32, [_P, _i32]),
    "lbm_velocity_skill": (_i32, [_P, _dp]),
    "lbm_conv3x3_test": (_i32, [_i32, _i32, _i32, _i32, _i32, _i32, _fp, _fp, _fp, _fp, _fp, _i32]),
    "lbm_f64_last_error": (ctypes.c_char_p, []),
    "lbm_f64_create": (_i32, [_u32, _u32, ctypes.c_double, ctypes.c_double, _i32, _i32, _PP]),
    "lbm_f64_destroy": (_i32, [_P]),
    "lbm_f64_set_mask": (_i32, [_P, _bp]),
    "lbm_f64_step": (_i32, [_P, _u32]),
    "lbm_f64_sync": (_i32, [_P]),
    "lbm_f64_step_count": (_i32, [_P, _u64p]),
    "lbm_f64_macro": (_i32, [_P, _dp, _dp, _dp]),
    "lbm_f64_read_f": (_i32, [_P, _dp]),
    "lbm_splitmix64": (_u64, [_u64p]),
    "lbm_gen_circles": (_i32, [_bp, _u32, _u32, _u64, _u32, _u32, _u32]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the gfx950 library and bind every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m lbm_amd.build` "
            "(the product has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, handle=None):
    if code != LBM_OK:
        msg = load().lbm_last_error(handle)
        raise LbmError(code, msg.decode() if msg else "")
    return code


def check_unet(code: int):
    if code != LBM_OK:
        msg = load().lbm_unet_last_error()
        raise LbmError(code, msg.decode() if msg else "")
    return code


def check_f64(code: int):
    if code != LBM_OK:
        msg = load().lbm_f64_last_error()
        raise LbmError(code, msg.decode() if msg else "")
    return code


def check_batch(code: int, handle=None):
    if code != LBM_OK:
        msg = load().lbm_batch_last_error(handle)
        raise LbmError(code, msg.decode() if msg else "")
    return code
