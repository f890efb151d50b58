"""ctypes binding of libcsx.so (the C ABI declared in include/csx.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the HIP library is missing, importing anything that needs it
raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcsx.so")

CSX_OK = 0
CSX_FP64, CSX_FP32, CSX_BF16X3, CSX_BF16X6 = 0, 1, 2, 3
PRECISIONS = {"fp64": CSX_FP64, "fp32": CSX_FP32, "bf16x3": CSX_BF16X3, "bf16x6": CSX_BF16X6}

_d, _i, _p, _u = C.c_double, C.c_int, C.c_void_p, C.c_uint
_sz, _u64 = C.c_size_t, C.c_uint64


class Grid(C.Structure):
    _fields_ = [("n", _i), ("x0", _d), ("dx", _d)]


class ChainInfo(C.Structure):
    _fields_ = [("n_sources", _i), ("n_lambdas", _i), ("n_candidates", _i),
                ("in_rows", _i), ("in_cols", _i), ("out_rows", _i), ("out_cols", _i),
                ("out_pixel_size", _d), ("lambda_over_d", _d)]


class Report(C.Structure):
    _fields_ = [("calib_peak", _d), ("n_region_pixels", _i),
                ("seconds_calibration", _d), ("seconds_propagation", _d)]


class Region(C.Structure):
    _fields_ = [("radius1", _d), ("radius2", _d), ("angle1", _d), ("angle2", _d),
                ("xlim1", _d), ("xlim2", _d), ("reference_lambda", _d)]


# name -> (restype, argtypes); every symbol include/csx.h declares.
SIGNATURES = {
    "csx_ctx_create": (_i, [_i, _u, C.POINTER(_p)]),
    "csx_ctx_destroy": (None, [_p]),
    "csx_ctx_set_stream": (_i, [_p, _p]),
    "csx_ctx_synchronize": (_i, [_p]),
    "csx_last_error": (C.c_char_p, []),
    "csx_version": (C.c_char_p, []),
    "csx_tune_report": (_i, [C.c_char_p, _sz]),
    "csx_chain_from_script": (_i, [_p, C.c_char_p, C.POINTER(_p)]),
    "csx_chain_destroy": (None, [_p]),
    "csx_chain_set_array": (_i, [_p, C.c_char_p, C.c_char_p, _p, _i, _i]),
    "csx_chain_set_complex_array": (_i, [_p, C.c_char_p, C.c_char_p, _p, _i, _i, _i]),
    "csx_chain_set_opt_data": (_i, [_p, C.c_char_p, C.c_char_p, _p, _sz]),
    "csx_chain_get_opt_data": (_i, [_p, C.c_char_p, C.c_char_p, _p, C.POINTER(_sz)]),
    "csx_chain_set_candidates": (_i, [_p, C.c_char_p, _p, _i, _i]),
    "csx_chain_set_precision": (_i, [_p, _i]),
    "csx_chain_get_info": (_i, [_p, C.POINTER(ChainInfo)]),
    "csx_chain_get_lambdas": (_i, [_p, _p, _i]),
    "csx_chain_propagate": (_i, [_p, _i]),
    "csx_chain_get_field": (_i, [_p, _i, _i, _i, _p]),
    "csx_chain_get_intensity": (_i, [_p, _i, _p]),
    "csx_chain_contrast": (_i, [_p, C.POINTER(Region), _p, C.POINTER(Report)]),
    "csx_chain_get_normalized_image": (_i, [_p, _i, _p]),
    "csx_chain_field_profiles": (_i, [_p, _p, _i, _p]),
    "csx_chain_contrast_curve": (_i, [_p, _i, _p, _p, _i]),
    "csx_chain_region_fields": (_i, [_p, _p, _p, _p, _p, _p]),
    "csx_fits_read": (_i, [C.c_char_p, _p, _sz, _p, _p, _p]),
    "csx_fits_write": (_i, [C.c_char_p, _p, _i, _i]),
    "csx_chain_set_profiling": (_i, [_p, _i]),
    "csx_chain_set_graph": (_i, [_p, _i]),
    "csx_chain_hex_response": (_i, [_p, C.c_char_p, C.POINTER(Region), _p, C.POINTER(_i), C.POINTER(_i),
                                    C.POINTER(_i), C.POINTER(_d)]),
    "csx_group_create": (_i, [_p, _i, C.POINTER(_p)]),
    "csx_group_destroy": (None, [_p]),
    "csx_group_candidates": (_i, [_p]),
    "csx_group_contrast": (_i, [_p, C.POINTER(Region), _p]),
    "csx_group_merged_launches": (_i, [_p]),
    "csx_chain_profile_json": (_i, [_p, C.c_char_p, _sz]),
    "csx_mft_workspace_size": (_sz, [Grid, Grid, Grid, Grid, _i, _i]),
    "csx_mft": (_i, [_p, Grid, Grid, Grid, Grid, _p, _i, _p, _p, _i, _p]),
    "csx_mft_noalloc": (_i, [_p, Grid, Grid, Grid, Grid, _p, _i, _p, _p, _i]),
    "csx_mft_noalloc_capacity": (_sz, []),
    "csx_fresnel_as": (_i, [_p, _p, _i, _i, _p, _d, _d]),
    "csx_fft2": (_i, [_p, _p, _i, _i, _i]),
    "csx_downsample": (_i, [_p, _p, _i, _i, _i, _i, _d, _p]),
    "csx_apply_mask": (_i, [_p, _p, _i, _i, _p]),
    "csx_apply_mirror": (_i, [_p, _p, _i, _i, _p, _d, _p]),
    "csx_hex_mask": (_i, [_p, _p, _p, _i, _i, _p, _i, _i, _p, _i, _d, _p]),
    "csx_hex_layout": (_i, [_i, _i, _d, _p, _p, _i]),
    "csx_hex_rasterize": (_i, [_i, _i, _d, _i, _d, _d, _i, _i, _d, _d, _p, _p]),
    "csx_intensity": (_i, [_p, _p, _i, _i, _i, _p]),
    "csx_region_mean": (_i, [_p, _p, _i, _i, _p, _i, _d, _p]),
    "csx_hex_resp_objective": (_i, [_p, _p, _i, _i, _i, _p, _d, _d, _p, _i, _i, _i, _p, _i]),
    "csx_script_parse": (_i, [C.c_char_p, C.c_char_p, _sz]),
    "csx_run_script": (_i, [_p, C.c_char_p, C.c_char_p, C.c_char_p, _sz]),
    "csx_response_save": (_i, [C.c_char_p, _p, _i, _i, _i, _i, _i, _d, _p]),
    "csx_response_load": (_i, [C.c_char_p, _p, _sz, _p, C.POINTER(_d), _p, _i]),
    "csx_de_optimize": (_i, [C.c_char_p, _i, _p, _p, _p, C.POINTER(_d), C.POINTER(_i)]),
    "csx_philox_fill": (_i, [_u64, _u64, _p, _sz, _i]),
    "csx_philox4x32_10": (None, [_p, _p, _p]),
    "csx_hex_phasor_features": (_i, [_p, _p, _i, _i, _p, _i, _d, _d, _d, _p]),
    "csx_mlp_create": (_i, [_p, _i, _i, _i, _i, _u64, C.POINTER(_p)]),
    "csx_mlp_destroy": (None, [_p]),
    "csx_mlp_get_params": (_i, [_p, _p]),
    "csx_mlp_set_params": (_i, [_p, _p]),
    "csx_mlp_param_count": (_sz, [_p]),
    "csx_mlp_forward": (_i, [_p, _p, _i, _p]),
    "csx_mlp_loss_grad": (_i, [_p, _p, _p, _i, C.POINTER(_d), _p]),
    "csx_mlp_train_step": (_i, [_p, _p, _p, _i, _d, C.POINTER(_d)]),
}

_lib = None


class CsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"csx error {code}: {msg}")
        self.code = code


def lib():
    """Load libcsx.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libcsx.so not built ({LIB_PATH}); run `python -m csim_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def missing_symbols():
    """Declared entry points that the built library does not export."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is None]


def check(rc):
    if isinstance(rc, int) and rc < 0:
        msg = lib().csx_last_error()
        raise CsxError(rc, msg.decode() if msg else "")
    return rc


def tune_report() -> dict:
    """The contraction tuner's choices in this process (csx_tune_report)."""
    import json
    n = check(lib().csx_tune_report(None, 0))
    buf = C.create_string_buffer(n + 1)
    check(lib().csx_tune_report(buf, n + 1))
    return json.loads(buf.value.decode())


def last_error() -> str:
    m = lib().csx_last_error()
    return m.decode() if m else ""
