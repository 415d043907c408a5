"""ctypes binding of ``libimops_gpu.so`` (the C ABI declared in include/imops_gpu.h).

The library is loaded from the package directory only; a missing library is an
error (there is no CPU fallback for any stage).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libimops_gpu.so"

OK = 0
E_INVALID = -1
E_HIP = -2
E_STATE = -3
E_UNSUPPORTED = -4

CS_NONE = -1
CS_SRGB = 0
CS_LINEAR_SRGB = 1
CS_ACESCG = 2
CS_XYZD65 = 3

DEMOSAIC_AMAZE = 0
DEMOSAIC_MARKESTEIJN = 1
DEMOSAIC_FAST = 2
DEMOSAIC_SUPERFAST = 3
DEMOSAIC_SUPERSUPERFAST = 4
DEMOSAIC_MARKESTEIJN3 = 5
DEMOSAIC_MARKESTEIJN3_1PASS = 6
BM3D_FP32 = 0
BM3D_MIXED = 1
BM3D_MIXED_MATCH = 2
BM3D_HADAMARD = 0
BM3D_HAAR = 1

_vp = C.c_void_p
_sz = C.c_size_t
_i = C.c_int
_f = C.c_float
_fp = C.POINTER(C.c_float)

class Bm3dParams(C.Structure):
    """imops_bm3d_params (include/imops_gpu.h)."""
    _fields_ = [("sigma", C.c_float), ("lambda_", C.c_float), ("max_dist", C.c_float), ("chroma_scale", C.c_float),
                ("detail", C.c_float), ("stride", C.c_int), ("search_window", C.c_int), ("max_group", C.c_int),
                ("max_candidates", C.c_int), ("precision", C.c_int), ("group_transform", C.c_int)]


class FilmFitModel(C.Structure):
    """imops_film_fit_model (include/imops_gpu.h)."""
    _fields_ = [("stock", C.c_int), ("spectral_bins", C.c_int), ("stock_seed", C.c_uint64), ("pitch_um", C.c_float)]


class FilmFitParams(C.Structure):
    """imops_film_fit_params (include/imops_gpu.h)."""
    _fields_ = [("n_emulsions", C.c_int), ("n_bins", C.c_int), ("n_knots", C.c_int), ("sens", (C.c_float * 81) * 8),
                ("knot_frac", (C.c_float * 8) * 8), ("d_max", C.c_float * 8), ("gamma", C.c_float * 8),
                ("hal_strength", C.c_float), ("hal_radius", C.c_float)]


class FilmParams(C.Structure):
    """imops_film_params (include/imops_gpu.h)."""
    _fields_ = [("stock", C.c_int), ("film_format", C.c_int), ("output", C.c_int), ("precision", C.c_int),
                ("spectral_bins", C.c_int), ("grain", C.c_int), ("seed", C.c_uint64), ("stock_seed", C.c_uint64),
                ("grain_sigma", C.c_float), ("halation_radius", C.c_float), ("halation_strength", C.c_float)]


# name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "imops_last_error": (C.c_char_p, []),
    "imops_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "imops_ctx_destroy": (_i, [_vp]),
    "imops_ctx_synchronize": (_i, [_vp]),
    "imops_ctx_stream": (_vp, [_vp]),
    "imops_ctx_set_timing": (_i, [_vp, _i]),
    "imops_ctx_timing": (_i, [_vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(_i)]),
    "imops_image_create": (_i, [_vp, C.POINTER(_vp)]),
    "imops_image_free": (_i, [_vp]),
    "imops_image_set_raw_u16": (_i, [_vp, _vp, _vp, _i, _sz, _sz, _sz, _sz, _sz, _sz, _f, _f, C.c_char_p]),
    "imops_image_set_raw_f32": (_i, [_vp, _vp, _vp, _i, _sz, _sz, C.c_char_p, _sz, _sz]),
    "imops_image_bind_raw_u16": (_i, [_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz, _sz, _f, _f, C.c_char_p]),
    "imops_crop_bayer_center": (_i, [_vp, _vp, _sz]),
    "imops_dng_calibration_matrix": (_i, [_fp, _fp, _fp, _fp, _fp]),
    "imops_image_set_rgb_f32": (_i, [_vp, _vp, _vp, _i, _sz, _sz, _i]),
    "imops_image_get_rgb_f32": (_i, [_vp, _vp, _vp, _i, _sz]),
    "imops_image_get_info": (_i, [_vp, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_i), C.POINTER(_i)]),
    "imops_image_get_rgb_u8": (_i, [_vp, _vp, _vp, _i, _i, _sz]),
    "imops_image_get_rgb_f16": (_i, [_vp, _vp, _vp, _i, _i, _sz]),
    "imops_image_rgb_device_ptr": (_i, [_vp, C.POINTER(_vp)]),
    "imops_image_set_wb_coeffs": (_i, [_vp, _fp]),
    "imops_image_set_calibration_matrix": (_i, [_vp, _fp, _sz]),
    "imops_image_set_shutter_seconds": (_i, [_vp, _f]),
    "imops_image_set_exposure_metadata": (_i, [_vp, _f, _f, _f]),
    "imops_image_set_opcode_list3": (_i, [_vp, _vp, _sz]),
    "imops_demosaic": (_i, [_vp, _vp, _i]),
    "imops_demosaic_amaze_noalloc": (_i, [_vp, _vp, _vp, _sz, _sz, C.c_char_p, _sz, _sz]),
    "imops_cfa_coeffs": (_i, [_vp, _vp, _fp]),
    "imops_exp": (_i, [_vp, _vp, _f]),
    "imops_gamma": (_i, [_vp, _vp, _f]),
    "imops_cst": (_i, [_vp, _vp, _i]),
    "imops_camera_cst": (_i, [_vp, _vp, _i, _fp, _sz]),
    "imops_bm3d_default_params": (_i, [_f, C.POINTER(Bm3dParams)]),
    "imops_bm3d": (_i, [_vp, _vp, _f]),
    "imops_bm3d_ex": (_i, [_vp, _vp, C.POINTER(Bm3dParams)]),
    "imops_chroma_bm3d": (_i, [_vp, _vp, _f]),
    "imops_luma_chroma_denoise": (_i, [_vp, _vp, _i, _f]),
    "imops_contrast": (_i, [_vp, _vp, _f]),
    "imops_highlight_reconstruction": (_i, [_vp, _vp]),
    "imops_lch": (_i, [_vp, _vp, _f, _f, _f]),
    "imops_sigmoid_tone_map": (_i, [_vp, _vp]),
    "imops_baseline_exposure": (_i, [_vp, _vp]),
    "imops_vignette": (_i, [_vp, _vp, _f]),
    "imops_film": (_i, [_vp, _vp, _i, _i, C.c_uint64, _i]),
    "imops_film_default_params": (_i, [C.POINTER(FilmParams)]),
    "imops_film_ex": (_i, [_vp, _vp, C.POINTER(FilmParams)]),
    "imops_film_fit_nominal": (_i, [C.POINTER(FilmFitModel), C.POINTER(FilmFitParams)]),
    "imops_film_fit_knots": (_i, [C.POINTER(FilmFitModel), _i, C.POINTER(C.c_double)]),
    "imops_film_fit_state_floats": (_i, [C.POINTER(FilmFitModel), _sz, _sz, _sz, C.POINTER(_sz)]),
    "imops_film_fit_forward": (_i, [_vp, C.POINTER(FilmFitModel), C.POINTER(FilmFitParams), _vp, _vp, _sz, _sz, _sz,
                                    _vp]),
    "imops_film_fit_backward": (_i, [_vp, C.POINTER(FilmFitModel), C.POINTER(FilmFitParams), _vp, _vp, _sz, _sz,
                                     _sz, _vp, C.POINTER(FilmFitParams)]),
    "imops_film_fit_l2_step": (_i, [_vp, C.POINTER(FilmFitModel), C.POINTER(FilmFitParams), _vp, _vp, _sz, _sz,
                                    _sz, C.POINTER(FilmFitParams), C.POINTER(C.c_double)]),
    "imops_pipeline_parse": (_i, [C.c_char_p, _sz, C.POINTER(_vp)]),
    "imops_pipeline_free": (_i, [_vp]),
    "imops_pipeline_num_stages": (_sz, [_vp]),
    "imops_pipeline_stage_name": (C.c_char_p, [_vp, _sz]),
    "imops_pipeline_run": (_i, [_vp, _vp, _vp]),
    "imops_pipeline_run_burst": (_i, [_vp, _vp, C.POINTER(_vp), _sz, _i]),
    "imops_ctx_set_autotune": (_i, [_vp, _i]),
    "imops_ctx_tuning": (_i, [_vp, C.c_char_p, _sz, C.POINTER(_sz)]),
    "imops_measure_peaks": (_i, [_vp, C.POINTER(C.c_double), _i]),
    "imops_debug_fail_allocations": (_i, [_i]),
    "imops_synth_scene_rgb": (_i, [_fp, _sz, _sz, C.c_uint64]),
    "imops_synth_raw_u16": (_i, [_vp, _sz, _sz, C.c_char_p, C.c_uint64, _f, _f, _f, _f]),
}


class ImopsError(RuntimeError):
    """A failed library call: status code + the library's last error message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m imops_amd.build` (no CPU fallback exists)"
            )
        # PyTorch-ROCm wheels carry their own HIP runtime under the same soname as the
        # system one this library links.  Whichever loads first serves the process; if it
        # were the system one, torch's later GPU initialisation finds no devices.  Loading
        # torch first keeps one runtime, torch's, whatever order the caller imports in.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().imops_last_error()
        raise ImopsError(status, msg.decode() if msg else "")


def fptr(arr) -> C.POINTER(C.c_float):
    return arr.ctypes.data_as(_fp)
