"""ctypes bindings of libvszipcu_amd.so (include/vszipcu_bm3d.h).

The shared library is built in-tree (``make -C vapoursynth-zipcu_amd``) into
``vapoursynth-zipcu_amd/lib/``. There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libvszipcu_amd.so")

VSZ_OK, VSZ_EINVAL, VSZ_EDEVICE, VSZ_ENOMEM, VSZ_EFRAME = 0, -1, -2, -3, -4


class KParams(C.Structure):
    """vsz_bm3d_kparams"""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int),
        ("block_step", C.c_int), ("bm_range", C.c_int),
        ("radius", C.c_int), ("ps_num", C.c_int), ("ps_range", C.c_int),
        ("chroma", C.c_int), ("final_pass", C.c_int),
        ("sigma", C.c_float * 3),
        ("extractor_exp", C.c_int),
        ("metric", C.c_int), ("t2d", C.c_int), ("t1d", C.c_int),
        ("proc_mask", C.c_uint),
        ("transform_precision", C.c_int),
    ]


class ClipInfo(C.Structure):
    """vsz_clip_info"""
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("num_frames", C.c_int),
        ("color_family", C.c_int), ("sample_type", C.c_int), ("bits_per_sample", C.c_int),
        ("sub_sampling_w", C.c_int), ("sub_sampling_h", C.c_int), ("num_planes", C.c_int),
    ]


class Bm3dArgs(C.Structure):
    """vsz_bm3d_args"""
    _fields_ = [
        ("clip", ClipInfo),
        ("has_ref", C.c_int), ("ref", ClipInfo),
        ("n_sigma", C.c_int), ("sigma", C.c_float * 3),
        ("n_block_step", C.c_int), ("block_step", C.c_int * 3),
        ("n_bm_range", C.c_int), ("bm_range", C.c_int * 3),
        ("has_radius", C.c_int), ("radius", C.c_int),
        ("n_ps_num", C.c_int), ("ps_num", C.c_int * 3),
        ("n_ps_range", C.c_int), ("ps_range", C.c_int * 3),
        ("has_chroma", C.c_int), ("chroma", C.c_int),
        ("has_device_id", C.c_int), ("device_id", C.c_int),
        ("has_num_streams", C.c_int), ("num_streams", C.c_int),
        ("has_extractor_exp", C.c_int), ("extractor_exp", C.c_int),
        ("n_bm_error_s", C.c_int), ("bm_error_s", C.c_char_p * 3),
        ("n_transform_2d_s", C.c_int), ("transform_2d_s", C.c_char_p * 3),
        ("n_transform_1d_s", C.c_int), ("transform_1d_s", C.c_char_p * 3),
        ("has_zero_init", C.c_int), ("zero_init", C.c_int),
        ("has_fast_fused", C.c_int), ("fast_fused", C.c_int),
    ]


class OutputInfo(C.Structure):
    """vsz_output_info"""
    _fields_ = [
        ("vi", ClipInfo), ("passthrough", C.c_int), ("stacked", C.c_int),
        ("radius", C.c_int), ("process", C.c_int * 3),
    ]


class Plane(C.Structure):
    """vsz_bm3d_plane"""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("stride", C.c_int), ("sigma", C.c_float)]


class BatchParams(C.Structure):
    """vsz_bm3d_batch_params"""
    _fields_ = [
        ("block_step", C.c_int), ("bm_range", C.c_int), ("extractor_exp", C.c_int),
        ("metric", C.c_int), ("t2d", C.c_int), ("t1d", C.c_int), ("transform_precision", C.c_int),
        ("two_pass", C.c_int), ("autotune", C.c_int), ("schedule", C.c_int),
    ]


class CnnModel(C.Structure):
    """vsz_cnn_model"""
    _fields_ = [("w", C.c_void_p * 6), ("b", C.c_void_p * 6)]


GET_PLANE_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_ssize_t))

# Every function include/vszipcu_bm3d.h declares: name -> (restype, argtypes)
PROTOTYPES = {
    "vsz_last_error": (C.c_char_p, []),
    "vsz_version": (C.c_int, []),
    "vsz_device_count": (C.c_int, []),
    "vsz_bm3d_num_blocks": (C.c_int, [C.POINTER(KParams), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vsz_bm3d_workspace_size": (C.c_size_t, [C.POINTER(KParams)]),
    "vsz_bm3d_match": (C.c_int, [C.POINTER(KParams), C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsz_bm3d_match_plan": (C.c_char_p, [C.POINTER(KParams)]),
    "vsz_bm3d_filter_plan": (C.c_int, [C.POINTER(KParams), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vsz_bm3d_run": (C.c_int, [C.POINTER(KParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                               C.c_void_p]),
    "vsz_bm3d_run_noalloc": (C.c_int, [C.POINTER(KParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_void_p]),
    "vsz_vaggregate_run": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p]),
    "vsz_bm3d_create": (C.c_int, [C.POINTER(Bm3dArgs), C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "vsz_bm3dv2_create": (C.c_int, [C.POINTER(Bm3dArgs), C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "vsz_vaggregate_create": (C.c_int, [C.POINTER(ClipInfo), C.POINTER(ClipInfo), C.POINTER(C.c_int), C.c_int,
                                        C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "vsz_filter_output_info": (C.c_int, [C.c_void_p, C.POINTER(OutputInfo)]),
    "vsz_filter_get_frame": (C.c_int, [C.c_void_p, C.c_int, GET_PLANE_FN, C.c_void_p, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_ssize_t), C.c_char_p, C.c_size_t]),
    "vsz_bm3d_batch_create": (C.c_int, [C.POINTER(BatchParams), C.POINTER(Plane), C.c_int, C.POINTER(C.c_void_p)]),
    "vsz_bm3d_batch_workspace_size": (C.c_size_t, [C.c_void_p]),
    "vsz_bm3d_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p,
                                     C.c_void_p]),
    "vsz_bm3d_batch_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vsz_bm3d_batch_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vsz_bm3d_batch_check": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vsz_bm3d_batch_free": (None, [C.c_void_p]),
    "vsz_cnn_conv3x3": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "vsz_cnn_conv3x3_dw": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "vsz_cnn_bias_grad": (C.c_int, [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p]),
    "vsz_cnn_pack_input": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "vsz_cnn_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsz_cnn_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float,
                               C.c_float, C.c_float, C.c_int, C.c_void_p]),
    "vsz_cnn_pack_weights": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsz_cnn_workspace_size": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "vsz_cnn_run": (C.c_int, [C.POINTER(CnnModel), C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p]),
    "vsz_filter_frame_range": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vsz_filter_free": (None, [C.c_void_p]),
}

_lib = None


def load():
    """Loads the native library (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libvszipcu_amd.so is not built ({LIB_PATH}); run `make -C vapoursynth-zipcu_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return load().vsz_last_error().decode()
