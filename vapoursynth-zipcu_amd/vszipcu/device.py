"""Kernel-layer calls on device-resident tensors.

PyTorch-ROCm only provides device memory and the stream here; every computation runs in the
native library's HIP kernels. Tensors must be contiguous float32 (int32 for match lists) on a
HIP device; shapes are checked on the host before any launch.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

from . import _lib

METRICS = {"ssd": 0, "sad": 1, "zssd": 2, "zsad": 3, "ssd/norm": 4}
TRANSFORMS = {"dct": 0, "haar": 1, "wht": 2, "bior1.5": 3}

# f32 bit patterns of the stepwise products (3/4)/255*64*2.7 and (3/4)/255*64*1.0
# (src/bm3d.zig:1019-1024)
SIGMA_FACTOR_BASIC_BITS = 0x3F021BB6
SIGMA_FACTOR_FINAL_BITS = 0x3E40C0C1


def sigma_factor(final: bool) -> float:
    import struct
    return struct.unpack("<f", struct.pack("<I", SIGMA_FACTOR_FINAL_BITS if final else SIGMA_FACTOR_BASIC_BITS))[0]


def scaled_sigma(sigma255, final: bool):
    """FP32 product sigma * factor, as the filter host computes it."""
    import numpy as np
    s = np.asarray(sigma255, dtype=np.float32) * np.float32(sigma_factor(final))
    return [float(v) for v in np.atleast_1d(s)]


def kparams(width, height, stride=None, *, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4,
            chroma=False, final=False, sigma=(0.0, 0.0, 0.0), extractor_exp=0, metric="ssd", t2d="dct",
            t1d="dct", proc_mask=0b111, precision="fp32") -> _lib.KParams:
    """Kernel-level parameters; `sigma` are the SCALED thresholds. precision "fp32" keeps the
    reference arithmetic; "split-bf16" runs the DCT stages as compensated BF16 MFMA products."""
    p = _lib.KParams()
    p.width, p.height, p.stride = width, height, stride or width
    p.block_step, p.bm_range = block_step, bm_range
    p.radius, p.ps_num, p.ps_range = radius, ps_num, ps_range
    p.chroma, p.final_pass = int(bool(chroma)), int(bool(final))
    s = list(sigma) if isinstance(sigma, (list, tuple)) else [sigma] * 3
    for i in range(3):
        p.sigma[i] = float(s[i])
    p.extractor_exp = extractor_exp
    p.metric = METRICS[metric.lower()] if isinstance(metric, str) else int(metric)
    p.t2d = TRANSFORMS[t2d.lower()] if isinstance(t2d, str) else int(t2d)
    p.t1d = TRANSFORMS[t1d.lower()] if isinstance(t1d, str) else int(t1d)
    p.proc_mask = proc_mask
    p.transform_precision = {"fp32": 0, "split-bf16": 1}[precision]
    return p


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"vszipcu: {_lib.last_error()} (code {rc})")


def num_blocks(p: _lib.KParams):
    nbx, nby = C.c_int(), C.c_int()
    _check(_lib.load().vsz_bm3d_num_blocks(C.byref(p), C.byref(nbx), C.byref(nby)))
    return nbx.value, nby.value


def match_plan(p: _lib.KParams) -> str:
    """Name of the matcher vsz_bm3d_match launches for these parameters (no GPU needed)."""
    name = _lib.load().vsz_bm3d_match_plan(C.byref(p)).decode()
    if not name:
        raise RuntimeError(f"vszipcu: {_lib.last_error()}")
    return name


def filter_plan(p: _lib.KParams):
    """(tbx, tby, workgroups per CU) of the LDS-tiled filter vsz_bm3d_run launches for these
    parameters, or (0, 0, 0) for the direct-atomic filter (no GPU needed)."""
    tx, ty, n = C.c_int(), C.c_int(), C.c_int()
    _check(_lib.load().vsz_bm3d_filter_plan(C.byref(p), C.byref(tx), C.byref(ty), C.byref(n)))
    return tx.value, ty.value, n.value


def workspace_size(p: _lib.KParams) -> int:
    n = _lib.load().vsz_bm3d_workspace_size(C.byref(p))
    if n == 0:
        raise RuntimeError("vszipcu: invalid kernel parameters")
    return n


def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _need(t, shape, dtype_name):
    import torch
    dt = {"f32": torch.float32, "i32": torch.int32, "u8": torch.uint8}[dtype_name]
    if t.device.type != "cuda" or t.dtype != dt or not t.is_contiguous():
        raise ValueError(f"expected a contiguous {dtype_name} device tensor, got {t.dtype} on {t.device}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"shape {tuple(t.shape)} != expected {tuple(shape)}")


def match(p: _lib.KParams, img, matches, stream=None):
    """img: [TW, H, S] f32; matches: [nby, nbx, 8, 4] i32."""
    nbx, nby = num_blocks(p)
    _need(img, (2 * p.radius + 1, p.height, p.stride), "f32")
    _need(matches, (nby, nbx, 8, 4), "i32")
    _check(_lib.load().vsz_bm3d_match(C.byref(p), C.c_void_p(img.data_ptr()), C.c_void_p(matches.data_ptr()),
                                      _stream_handle(stream)))


def run(p: _lib.KParams, clip0, clip1, work, out, out_kind: int = 0, stream=None):
    """One pass. clip0/clip1: [P, TW, H, S] f32; work: workspace_size bytes (u8);
    out: [P, H, S] (out_kind 0) or [P, TW, 2, H, S] (out_kind 1)."""
    P = 3 if p.chroma else 1
    TW = 2 * p.radius + 1
    _need(clip0, (P, TW, p.height, p.stride), "f32")
    if p.final_pass:
        _need(clip1, (P, TW, p.height, p.stride), "f32")
    _need(work, None, "u8")
    if work.numel() < workspace_size(p):
        raise ValueError("workspace too small")
    _need(out, (P, p.height, p.stride) if out_kind == 0 else (P, TW, 2, p.height, p.stride), "f32")
    _check(_lib.load().vsz_bm3d_run(C.byref(p), C.c_void_p(clip0.data_ptr()),
                                    C.c_void_p(clip1.data_ptr() if p.final_pass else 0),
                                    C.c_void_p(work.data_ptr()), C.c_void_p(out.data_ptr()), out_kind,
                                    _stream_handle(stream)))


def vaggregate(width: int, radius: int, n: int, num_frames: int, slots: Sequence, dst, stream=None):
    """slots: 2r+1 tensors [TW, 2, H, S] f32 (centres clamp(n-r+i)); dst: [H, S] f32."""
    TW = 2 * radius + 1
    if len(slots) != TW:
        raise ValueError("need 2r+1 slots")
    H, S = dst.shape
    for s in slots:
        _need(s, (TW, 2, H, S), "f32")
    _need(dst, (H, S), "f32")
    arr = (C.c_void_p * TW)(*[s.data_ptr() for s in slots])
    _check(_lib.load().vsz_vaggregate_run(width, H, S, radius, n, num_frames, arr, C.c_void_p(dst.data_ptr()),
                                          _stream_handle(stream)))


class Batch:
    """Many independent Gray planes, two-pass BM3D, one launch per stage, replayed as one hipGraph.

    planes: [(width, height, stride, sigma)] with the filter's raw sigma per plane."""

    SCHEDULES = {"staged": 0, "persistent": 1, "auto": 2}

    def __init__(self, planes, *, block_step=4, bm_range=12, extractor_exp=3, metric="ssd", precision="fp32",
                 two_pass=True, autotune=True, schedule="auto", transform="dct"):
        bp = _lib.BatchParams()
        bp.block_step, bp.bm_range, bp.extractor_exp = block_step, bm_range, extractor_exp
        bp.metric = METRICS[metric.lower()] if isinstance(metric, str) else int(metric)
        bp.t2d = bp.t1d = TRANSFORMS[transform.lower()] if isinstance(transform, str) else int(transform)
        bp.transform_precision = {"fp32": 0, "split-bf16": 1}[precision]
        bp.two_pass, bp.autotune = int(bool(two_pass)), int(bool(autotune))
        bp.schedule = self.SCHEDULES[schedule]
        arr = (_lib.Plane * len(planes))()
        for i, (w, h, st, sg) in enumerate(planes):
            arr[i].width, arr[i].height, arr[i].stride, arr[i].sigma = w, h, st, sg
        self.planes = [tuple(p) for p in planes]
        h = C.c_void_p()
        _check(_lib.load().vsz_bm3d_batch_create(C.byref(bp), arr, len(planes), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().vsz_bm3d_batch_free(h)
            self._h = C.c_void_p(0)

    @property
    def workspace_size(self) -> int:
        return _lib.load().vsz_bm3d_batch_workspace_size(self._h)

    def run(self, noisy: Sequence, out: Sequence, work, stream=None):
        """noisy[i], out[i]: [height, stride] f32 device tensors; work: workspace_size bytes (u8)."""
        n = len(self.planes)
        if len(noisy) != n or len(out) != n:
            raise ValueError(f"need {n} input and output planes")
        for (w, h, st, _), a, b in zip(self.planes, noisy, out):
            _need(a, (h, st), "f32")
            _need(b, (h, st), "f32")
        _need(work, None, "u8")
        if work.numel() < self.workspace_size:
            raise ValueError("workspace too small")
        ins = (C.c_void_p * n)(*[t.data_ptr() for t in noisy])
        outs = (C.c_void_p * n)(*[t.data_ptr() for t in out])
        _check(_lib.load().vsz_bm3d_batch_run(self._h, ins, outs, C.c_void_p(work.data_ptr()), _stream_handle(stream)))

    def check(self, stream=None):
        """Raises if a dependency wait of a persistent launch reached its time bound."""
        _check(_lib.load().vsz_bm3d_batch_check(self._h, _stream_handle(stream)))

    def schedule(self):
        """(schedule, matcher) the batch runs with, e.g. ("staged", "shared-column"); final after its
        first run. Matchers: "rows", "shared-column", "global" (windows beyond the LDS budget)."""
        sch, m = C.c_int(), C.c_int()
        _check(_lib.load().vsz_bm3d_batch_schedule(self._h, C.byref(sch), C.byref(m)))
        return {0: "staged", 1: "persistent"}[sch.value], {0: "rows", 1: "shared-column", 2: "global"}[m.value]

    def tile(self, plane: int, pass_: int):
        tx, ty = C.c_int(), C.c_int()
        _check(_lib.load().vsz_bm3d_batch_tile(self._h, plane, pass_, C.byref(tx), C.byref(ty)))
        return tx.value, ty.value


def device_count() -> int:
    return _lib.load().vsz_device_count()
