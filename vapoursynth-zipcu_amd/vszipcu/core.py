"""Host-side mirror of the plugin namespace ``core.vszipcu`` for the BM3D path.

``core.vszipcu.BM3D / BM3Dv2 / VAggregate`` take the reference's arguments (src/plugin.zig:84-98)
and return lazily evaluated clips. Argument checking, per-plane fallback, error strings and
frame semantics live in the native library (include/vszipcu_bm3d.h); this module only
marshals numpy frames through its frame callback. Errors raise :class:`Error` carrying the
library's message, the way VapourSynth raises ``vs.Error`` from ``mapSetError``.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib

GRAY, RGB, YUV = 1, 2, 3
INTEGER, FLOAT = 0, 1


class Error(Exception):
    """Raised with the filter's error message (the counterpart of vs.Error)."""


@dataclass(frozen=True)
class VideoFormat:
    name: str
    color_family: int
    sample_type: int
    bits_per_sample: int
    sub_sampling_w: int = 0
    sub_sampling_h: int = 0

    @property
    def num_planes(self) -> int:
        return 1 if self.color_family == GRAY else 3


GRAYS = VideoFormat("GRAYS", GRAY, FLOAT, 32)
YUV444PS = VideoFormat("YUV444PS", YUV, FLOAT, 32)
YUV420PS = VideoFormat("YUV420PS", YUV, FLOAT, 32, 1, 1)
RGBS = VideoFormat("RGBS", RGB, FLOAT, 32)
GRAY16 = VideoFormat("GRAY16", GRAY, INTEGER, 16)
GRAYH = VideoFormat("GRAYH", GRAY, FLOAT, 16)


@dataclass
class Frame:
    planes: list
    props: dict = field(default_factory=dict)


class Clip:
    """A clip of numpy frames. `source(n)` returns the frame's planes (float32 2-D arrays)."""

    def __init__(self, fmt: VideoFormat, width: int, height: int, num_frames: int,
                 source: Callable[[int], Frame]):
        self.format = fmt
        self.width = width
        self.height = height
        self.num_frames = num_frames
        self._source = source
        self._cache: dict = {}
        self._lock = threading.Lock()

    @classmethod
    def from_frames(cls, frames: Sequence, fmt: VideoFormat = GRAYS) -> "Clip":
        """frames: list of frames; each frame is a 2-D array (Gray) or a list of planes."""
        fr = []
        for f in frames:
            planes = [f] if isinstance(f, np.ndarray) and f.ndim == 2 else list(f)
            fr.append(Frame([np.ascontiguousarray(p) for p in planes]))
        h, w = fr[0].planes[0].shape
        return cls(fmt, w, h, len(fr), lambda n: fr[n])

    def plane_shape(self, p: int, height: Optional[int] = None):
        h = self.height if height is None else height
        if p == 0:
            return h, self.width
        return h >> self.format.sub_sampling_h, self.width >> self.format.sub_sampling_w

    def get_frame(self, n: int) -> Frame:
        if not 0 <= n < self.num_frames:
            raise Error(f"frame {n} out of range")
        with self._lock:
            f = self._cache.get(n)
        if f is None:
            f = self._source(n)
            with self._lock:
                self._cache[n] = f
                while len(self._cache) > 64:
                    self._cache.pop(next(iter(self._cache)))
        return f

    def info(self) -> _lib.ClipInfo:
        ci = _lib.ClipInfo()
        ci.width, ci.height, ci.num_frames = self.width, self.height, self.num_frames
        ci.color_family = self.format.color_family
        ci.sample_type = self.format.sample_type
        ci.bits_per_sample = self.format.bits_per_sample
        ci.sub_sampling_w = self.format.sub_sampling_w
        ci.sub_sampling_h = self.format.sub_sampling_h
        ci.num_planes = self.format.num_planes
        return ci


class _FilterClip(Clip):
    """Output of a native filter instance; frames are computed on demand."""

    def __init__(self, handle: int, inputs: Sequence[Optional[Clip]], out: _lib.OutputInfo, fmt: VideoFormat):
        self._handle = C.c_void_p(handle)
        self._inputs = list(inputs)
        self.out_info = out
        super().__init__(fmt, out.vi.width, out.vi.height, out.vi.num_frames, self._render)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _lib.load().vsz_filter_free(h)
            self._handle = C.c_void_p(0)

    def _render(self, n: int) -> Frame:
        L = _lib.load()
        keep = []

        def get_plane(_user, which, fn_n, plane, stride_p):
            try:
                clip = self._inputs[which]
                if clip is None:
                    return None
                arr = np.ascontiguousarray(clip.get_frame(fn_n).planes[plane], dtype=np.float32)
                keep.append(arr)
                stride_p[0] = arr.strides[0]
                return arr.ctypes.data
            except Exception:  # noqa: BLE001 - surfaces as a frame error in the library
                return None

        cb = _lib.GET_PLANE_FN(get_plane)
        planes = [np.zeros(self.plane_shape(p), np.float32) for p in range(self.format.num_planes)]
        ptrs = (C.c_void_p * 3)(*([p.ctypes.data for p in planes] + [None] * (3 - len(planes))))
        strides = (C.c_ssize_t * 3)(*([p.strides[0] for p in planes] + [0] * (3 - len(planes))))
        err = C.create_string_buffer(512)
        rc = L.vsz_filter_get_frame(self._handle, n, cb, None, ptrs, strides, err, 512)
        if rc != 0:
            raise Error(err.value.decode())
        props = {}
        if self.out_info.stacked:
            props["BM3D_V_radius"] = self.out_info.radius
            props["BM3D_V_process"] = [int(v) for v in self.out_info.process]
        return Frame(planes, props)


def _opt(kw, name):
    return kw.get(name) is not None


def _seq(v):
    if v is None:
        return []
    if isinstance(v, (list, tuple)):
        return list(v)
    return [v]


def _bm3d_args(clip: Clip, ref: Optional[Clip], kw: dict) -> _lib.Bm3dArgs:
    a = _lib.Bm3dArgs()
    a.clip = clip.info()
    if ref is not None:
        a.has_ref = 1
        a.ref = ref.info()

    def per_plane(name, ctype_arr, conv):
        vals = _seq(kw.get(name))[:3]
        for i, v in enumerate(vals):
            ctype_arr[i] = conv(v)
        return len(vals)

    a.n_sigma = per_plane("sigma", a.sigma, float)
    a.n_block_step = per_plane("block_step", a.block_step, int)
    a.n_bm_range = per_plane("bm_range", a.bm_range, int)
    a.n_ps_num = per_plane("ps_num", a.ps_num, int)
    a.n_ps_range = per_plane("ps_range", a.ps_range, int)
    for name in ("radius", "chroma", "device_id", "num_streams", "extractor_exp", "zero_init", "fast_fused"):
        if _opt(kw, name):
            setattr(a, "has_" + name, 1)
            setattr(a, name, int(kw[name]))
    keep = []
    for name in ("bm_error_s", "transform_2d_s", "transform_1d_s"):
        vals = _seq(kw.get(name))[:3]
        setattr(a, "n_" + name, len(vals))
        arr = getattr(a, name)
        for i, v in enumerate(vals):
            b = v.encode() if isinstance(v, str) else bytes(v)
            keep.append(b)
            arr[i] = b
    a._keep = keep  # keep the string buffers alive
    return a


_BM3D_KEYS = {"sigma", "block_step", "bm_range", "radius", "ps_num", "ps_range", "chroma", "device_id",
              "num_streams", "extractor_exp", "bm_error_s", "transform_2d_s", "transform_1d_s", "zero_init",
              "fast_fused"}


def _create(fn_name: str, clip: Clip, ref: Optional[Clip], kw: dict) -> Clip:
    unknown = set(kw) - _BM3D_KEYS
    if unknown:
        raise Error(f"{fn_name}: unknown argument(s) {sorted(unknown)}")
    L = _lib.load()
    args = _bm3d_args(clip, ref, kw)
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = getattr(L, "vsz_bm3d_create" if fn_name == "BM3D" else "vsz_bm3dv2_create")(C.byref(args), C.byref(h), err, 512)
    if rc != 0:
        raise Error(err.value.decode())
    out = _lib.OutputInfo()
    L.vsz_filter_output_info(h, C.byref(out))
    if out.passthrough:
        L.vsz_filter_free(h)
        return clip
    # getter indices: 0 = clip, 1 = ref
    return _FilterClip(h.value, [clip, ref], out, clip.format)


class _Namespace:
    """core.vszipcu"""

    @staticmethod
    def BM3D(clip: Clip, ref: Optional[Clip] = None, **kw) -> Clip:
        return _create("BM3D", clip, ref, kw)

    @staticmethod
    def BM3Dv2(clip: Clip, ref: Optional[Clip] = None, **kw) -> Clip:
        return _create("BM3Dv2", clip, ref, kw)

    @staticmethod
    def VAggregate(clip: Clip, src: Clip, planes) -> Clip:
        L = _lib.load()
        pl = _seq(planes)
        arr = (C.c_int * max(len(pl), 1))(*pl)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        ci, si = clip.info(), src.info()
        rc = L.vsz_vaggregate_create(C.byref(ci), C.byref(si), arr, len(pl), C.byref(h), err, 512)
        if rc != 0:
            raise Error(err.value.decode())
        out = _lib.OutputInfo()
        L.vsz_filter_output_info(h, C.byref(out))
        # getter indices: 0 = stacked clip, 1 = src
        return _FilterClip(h.value, [clip, src], out, src.format)


class _Core:
    vszipcu = _Namespace()


core = _Core()
