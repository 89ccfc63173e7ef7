"""TEST INFRASTRUCTURE — ctypes bindings of the BM3D oracle (oracle/bm3d_oracle.cpp).

Only tests/, the smoke check and the CPU-baseline leg of bench.py use this module, and only
as the checker. It builds the oracle with `make -C oracle` when the shared library is absent.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_bm3d.so")

METRICS = {"ssd": 0, "sad": 1, "zssd": 2, "zsad": 3, "ssd/norm": 4}
TRANSFORMS = {"dct": 0, "haar": 1, "wht": 2, "bior1.5": 3}


class OracleParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int),
        ("block_step", C.c_int), ("bm_range", C.c_int), ("radius", C.c_int),
        ("ps_num", C.c_int), ("ps_range", C.c_int),
        ("chroma", C.c_int), ("final_pass", C.c_int),
        ("sigma", C.c_float * 3),
        ("extractor", C.c_float), ("extractor_exp", C.c_int),
        ("metric", C.c_int), ("t2d", C.c_int), ("t1d", C.c_int),
        ("threads", C.c_int),
        ("by_begin", C.c_int), ("by_end", C.c_int),
        ("float_sums", C.c_int),
    ]


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oracle_transform_f64.argtypes = [C.c_int, C.c_int, dp]
        L.oracle_transform_f32.argtypes = [C.c_int, C.c_int, fp]
        L.oracle_block_distance_f32.argtypes = [C.c_int, fp, C.c_long, fp, C.c_long]
        L.oracle_block_distance_f32.restype = C.c_float
        L.oracle_block_distance_f64.argtypes = [C.c_int, dp, C.c_long, dp, C.c_long]
        L.oracle_block_distance_f64.restype = C.c_double
        L.oracle_filter_group_f32.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, fp, fp]
        L.oracle_filter_group_f32.restype = C.c_float
        L.oracle_num_blocks.argtypes = [C.POINTER(OracleParams), ip, ip]
        L.oracle_bm3d_f32.argtypes = [C.POINTER(OracleParams), fp, fp, ip, fp, ip]
        L.oracle_bm3d_f64.argtypes = [C.POINTER(OracleParams), dp, dp, ip, dp, ip]
        L.oracle_match_f32.argtypes = [C.POINTER(OracleParams), fp, ip, fp]
        L.oracle_match_blocks_f32.argtypes = [C.POINTER(OracleParams), fp, ip, C.c_int, ip, fp]
        L.oracle_aggregate_f32.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, fp, fp]
        L.oracle_aggregate_f64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, dp, dp]
        L.oracle_agg_z.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_vaggregate_f32.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(fp), fp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(width, height, stride=None, *, block_step=8, bm_range=9, radius=0, ps_num=2,
                ps_range=4, chroma=False, final=False, sigma=(0.0, 0.0, 0.0), extractor_exp=0,
                metric="ssd", t2d="dct", t1d="dct", threads=0, block_rows=None, float_sums=False) -> OracleParams:
    """Kernel-level parameters. `sigma` holds the SCALED thresholds (sigma * 0x3f021bb6 for
    the basic pass, * 0x3e40c0c1 for the final pass), as the device code receives them."""
    p = OracleParams()
    p.width, p.height, p.stride = width, height, stride or width
    p.block_step, p.bm_range, p.radius = block_step, bm_range, radius
    p.ps_num, p.ps_range = ps_num, ps_range
    p.chroma, p.final_pass = int(bool(chroma)), int(bool(final))
    s = list(sigma) if np.ndim(sigma) else [sigma] * 3
    for i in range(3):
        p.sigma[i] = float(s[i])
    p.extractor = float(np.ldexp(np.float32(1.0), extractor_exp)) if extractor_exp != 0 else 0.0
    p.extractor_exp = extractor_exp
    p.metric = METRICS[metric.lower()] if isinstance(metric, str) else int(metric)
    p.t2d = TRANSFORMS[t2d.lower()] if isinstance(t2d, str) else int(t2d)
    p.t1d = TRANSFORMS[t1d.lower()] if isinstance(t1d, str) else int(t1d)
    p.threads = threads
    if block_rows is not None:
        p.by_begin, p.by_end = block_rows
    p.float_sums = int(bool(float_sums))
    return p


def transform(kind, forward, v, dtype=np.float64):
    v = np.ascontiguousarray(v, dtype=dtype).copy()
    k = TRANSFORMS[kind] if isinstance(kind, str) else kind
    if dtype == np.float64:
        lib().oracle_transform_f64(k, int(forward), _ptr(v, C.c_double))
    else:
        lib().oracle_transform_f32(k, int(forward), _ptr(v, C.c_float))
    return v


def transform_matrix(kind, forward, dtype=np.float64):
    """Columns are the images of the unit vectors."""
    return np.stack([transform(kind, forward, np.eye(8)[i], dtype) for i in range(8)], axis=1)


def block_distance(metric, a, b, dtype=np.float32):
    a = np.ascontiguousarray(a, dtype=dtype)
    b = np.ascontiguousarray(b, dtype=dtype)
    m = METRICS[metric] if isinstance(metric, str) else metric
    if dtype == np.float32:
        return lib().oracle_block_distance_f32(m, _ptr(a, C.c_float), a.shape[1], _ptr(b, C.c_float), b.shape[1])
    return lib().oracle_block_distance_f64(m, _ptr(a, C.c_double), a.shape[1], _ptr(b, C.c_double), b.shape[1])


def filter_group(group, sigma, ref=None, t2d="dct", t1d="dct"):
    g = np.ascontiguousarray(group, dtype=np.float32).copy()
    r = None if ref is None else np.ascontiguousarray(ref, dtype=np.float32).copy()
    w = lib().oracle_filter_group_f32(TRANSFORMS[t2d], TRANSFORMS[t1d], int(ref is not None), float(sigma),
                                      _ptr(g, C.c_float), None if r is None else _ptr(r, C.c_float))
    return g, w


def num_blocks(p: OracleParams):
    nbx, nby = C.c_int(), C.c_int()
    if lib().oracle_num_blocks(C.byref(p), C.byref(nbx), C.byref(nby)):
        raise ValueError("bad oracle params")
    return nbx.value, nby.value


def bm3d(p: OracleParams, src: np.ndarray, *, dtype=np.float32, given_matches=None):
    """Run one kernel-level BM3D pass.

    src: [(final?2:1), P, TW, H, S]. Returns (res [P, TW, 2, H, S], matches [nby, nbx, 8, 3],
    match errors [nby, nbx, 8])."""
    P = 3 if p.chroma else 1
    TW = 2 * p.radius + 1
    nclips = 2 if p.final_pass else 1
    src = np.ascontiguousarray(src, dtype=dtype)
    assert src.shape == (nclips, P, TW, p.height, p.stride), src.shape
    nbx, nby = num_blocks(p)
    res = np.zeros((P, TW, 2, p.height, p.stride), dtype=dtype)
    matches = np.zeros((nby, nbx, 8, 3), dtype=np.int32)
    errs = np.zeros((nby, nbx, 8), dtype=dtype)
    gm = None
    if given_matches is not None:
        gm = np.ascontiguousarray(given_matches, dtype=np.int32)
        assert gm.shape == matches.shape
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_bm3d_f32 if dtype == np.float32 else lib().oracle_bm3d_f64
    rc = fn(C.byref(p), _ptr(src, ct), _ptr(res, ct), _ptr(matches, C.c_int32), _ptr(errs, ct),
            None if gm is None else _ptr(gm, C.c_int32))
    if rc:
        raise ValueError("oracle rejected the parameters")
    return res, matches, errs


def match(p: OracleParams, img: np.ndarray):
    """Matching only. img: [TW, H, S] float32."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    nbx, nby = num_blocks(p)
    matches = np.zeros((nby, nbx, 8, 3), dtype=np.int32)
    errs = np.zeros((nby, nbx, 8), dtype=np.float32)
    if lib().oracle_match_f32(C.byref(p), _ptr(img, C.c_float), _ptr(matches, C.c_int32), _ptr(errs, C.c_float)):
        raise ValueError("oracle rejected the parameters")
    return matches, errs


def match_blocks(p: OracleParams, img: np.ndarray, ids):
    """Matching of the raster block indices `ids` only. Returns ([n, 8, 3], [n, 8])."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    matches = np.zeros((len(ids), 8, 3), dtype=np.int32)
    errs = np.zeros((len(ids), 8), dtype=np.float32)
    if lib().oracle_match_blocks_f32(C.byref(p), _ptr(img, C.c_float), _ptr(ids, C.c_int32), len(ids),
                                     _ptr(matches, C.c_int32), _ptr(errs, C.c_float)):
        raise ValueError("oracle rejected the parameters")
    return matches, errs


def aggregate(res: np.ndarray, width: int, proc_mask: int = 0b111, dtype=np.float32):
    """res: [P, 2, H, S] (radius 0) -> dst [P, H, S]; unprocessed planes stay zero."""
    res = np.ascontiguousarray(res, dtype=dtype)
    P, two, H, S = res.shape
    assert two == 2
    dst = np.zeros((P, H, S), dtype=dtype)
    if dtype == np.float32:
        lib().oracle_aggregate_f32(P, width, H, S, proc_mask, _ptr(res, C.c_float), _ptr(dst, C.c_float))
    else:
        lib().oracle_aggregate_f64(P, width, H, S, proc_mask, _ptr(res, C.c_double), _ptr(dst, C.c_double))
    return dst


def agg_z(i, n, nframes, radius):
    return lib().oracle_agg_z(i, n, nframes, radius)


def vaggregate(slots, width, radius, n, nframes):
    """slots: list of 2r+1 arrays [TW, 2, H, S] float32 for centres clamp(n-r+i)."""
    slots = [np.ascontiguousarray(s, dtype=np.float32) for s in slots]
    TW, two, H, S = slots[0].shape
    arr = (C.POINTER(C.c_float) * len(slots))(*[_ptr(s, C.c_float) for s in slots])
    dst = np.zeros((H, S), dtype=np.float32)
    lib().oracle_vaggregate_f32(width, H, S, radius, n, nframes, arr, _ptr(dst, C.c_float))
    return dst
