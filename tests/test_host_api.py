"""CPU tier for the native library (no kernel launches).

* the library loads and exports every function include/vszipcu_bm3d.h declares;
* without a GPU the library reports zero devices and filter creation fails with a clean
  error instead of crashing (the reference CI's no-GPU smoke semantics, .github/smoke.py:50-59);
* argument validation reproduces the reference's rules and verbatim messages
  (src/bm3d.zig:991-1201, 1372-1417), checked before any device work;
* BM3Dv2 with every sigma below FLT_EPSILON returns its input (src/bm3d.zig:1437-1456);
* VAggregate runs on the CPU (src/bm3d.zig:1284-1363) and matches the oracle bit for bit.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vszipcu
from vszipcu import GRAYS, RGBS, YUV420PS, YUV444PS, Clip, Error, core
from vszipcu import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = vszipcu.load()
    names = set()
    for h in ("vszipcu_bm3d.h", "vszipcu_cnn.h"):
        header = open(os.path.join(REPO, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vsz_\w+)\s*\(", header, re.M))
    assert len(names) == 34
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.PROTOTYPES) == names
    assert "void VapourSynthPluginInit2(" in open(os.path.join(REPO, "include", "vszipcu_vs4.h")).read()
    assert hasattr(L, "VapourSynthPluginInit2")


def _gray(w=32, h=24, n=3, fmt=GRAYS):
    rng = np.random.default_rng(0)
    frames = []
    for _ in range(n):
        planes = [rng.random((h, w), dtype=np.float32)]
        if fmt.num_planes == 3:
            cw, ch = w >> fmt.sub_sampling_w, h >> fmt.sub_sampling_h
            planes += [rng.random((ch, cw), dtype=np.float32) for _ in range(2)]
        frames.append(planes)
    return Clip.from_frames(frames, fmt)


def _no_gpu():
    return vszipcu.load().vsz_device_count() == 0


@pytest.mark.parametrize("kwargs,msg", [
    (dict(sigma=[-1.0]), 'BM3D: "sigma" must be non-negative.'),
    (dict(block_step=0), 'BM3D: "block_step" must be in range [1, 8].'),
    (dict(block_step=[4, 9]), 'BM3D: "block_step" must be in range [1, 8].'),
    (dict(bm_range=0), 'BM3D: "bm_range" must be positive.'),
    (dict(ps_num=9), 'BM3D: "ps_num" must be in range [1, 8].'),
    (dict(ps_range=0), 'BM3D: "ps_range" must be positive.'),
    (dict(radius=-1), 'BM3D: "radius" must be non-negative.'),
    (dict(radius=17), 'BM3D: "radius" must be <= 16.'),
    (dict(chroma=1), 'BM3D: clip format must be YUV444 when "chroma" is true.'),
    (dict(bm_error_s=["ssd", "nope"]), 'BM3D: invalid "bm_error_s".'),
    (dict(transform_2d_s="dft"), 'BM3D: invalid "transform_2d_s".'),
    (dict(transform_1d_s=["DCT", "Bior1.5", "x"]), 'BM3D: invalid "transform_1d_s".'),
    (dict(device_id=-1), "BM3D: invalid device ID."),
    (dict(num_streams=0), "BM3D: num_streams must be 1..32."),
    (dict(num_streams=33), "BM3D: num_streams must be 1..32."),
    (dict(sigma=0.0), "BM3D: all planes have sigma < FLT_EPSILON (nothing to process)."),
])
def test_bm3d_validation_messages(kwargs, msg):
    with pytest.raises(Error) as ei:
        core.vszipcu.BM3D(_gray(), **kwargs)
    assert str(ei.value) == msg


def test_format_and_ref_rules():
    with pytest.raises(Error, match=re.escape("BM3D: only constant format 32 bit float input supported.")):
        core.vszipcu.BM3D(_gray(fmt=vszipcu.GRAY16))
    g = _gray()
    with pytest.raises(Error, match=re.escape('BM3D: "ref" must be of the same format as "clip".')):
        core.vszipcu.BM3D(g, ref=_gray(fmt=RGBS))
    with pytest.raises(Error, match=re.escape('BM3D: "ref" must be of the same dimensions as "clip".')):
        core.vszipcu.BM3D(g, ref=_gray(w=40))
    with pytest.raises(Error, match=re.escape('BM3D: "ref" must be of the same number of frames as "clip".')):
        core.vszipcu.BM3D(g, ref=_gray(n=5))
    with pytest.raises(Error, match=re.escape("BM3D: every processed plane must be at least 8x8.")):
        core.vszipcu.BM3D(_gray(w=7))
    # 4:2:0 chroma planes 4x6 < 8x8 are rejected only when processed
    with pytest.raises(Error, match=re.escape("BM3D: every processed plane must be at least 8x8.")):
        core.vszipcu.BM3D(_gray(w=8, h=12, fmt=YUV420PS))


def test_per_plane_fallback_skips_plane_errors():
    # sigma=[3, 0] -> planes 1 and 2 unprocessed; block_step=[4] falls back for all planes
    clip = _gray(w=8, h=12, fmt=YUV420PS)
    try:
        core.vszipcu.BM3D(clip, sigma=[3.0, 0.0], block_step=[4])
    except Error as e:
        assert "at least 8x8" not in str(e)


def test_clean_error_without_gpu():
    if not _no_gpu():
        pytest.skip("a GPU is visible")
    with pytest.raises(Error, match="HIP initialization failed"):
        core.vszipcu.BM3D(_gray())


def test_bm3dv2_passthrough_when_nothing_processed():
    g = _gray()
    assert core.vszipcu.BM3Dv2(g, sigma=[0.0]) is g


def test_vaggregate_validation():
    src = _gray(n=4)
    stacked = Clip.from_frames([np.zeros((24 * 6, 32), np.float32)] * 4, GRAYS)
    with pytest.raises(Error, match=re.escape('VAggregate: "planes" is required.')):
        core.vszipcu.VAggregate(stacked, src, [])
    with pytest.raises(Error, match=re.escape('VAggregate: "planes" must be 0..2.')):
        core.vszipcu.VAggregate(stacked, src, [3])
    bad = Clip.from_frames([np.zeros((24 * 5, 32), np.float32)] * 4, GRAYS)
    with pytest.raises(Error, match=re.escape('VAggregate: "clip" is not a BM3D stacked clip of "src" (radius >= 1).')):
        core.vszipcu.VAggregate(bad, src, [0])
    with pytest.raises(Error, match=re.escape('VAggregate: "clip" must have the same format and frame count as "src".')):
        core.vszipcu.VAggregate(Clip.from_frames([np.zeros((24 * 6, 32), np.float32)] * 3, GRAYS), src, [0])
    with pytest.raises(Error, match=re.escape('VAggregate: "src" must be 32 bit float.')):
        core.vszipcu.VAggregate(stacked, _gray(n=4, fmt=vszipcu.GRAY16), [0])


def test_vaggregate_cpu_matches_oracle(orc):
    rng = np.random.default_rng(3)
    r, nf, H, W = 1, 5, 16, 24
    TW = 2 * r + 1
    stacks = [rng.random((TW, 2, H, W), dtype=np.float32) + 0.2 for _ in range(nf)]
    src = Clip.from_frames([rng.random((H, W), dtype=np.float32) for _ in range(nf)], GRAYS)
    stacked = Clip.from_frames([s.reshape(TW * 2 * H, W) for s in stacks], GRAYS)
    out = core.vszipcu.VAggregate(stacked, src, [0])
    for n in range(nf):
        slots = [stacks[min(max(n - r + i, 0), nf - 1)] for i in range(TW)]
        want = orc.vaggregate(slots, W, r, n, nf)
        assert np.array_equal(out.get_frame(n).planes[0], want)


def test_kernel_params_geometry_matches_oracle(orc):
    from vszipcu import device
    for (w, h, bs) in [(640, 360, 8), (1920, 1080, 4), (37, 29, 3), (8, 8, 1)]:
        kp = device.kparams(w, h, block_step=bs)
        assert device.num_blocks(kp) == orc.num_blocks(orc.make_params(w, h, block_step=bs))


@pytest.mark.parametrize("kw,plan", [
    (dict(width=1920, height=1080, block_step=4, bm_range=16), "lds4"),
    (dict(width=1920, height=1080, block_step=4, bm_range=16, metric="zssd"), "lds_rows2"),
    (dict(width=640, height=360, block_step=4, bm_range=9, metric="sad", radius=0), "lds4"),
    (dict(width=640, height=360, block_step=5, bm_range=9), "lds_rows4"),
    (dict(width=200, height=150, block_step=8, bm_range=9), "lds_rows4"),
    (dict(width=3840, height=2160, block_step=7, bm_range=9), "lds_short"),
    (dict(width=1920, height=1080, block_step=3, bm_range=9, metric="zsad"), "lds_rows1"),
    (dict(width=200, height=150, block_step=8, bm_range=48), "global"),
    (dict(width=120, height=96, block_step=8, bm_range=48, radius=1), "global+temporal"),
    (dict(width=264, height=264, block_step=1, bm_range=57, radius=1), "global+temporal"),
    (dict(width=3840, height=2160, block_step=8, bm_range=9, radius=1), "lds_rows2+temporal"),
    (dict(width=72, height=64, block_step=4, bm_range=16, radius=2), "lds4+temporal"),
])
def test_match_plan_selects_the_parity_tested_kernels(kw, plan):
    """The GPU parity cases of the global-memory and short-tile matchers rely on these choices."""
    from vszipcu import device
    w, h = kw.pop("width"), kw.pop("height")
    assert device.match_plan(device.kparams(w, h, **kw)) == plan


@pytest.mark.parametrize("kw,plan", [
    # the 1080p bench frame's hard pass: one round of 768 workgroups, three per CU
    (dict(width=1920, height=1080, block_step=4, bm_range=16, chroma=True), (10, 17, 3)),
    # its Wiener pass keeps two 16 x 16-block tiles per CU (245 VGPRs)
    (dict(width=1920, height=1080, block_step=4, bm_range=16, chroma=True, final=True), (16, 16, 2)),
    # small planes cannot fill a round of three: the planner for two per CU (small tiles, which
    # then fit three per CU anyway)
    (dict(width=640, height=360, block_step=8, bm_range=9), (2, 4, 3)),
    (dict(width=640, height=360, block_step=8, bm_range=9, final=True), (2, 4, 2)),
    # split-BF16 tiles
    (dict(width=1920, height=1080, block_step=4, bm_range=16, chroma=True, precision="split-bf16"), None),
    # radius > 0 and windows beyond the LDS budget filter with direct atomics
    (dict(width=1920, height=1080, block_step=8, bm_range=9, radius=2), (0, 0, 0)),
    (dict(width=640, height=480, block_step=8, bm_range=120), (0, 0, 0)),
    # a window beyond the matcher's LDS budget still filters in single-block tiles
    (dict(width=200, height=150, block_step=8, bm_range=48), (1, 1, 3)),
])
def test_filter_plan(kw, plan):
    """Tile plans of the LDS-tiled filter (the three-per-CU hard-pass plan is parity-tested on the
    GPU by the full-size 1080p test)."""
    from vszipcu import device
    w, h = kw.pop("width"), kw.pop("height")
    got = device.filter_plan(device.kparams(w, h, **kw))
    if plan is None:  # split-BF16 hard pass: 32 KB regions, five workgroups per CU
        assert got[0] > 0 and got[1] > 0 and got[2] == 5, got
    else:
        assert got == plan
