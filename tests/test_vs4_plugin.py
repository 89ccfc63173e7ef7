"""The VapourSynth API v4 plugin entry (VapourSynthPluginInit2) driven by a test host.

The host (tests/vs4host) loads libvszipcu_amd.so as a VapourSynth plugin and plays the core:
signature checks in invoke, the arInitial / arAllFramesReady protocol, StrictSpatial request
checks, and reference counts of frames and nodes. CPU tests cover registration (identifier,
namespace, signatures of src/plugin.zig:84-98), create-time errors, BM3Dv2 passthrough and the
CPU VAggregate; GPU tests compare full frames with the ctypes mirror and the oracle.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "vs4host"))

import host  # noqa: E402
import vszipcu  # noqa: E402
from vszipcu import GRAY16, GRAYS, YUV444PS, Clip, _lib, core, synth  # noqa: E402

BM3D_SIG = ("clip:vnode;ref:vnode:opt;sigma:float[]:opt;block_step:int[]:opt;"
            "bm_range:int[]:opt;radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;"
            "chroma:int:opt;device_id:int:opt;num_streams:int:opt;"
            "extractor_exp:int:opt;bm_error_s:data[]:opt;transform_2d_s:data[]:opt;"
            "transform_1d_s:data[]:opt;zero_init:int:opt;fast_fused:int:opt;")


@pytest.fixture(scope="module")
def vs():
    h = host.Host()
    assert h.load(_lib.LIB_PATH) == 3
    return h


@pytest.fixture(autouse=True)
def clean_host():
    before = host.live()
    yield
    assert host.live() == before, "frame or node references leaked"
    n, log = host.misuse()
    assert n == 0, log


def _frames(w, h, n, fmt=GRAYS, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        planes = [rng.random((h, w), dtype=np.float32)]
        if fmt.num_planes == 3:
            planes += [rng.random((h >> fmt.sub_sampling_h, w >> fmt.sub_sampling_w), dtype=np.float32)
                       for _ in range(2)]
        out.append(planes)
    return out


def test_registration(vs):
    p = vs.plugin()
    assert (p["id"], p["namespace"], p["name"]) == ("com.julek.vszipcu", "vszipcu", "VapourSynth Zig Image Process CUDA")
    assert p["version"] == 1 and p["api"] == (4 << 16)
    assert vs.functions["BM3D"] == (BM3D_SIG, "clip:vnode;")
    assert vs.functions["BM3Dv2"] == (BM3D_SIG, "clip:vnode;")
    assert vs.functions["VAggregate"] == ("clip:vnode;src:vnode;planes:int[];", "clip:vnode;")


@pytest.mark.parametrize("kwargs,msg", [
    (dict(block_step=9), 'BM3D: "block_step" must be in range [1, 8].'),
    (dict(sigma=[3, -1]), 'BM3D: "sigma" must be non-negative.'),
    (dict(radius=17), 'BM3D: "radius" must be <= 16.'),
    (dict(chroma=1), 'BM3D: clip format must be YUV444 when "chroma" is true.'),
    (dict(bm_error_s=["ssd", "bad"]), 'BM3D: invalid "bm_error_s".'),
    (dict(num_streams=40), "BM3D: num_streams must be 1..32."),
    (dict(nonsense=1), "BM3D: argument nonsense does not exist"),
    (dict(radius=[1, 2]), "BM3D: argument radius is not an array"),
    (dict(sigma="x"), "BM3D: argument sigma is not of the correct type"),
])
def test_create_errors_reach_the_output_map(vs, kwargs, msg):
    src = vs.source(_frames(32, 24, 2), GRAYS)
    try:
        with pytest.raises(host.HostError) as ei:
            vs.invoke("BM3D", clip=src, **kwargs)
        assert str(ei.value) == msg
    finally:
        src.free()


@pytest.mark.parametrize("kwargs", [dict(block_step=9), dict(radius=17, sigma=3), dict(chroma=1)])
def test_bm3dv2_reports_the_invoked_bm3d_error(vs, kwargs):
    """The composed BM3Dv2 returns the error map of the BM3D it invokes (src/bm3d.zig:1469-1472)."""
    src = vs.source(_frames(32, 24, 2), GRAYS)
    try:
        with pytest.raises(host.HostError) as e1:
            vs.invoke("BM3D", clip=src, **kwargs)
        with pytest.raises(host.HostError) as e2:
            vs.invoke("BM3Dv2", clip=src, **kwargs)
        assert str(e2.value) == str(e1.value) and str(e1.value).startswith("BM3D: ")
    finally:
        src.free()


def test_format_error_and_required_clip(vs):
    src = vs.source([[np.zeros((24, 32), np.float32)]], GRAY16)
    try:
        with pytest.raises(host.HostError, match="only constant format 32 bit float input supported"):
            vs.invoke("BM3Dv2", clip=src, sigma=3)
        with pytest.raises(host.HostError, match="argument clip is required"):
            vs.invoke("BM3D", sigma=3)
    finally:
        src.free()


def test_without_gpu_creation_fails_cleanly(vs):
    if vszipcu.load().vsz_device_count() != 0:
        pytest.skip("a GPU is visible")
    src = vs.source(_frames(32, 24, 2), GRAYS)
    try:
        with pytest.raises(host.HostError, match="HIP initialization failed"):
            vs.invoke("BM3D", clip=src, sigma=3)
    finally:
        src.free()


def test_bm3dv2_passthrough_returns_the_input_node(vs):
    fr = _frames(32, 24, 3, YUV444PS)
    src = vs.source(fr, YUV444PS)
    out = vs.invoke("BM3Dv2", clip=src, sigma=[0, 0, 0])
    try:
        assert out.ptr == src.ptr
        planes, props = out.frame(1)
        for a, b in zip(planes, fr[1]):
            assert np.array_equal(a, b)
    finally:
        out.free()
        src.free()


def test_vaggregate_through_the_plugin(vs, orc):
    rng = np.random.default_rng(5)
    r, nf, H, W = 2, 6, 16, 24
    TW = 2 * r + 1
    stacks = [rng.random((TW, 2, H, W), dtype=np.float32) + 0.25 for _ in range(nf)]
    srcf = [[rng.random((H, W), dtype=np.float32)] for _ in range(nf)]
    stacked = vs.source([[s.reshape(TW * 2 * H, W)] for s in stacks], GRAYS)
    src = vs.source(srcf, GRAYS)
    out = vs.invoke("VAggregate", clip=stacked, src=src, planes=[0])
    try:
        assert out.name == "VAggregate" and out.deps() == [0, 2]   # General, StrictSpatial
        assert out.info()["height"] == H
        for n in range(nf):
            planes, props = out.frame(n)
            slots = [stacks[min(max(n - r + i, 0), nf - 1)] for i in range(TW)]
            assert np.array_equal(planes[0], orc.vaggregate(slots, W, r, n, nf))
            assert props["_SourceFrame"] == [n]       # props come from src frame n
        with pytest.raises(host.HostError, match='VAggregate: "planes" is required'):
            vs.invoke("VAggregate", clip=stacked, src=src, planes=[])
    finally:
        out.free()
        stacked.free()
        src.free()


# ---- GPU --------------------------------------------------------------------------------------

def _mirror_frames(clip, n):
    return [np.asarray(p) for p in clip.get_frame(n).planes]


@pytest.mark.gpu
def test_bm3d_radius0_matches_oracle(vs, orc, torch_gpu):
    W, H = 64, 48
    clean, nz = synth.gray_frame(W, H, 10.0, seed=31)
    src = vs.source([[nz]], GRAYS)
    out = vs.invoke("BM3D", clip=src, sigma=10, block_step=4, bm_range=9, extractor_exp=3)
    try:
        assert out.name == "BM3D" and out.deps() == [2]            # StrictSpatial at radius 0
        planes, props = out.frame(0)
        from vszipcu import device
        p = orc.make_params(W, H, block_step=4, bm_range=9, sigma=tuple(device.scaled_sigma([10.0] * 3, False)),
                            extractor_exp=3)
        res, _, _ = orc.bm3d(p, nz[None, None, None])
        assert np.array_equal(planes[0], orc.aggregate(res[:, 0], W)[0])
        assert props["_SourceFrame"] == [0]
    finally:
        out.free()
        src.free()


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [0, 1])
def test_bm3dv2_temporal_two_pass_matches_mirror(vs, torch_gpu, fused):
    W, H, nf = 48, 40, 4
    _, seq = synth.moving_sequence(W, H, nf, 8.0, seed=32)
    frames = [[f] for f in seq]
    src = vs.source(frames, GRAYS)
    mclip = Clip.from_frames(frames, GRAYS)
    kw = dict(sigma=8, radius=1, block_step=4, bm_range=6, extractor_exp=3, fast_fused=fused)
    basic = vs.invoke("BM3Dv2", clip=src, **kw)
    final = vs.invoke("BM3Dv2", clip=src, ref=basic, **kw)
    mb = core.vszipcu.BM3Dv2(mclip, **kw)
    mf = core.vszipcu.BM3Dv2(mclip, ref=mb, **kw)
    try:
        if fused:   # the fused chain is one filter over clip and ref
            assert final.name == "BM3Dv2" and final.deps() == [0, 0]
        else:       # composed: VAggregate (General over BM3D's stack, StrictSpatial over src)
            assert final.name == "VAggregate" and final.deps() == [0, 2]
        for n in range(nf):
            got, _ = final.frame(n)
            assert np.array_equal(got[0], _mirror_frames(mf, n)[0])
    finally:
        for x in (final, basic, src):
            x.free()


@pytest.mark.gpu
def test_bm3dv2_radius0_is_the_registered_bm3d(vs, torch_gpu):
    """BM3Dv2 at radius 0 returns the node the registered BM3D creates (src/bm3d.zig:1474-1479)."""
    W, H = 48, 40
    _, nz = synth.gray_frame(W, H, 8.0, seed=34)
    src = vs.source([[nz]], GRAYS)
    kw = dict(sigma=8, block_step=4, bm_range=6, extractor_exp=3)
    v2 = vs.invoke("BM3Dv2", clip=src, **kw)
    v1 = vs.invoke("BM3D", clip=src, **kw)
    try:
        assert v2.name == "BM3D" and v2.deps() == [2]
        assert np.array_equal(v2.frame(0)[0][0], v1.frame(0)[0][0])
    finally:
        for x in (v2, v1, src):
            x.free()


@pytest.mark.gpu
def test_stacked_output_props_and_zero_init(vs, torch_gpu):
    W, H, nf = 40, 32, 3
    fr = _frames(W, H, nf, YUV444PS, seed=33)
    src = vs.source(fr, YUV444PS)
    out = vs.invoke("BM3D", clip=src, sigma=[6, 0], radius=1, block_step=4, bm_range=5, extractor_exp=3)
    try:
        assert out.info()["height"] == H * 2 * 3 and out.deps() == [0]
        planes, props = out.frame(1)
        assert props["BM3D_V_radius"] == [1] and props["BM3D_V_process"] == [1, 0, 0]
        assert not planes[1].any() and not planes[2].any()      # zero_init on unprocessed planes
        want = _mirror_frames(core.vszipcu.BM3D(Clip.from_frames(fr, YUV444PS), sigma=[6, 0], radius=1,
                                                block_step=4, bm_range=5, extractor_exp=3), 1)
        assert np.array_equal(planes[0], want[0])
    finally:
        out.free()
        src.free()
