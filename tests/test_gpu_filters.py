"""GPU tests of the filter layer (vsz_bm3d_create / vsz_bm3dv2_create / vsz_vaggregate_create /
vsz_filter_get_frame) through the host mirror `core.vszipcu`, against oracle expectations built
with the reference's host rules (sigma scaling, per-plane entries, frame windows, stacked
layout, aggZ). extractor_exp=3 everywhere, so every comparison is bit-exact.
"""
import threading

import numpy as np
import pytest

from vszipcu import GRAYS, YUV420PS, YUV444PS, Clip, core, device, synth

pytestmark = pytest.mark.gpu


def _seq(W, H, nf, seed=0, sigma=8.0):
    clean, nz = synth.moving_sequence(W, H, nf, sigma, seed=seed)
    return clean, nz


def _oracle_plane(orc, frames_tw, *, sigma255, final=False, ref_tw=None, radius=0, **kw):
    """frames_tw: [TW, H, W] window of one plane -> res [TW, 2, H, W] (exact)."""
    TW, H, W = frames_tw.shape
    sig = device.scaled_sigma([sigma255] * 3, final)
    p = orc.make_params(W, H, radius=radius, sigma=tuple(sig), final=final, extractor_exp=3, **kw)
    src = np.stack([ref_tw, frames_tw])[:, None] if final else frames_tw[None, None]
    res, _, _ = orc.bm3d(p, src)
    return res[0]


def _window(frames, n, r):
    nf = len(frames)
    return np.stack([frames[min(max(n - r + t, 0), nf - 1)] for t in range(2 * r + 1)])


def test_bm3d_gray_radius0(orc, torch_gpu):
    _, nz = _seq(72, 40, 3)
    clip = Clip.from_frames(list(nz), GRAYS)
    out = core.vszipcu.BM3D(clip, sigma=8.0, block_step=4, bm_range=7, extractor_exp=3)
    for n in range(3):
        res = _oracle_plane(orc, nz[n][None], sigma255=8.0, block_step=4, bm_range=7)
        want = res[0, 0] / res[0, 1]
        assert np.array_equal(out.get_frame(n).planes[0], want)


def test_bm3d_two_pass_chroma(orc, torch_gpu):
    W, H = 64, 48
    clean, nz = synth.yuv444_frame(W, H, seed=3)
    clip = Clip.from_frames([list(nz)], YUV444PS)
    kw = dict(sigma=[5.0, 3.0, 3.0], block_step=4, bm_range=8, chroma=1, extractor_exp=3)
    basic = core.vszipcu.BM3D(clip, **kw)
    final = core.vszipcu.BM3D(clip, ref=basic, **kw)
    gb = np.stack(basic.get_frame(0).planes)
    gf = np.stack(final.get_frame(0).planes)
    sb = device.scaled_sigma(kw["sigma"], False)
    sf = device.scaled_sigma(kw["sigma"], True)
    pb = orc.make_params(W, H, block_step=4, bm_range=8, chroma=True, sigma=tuple(sb), extractor_exp=3)
    rb, _, _ = orc.bm3d(pb, nz[None, :, None])
    ob = orc.aggregate(rb[:, 0], W)
    assert np.array_equal(gb, ob)
    pf = orc.make_params(W, H, block_step=4, bm_range=8, chroma=True, final=True, sigma=tuple(sf), extractor_exp=3)
    rf, _, _ = orc.bm3d(pf, np.stack([ob, nz])[:, :, None])
    of = orc.aggregate(rf[:, 0], W)
    assert np.array_equal(gf, of)
    assert synth.psnr(gf[0], clean[0]) > synth.psnr(nz[0], clean[0]) + 0.5


def test_unprocessed_planes_pass_through(orc, torch_gpu):
    W, H = 48, 32
    _, nz = synth.yuv444_frame(W, H, seed=4)
    clip = Clip.from_frames([list(nz)], YUV444PS)
    out = core.vszipcu.BM3D(clip, sigma=[5.0, 0.0], block_step=8, extractor_exp=3)
    f = out.get_frame(0).planes
    assert np.array_equal(f[1], nz[1]) and np.array_equal(f[2], nz[2])
    res = _oracle_plane(orc, nz[0][None], sigma255=5.0, block_step=8, bm_range=9)
    assert np.array_equal(f[0], res[0, 0] / res[0, 1])


def test_subsampled_planes_separately(orc, torch_gpu):
    W, H = 64, 48
    rng = np.random.default_rng(5)
    planes = [rng.random((H, W), dtype=np.float32), rng.random((H // 2, W // 2), dtype=np.float32),
              rng.random((H // 2, W // 2), dtype=np.float32)]
    clip = Clip.from_frames([planes], YUV420PS)
    out = core.vszipcu.BM3D(clip, sigma=[6.0, 4.0], block_step=[4, 2], bm_range=[8, 5], extractor_exp=3)
    f = out.get_frame(0).planes
    for p, (bs, rg, sg) in enumerate([(4, 8, 6.0), (2, 5, 4.0), (2, 5, 4.0)]):
        res = _oracle_plane(orc, planes[p][None], sigma255=sg, block_step=bs, bm_range=rg)
        assert np.array_equal(f[p], res[0, 0] / res[0, 1]), p


def test_stacked_radius1_and_props(orc, torch_gpu):
    W, H, nf, r = 40, 32, 4, 1
    _, nz = _seq(W, H, nf, seed=6)
    clip = Clip.from_frames(list(nz), GRAYS)
    out = core.vszipcu.BM3D(clip, sigma=8.0, block_step=4, bm_range=6, radius=r, extractor_exp=3)
    assert out.height == H * 2 * (2 * r + 1)
    for n in range(nf):
        fr = out.get_frame(n)
        assert fr.props["BM3D_V_radius"] == r
        assert fr.props["BM3D_V_process"] == [1, 1, 1]
        res = _oracle_plane(orc, _window(nz, n, r), sigma255=8.0, radius=r, block_step=4, bm_range=6)
        assert np.array_equal(fr.planes[0], res.reshape(-1, W)), n


@pytest.mark.parametrize("r", [1, 2])
def test_bm3dv2_composed_and_fused_identical(orc, torch_gpu, r):
    W, H, nf = 40, 32, 5
    _, nz = _seq(W, H, nf, seed=7 + r)
    clip = Clip.from_frames(list(nz), GRAYS)
    kw = dict(sigma=8.0, block_step=4, bm_range=6, radius=r, extractor_exp=3)
    composed = core.vszipcu.BM3Dv2(clip, **kw)
    fused = core.vszipcu.BM3Dv2(clip, fast_fused=1, **kw)
    stacks = [_oracle_plane(orc, _window(nz, m, r), sigma255=8.0, radius=r, block_step=4, bm_range=6)
              for m in range(nf)]
    for n in range(nf):
        a = composed.get_frame(n).planes[0]
        b = fused.get_frame(n).planes[0]
        assert np.array_equal(a, b), n
        slots = [stacks[min(max(n - r + i, 0), nf - 1)] for i in range(2 * r + 1)]
        assert np.array_equal(a, orc.vaggregate(slots, W, r, n, nf)), n


def test_concurrent_get_frame(torch_gpu):
    W, H, nf = 64, 48, 8
    _, nz = _seq(W, H, nf, seed=9)
    clip = Clip.from_frames(list(nz), GRAYS)
    kw = dict(sigma=8.0, block_step=4, bm_range=8, extractor_exp=3, num_streams=4)
    serial = [core.vszipcu.BM3D(clip, **kw).get_frame(n).planes[0] for n in range(nf)]
    out = core.vszipcu.BM3D(clip, **kw)
    got = [None] * nf

    def work(n):
        got[n] = out.get_frame(n).planes[0]

    ts = [threading.Thread(target=work, args=(n,)) for n in range(nf)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for n in range(nf):
        assert np.array_equal(got[n], serial[n])


def test_zero_copy_output_matches_copy_path(torch_gpu):
    """num_streams >= 2 writes the output frame from the kernels straight into mapped pinned
    memory (src/bm3d.zig:1203-1206); num_streams = 1 copies it device-to-host. Same bits, for a
    radius-0 pass with an unprocessed (passthrough) plane and for the fused temporal chain."""
    W, H, nf = 64, 48, 4
    clean, nz = synth.moving_sequence(W, H, nf, 8.0, seed=21)
    yuv = Clip.from_frames([[f, f[::-1].copy(), f[:, ::-1].copy()] for f in nz], YUV444PS)
    gray = Clip.from_frames(list(nz), GRAYS)
    cases = [(core.vszipcu.BM3D, yuv, dict(sigma=[8.0, 0.0, 6.0], block_step=4, bm_range=6, extractor_exp=3)),
             (core.vszipcu.BM3Dv2, gray, dict(sigma=8.0, block_step=4, bm_range=6, radius=1, fast_fused=1,
                                              extractor_exp=3))]
    for fn, clip, kw in cases:
        dma = fn(clip, num_streams=1, **kw)
        zc = fn(clip, num_streams=4, **kw)
        for n in range(nf):
            a, b = dma.get_frame(n).planes, zc.get_frame(n).planes
            for p in range(len(a)):
                assert np.array_equal(a[p], b[p]), (fn, n, p)
        if fn is core.vszipcu.BM3D:
            assert np.array_equal(zc.get_frame(1).planes[1], nz[1][::-1])  # passthrough plane
