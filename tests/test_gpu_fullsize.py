"""Parity at full size: the benchmark frame and the CPU-runnable configuration, against the FP32
oracle on the same seeded inputs (the oracle runs multithreaded; a 1080p two-pass takes a few
seconds). With extractor_exp 3 the sums are exact integers, so the bar is bit-identical match
lists, distances and outputs.

* 640x360 Gray basic pass, sigma 10, block_step 8, bm_range 9 (the CPU-runnable case);
* 1920x1080 YUV444 two-pass with chroma, sigma [5, 3, 3], block_step 4, bm_range 16 (the
  benchmark workload; its matching takes the shared-column kernel, its hard pass three 10 x 17-block
  filter tiles per CU);
* 1920x1080 Gray basic pass at block_step 3 and 8: the hard pass's other tile plans (20 x 15
  blocks three per CU, and 8 x 8 blocks from the two-per-CU planner).
"""
import numpy as np
import pytest

from gpu_util import assert_matches_equal, gpu_pass
from vszipcu import device, synth

pytestmark = pytest.mark.gpu


def _params(orc, W, H, sigma255, final, **kw):
    sig = tuple(device.scaled_sigma(sigma255, final))
    kp = device.kparams(W, H, sigma=sig, final=final, **kw)
    op = orc.make_params(W, H, sigma=sig, final=final, threads=0, **kw)
    return kp, op


def test_cpu_case_640x360_basic(orc, torch_gpu):
    print("full size: 640x360 Gray basic bs8 R9", flush=True)
    clean, nz = synth.gray_frame(640, 360, 10.0, seed=1)
    kp, op = _params(orc, 640, 360, [10.0] * 3, False, block_step=8, bm_range=9, extractor_exp=3)
    src = nz[None, None, None]
    ores, om, oe = orc.bm3d(op, src)
    gres, gm = gpu_pass(torch_gpu, kp, src, 1)
    assert_matches_equal(gm, om, oe)
    assert np.array_equal(gres, ores)
    gdst, _ = gpu_pass(torch_gpu, kp, src, 0)
    assert np.array_equal(gdst[0], orc.aggregate(ores[:, 0], 640)[0])


def test_bench_frame_1080p_yuv444_two_pass(orc, torch_gpu):
    print("full size: 1920x1080 YUV444 two-pass chroma bs4 R16", flush=True)
    W, H = 1920, 1080
    sig255 = [5.0, 3.0, 3.0]
    clean, nz = synth.yuv444_frame(W, H, sig255, seed=2)
    kw = dict(block_step=4, bm_range=16, ps_num=2, ps_range=4, chroma=True, extractor_exp=3)
    kb, ob = _params(orc, W, H, sig255, False, **kw)
    src = nz[None, :, None]                                   # [1, 3, 1, H, W]
    ores, om, oe = orc.bm3d(ob, src)
    gdst, gm = gpu_pass(torch_gpu, kb, src, 0)
    assert_matches_equal(gm, om, oe)
    basic = orc.aggregate(ores[:, 0], W)
    assert np.array_equal(gdst, basic)
    kf, of = _params(orc, W, H, sig255, True, **kw)
    src2 = np.stack([basic, nz])[:, :, None]                  # [2, 3, 1, H, W]
    ores2, om2, oe2 = orc.bm3d(of, src2)
    gdst2, gm2 = gpu_pass(torch_gpu, kf, src2, 0)
    assert_matches_equal(gm2, om2, oe2)
    final = orc.aggregate(ores2[:, 0], W)
    assert np.array_equal(gdst2, final)
    psnr = synth.psnr(final[0], clean[0])
    print(f"  bit-identical; PSNR Y {psnr:.3f} dB", flush=True)
    assert psnr > synth.psnr(nz[0], clean[0]) + 3.0


@pytest.mark.parametrize("bs,rng,tiles", [(3, 9, (20, 15, 3)), (8, 12, (8, 8, 3))])
def test_1080p_gray_basic_tile_plans(orc, torch_gpu, bs, rng, tiles):
    print(f"full size: 1920x1080 Gray basic bs{bs} R{rng}, filter tiles {tiles}", flush=True)
    W, H = 1920, 1080
    clean, nz = synth.gray_frame(W, H, 12.0, seed=3 + bs)
    kp, op = _params(orc, W, H, [12.0] * 3, False, block_step=bs, bm_range=rng, extractor_exp=3)
    assert device.filter_plan(kp) == tiles
    src = nz[None, None, None]
    ores, om, oe = orc.bm3d(op, src)
    gres, gm = gpu_pass(torch_gpu, kp, src, 1)
    assert_matches_equal(gm, om, oe)
    assert np.array_equal(gres, ores)
