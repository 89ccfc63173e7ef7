"""GPU parity of the kernel layer (vsz_bm3d_match / vsz_bm3d_run / vsz_vaggregate_run) against
the FP32 oracle on the same seeded inputs.

Bar (from the build's contract): block-match indices exact; pixels within 1e-5 max-abs. The
kernels follow the oracle's op sequence, so the checks below are stricter where they can be:
block distances bit-exact, and with a non-zero extractor (exact integer aggregation) the
num/den stack and the output bit-exact. With extractor 0 (FP32 atomics, order-dependent by
design, as in the reference) num/den are compared with a relative tolerance and the output
within 1e-5.
"""
import numpy as np
import pytest

from gpu_util import assert_matches_equal, gpu_pass, kp_and_op
from vszipcu import device, synth

pytestmark = pytest.mark.gpu


def _noisy(W, H, seed, sigma=10.0):
    c = synth.clean_luma(W, H, seed)
    return synth.noisy(c, sigma, seed + 7)


def _check(orc, torch, kp, op, src, *, exact=True):
    ores, om, oe = orc.bm3d(op, src)
    gres, gm = gpu_pass(torch, kp, src, 1)
    assert_matches_equal(gm, om, oe)
    if exact:
        assert np.array_equal(gres, ores), f"max diff {np.max(np.abs(gres - ores))}"
    else:
        np.testing.assert_allclose(gres, ores, rtol=2e-6, atol=1e-6)
    if kp.radius == 0:
        gdst, _ = gpu_pass(torch, kp, src, 0)
        odst = orc.aggregate(ores[:, 0], kp.width, kp.proc_mask)
        P = ores.shape[0]
        for p in range(P):
            if (kp.proc_mask >> p) & 1:
                g, o = gdst[p, :, :kp.width], odst[p, :, :kp.width]
                if exact:
                    assert np.array_equal(g, o, equal_nan=True), f"plane {p}: {np.nanmax(np.abs(g - o))}"
                else:
                    np.testing.assert_allclose(g, o, rtol=0, atol=1e-5, equal_nan=True)
    return ores


@pytest.mark.parametrize("W,H,bs,rng", [(64, 48, 8, 9), (37, 29, 3, 5), (80, 45, 4, 16), (8, 8, 1, 4)])
def test_basic_pass_gray(orc, torch_gpu, W, H, bs, rng):
    img = _noisy(W, H, W + H)
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, sigma=device.scaled_sigma([10.0] * 3, False))
    _check(orc, torch_gpu, kp, op, img[None, None, None])


@pytest.mark.parametrize("metric", ["sad", "zssd", "zsad", "ssd/norm"])
def test_metrics(orc, torch_gpu, metric):
    img = _noisy(48, 40, 3)
    kp, op = kp_and_op(orc, 48, 40, bs=4, rng=6, metric=metric, sigma=device.scaled_sigma([8.0] * 3, False))
    _check(orc, torch_gpu, kp, op, img[None, None, None])


@pytest.mark.parametrize("t2d", ["dct", "haar", "wht", "bior1.5"])
@pytest.mark.parametrize("t1d", ["dct", "haar", "wht", "bior1.5"])
def test_transforms(orc, torch_gpu, t2d, t1d):
    img = _noisy(40, 32, 4)
    kp, op = kp_and_op(orc, 40, 32, bs=6, rng=5, t2d=t2d, t1d=t1d, sigma=device.scaled_sigma([6.0] * 3, False))
    _check(orc, torch_gpu, kp, op, img[None, None, None])


def test_padded_stride(orc, torch_gpu):
    W, H, S = 45, 30, 64
    img = np.zeros((H, S), np.float32)
    img[:, :W] = _noisy(W, H, 5)
    img[:, W:] = np.nan  # padding must never be read
    kp, op = kp_and_op(orc, W, H, bs=4, rng=7, stride=S, sigma=device.scaled_sigma([10.0] * 3, False))
    _check(orc, torch_gpu, kp, op, img[None, None, None])


def test_chroma_with_skip_quirk(orc, torch_gpu):
    W, H = 48, 40
    clean, nz = synth.yuv444_frame(W, H, seed=6)
    # plane 2: raw sigma >= FLT_EPSILON (processed) but scaled below it (never accumulated ->
    # 0/0 = NaN, src/bm3d.cu:804-810)
    sig = device.scaled_sigma([5.0, 3.0, 1.5e-7], False)
    kp, op = kp_and_op(orc, W, H, bs=4, rng=8, chroma=True, sigma=sig)
    ores = _check(orc, torch_gpu, kp, op, nz[None, :, None])
    assert np.all(ores[2] == 0)
    gdst, _ = gpu_pass(torch_gpu, kp, nz[None, :, None], 0)
    assert np.all(np.isnan(gdst[2]))


@pytest.mark.parametrize("radius,ps_num,ps_range", [(1, 2, 4), (2, 3, 2)])
def test_temporal(orc, torch_gpu, radius, ps_num, ps_range):
    clean, nz = synth.moving_sequence(40, 32, 2 * radius + 1, 8.0, seed=radius)
    kp, op = kp_and_op(orc, 40, 32, bs=4, rng=6, radius=radius, ps_num=ps_num, ps_range=ps_range,
                       sigma=device.scaled_sigma([8.0] * 3, False))
    _check(orc, torch_gpu, kp, op, nz[None, None])


def test_wiener_pass_on_oracle_basic(orc, torch_gpu):
    W, H = 64, 48
    nz = _noisy(W, H, 8)
    kp_b, op_b = kp_and_op(orc, W, H, bs=4, rng=9, sigma=device.scaled_sigma([10.0] * 3, False))
    ores_b, _, _ = orc.bm3d(op_b, nz[None, None, None])
    basic = orc.aggregate(ores_b[:, 0], W)[0]
    kp, op = kp_and_op(orc, W, H, bs=4, rng=9, final=True, sigma=device.scaled_sigma([10.0] * 3, True))
    _check(orc, torch_gpu, kp, op, np.stack([basic, nz])[:, None, None])


def test_float_accumulation_extractor0(orc, torch_gpu):
    img = _noisy(64, 48, 9)
    kp, op = kp_and_op(orc, 64, 48, bs=4, rng=8, ext=0, sigma=device.scaled_sigma([10.0] * 3, False))
    _check(orc, torch_gpu, kp, op, img[None, None, None], exact=False)


def test_deterministic_runs(orc, torch_gpu):
    img = _noisy(96, 64, 10)
    kp, _ = kp_and_op(orc, 96, 64, bs=4, rng=12, sigma=device.scaled_sigma([10.0] * 3, False))
    a, ma = gpu_pass(torch_gpu, kp, img[None, None, None], 1)
    b, mb = gpu_pass(torch_gpu, kp, img[None, None, None], 1)
    assert np.array_equal(a, b) and np.array_equal(ma, mb)


def test_vaggregate_kernel(orc, torch_gpu):
    rng = np.random.default_rng(1)
    r, nf, n, H, W = 2, 6, 4, 20, 24
    TW = 2 * r + 1
    slots = [rng.random((TW, 2, H, W), dtype=np.float32) + 0.1 for _ in range(TW)]
    want = orc.vaggregate(slots, W, r, n, nf)
    ds = [torch_gpu.from_numpy(s).cuda() for s in slots]
    dst = torch_gpu.zeros((H, W), dtype=torch_gpu.float32, device="cuda")
    device.vaggregate(W, r, n, nf, ds, dst)
    torch_gpu.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), want)


@pytest.mark.parametrize("W,H,rng,metric", [(200, 150, 4, "ssd"), (68, 70, 16, "ssd"), (236, 101, 16, "sad"),
                                            (132, 90, 30, "ssd"), (320, 240, 12, "sad"), (20, 60, 9, "ssd")])
def test_shared_column_matcher_block_step4(orc, torch_gpu, W, H, rng, metric):
    """block_step 4 with (W-8) % 4 == 0 takes the shared-column matcher (columns 0..3 of a block
    taken from its left neighbour); several 15 x 16 block tiles, partial tiles, clamped rows
    ((H-8) % 4 != 0), windows larger than the plane."""
    print(f"shared-column matcher: {W}x{H} R={rng} {metric}", flush=True)
    torch = torch_gpu
    img = _noisy(W, H, W * 3 + H)
    kp, op = kp_and_op(orc, W, H, bs=4, rng=rng, metric=metric, sigma=device.scaled_sigma([10.0] * 3, False))
    om, oe = orc.match(op, img[None])
    nbx, nby = device.num_blocks(kp)
    m = torch.zeros((nby, nbx, 8, 4), dtype=torch.int32, device="cuda")
    device.match(kp, torch.from_numpy(img[None]).cuda(), m)
    torch.cuda.synchronize()
    assert_matches_equal(m.cpu().numpy(), om, oe)


@pytest.mark.parametrize("W,H,bs,rng,metric", [(200, 150, 8, 9, "ssd"), (333, 121, 8, 9, "zsad"), (140, 260, 7, 12, "ssd")])
def test_lds_matcher_split_rows(orc, torch_gpu, W, H, bs, rng, metric):
    """Small planes at other block steps (the plugin default block_step 8 with bm_range 9) take the
    LDS-staged matcher with each block's search rows split over 4 lanes."""
    print(f"split-row matcher: {W}x{H} bs={bs} R={rng} {metric}", flush=True)
    torch = torch_gpu
    img = _noisy(W, H, W + 5 * H)
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, metric=metric, sigma=device.scaled_sigma([10.0] * 3, False))
    om, oe = orc.match(op, img[None])
    nbx, nby = device.num_blocks(kp)
    m = torch.zeros((nby, nbx, 8, 4), dtype=torch.int32, device="cuda")
    device.match(kp, torch.from_numpy(img[None]).cuda(), m)
    torch.cuda.synchronize()
    assert_matches_equal(m.cpu().numpy(), om, oe)


@pytest.mark.parametrize("bs,rng,ps_num,ps_range", [(8, 9, 2, 3), (4, 16, 2, 3), (8, 9, 2, 4), (4, 16, 3, 4)])
def test_temporal_split_matcher(orc, torch_gpu, bs, rng, ps_num, ps_range):
    """radius > 0: centre-frame search on an LDS-staged matcher, then the predictive search
    (range 4 takes the register-band scanner, whose windows here often cross the frame edges)."""
    print(f"temporal split: bs={bs} R={rng} ps_num={ps_num} ps_range={ps_range}", flush=True)
    _, seq = synth.moving_sequence(72, 64, 5, 8.0, seed=41)
    src = np.stack(seq)[None, None]                            # [1, 1, 5, H, W]
    kp, op = kp_and_op(orc, 72, 64, bs=bs, rng=rng, radius=2, ps_num=ps_num, ps_range=ps_range,
                       sigma=device.scaled_sigma([8.0] * 3, False))
    _check(orc, torch_gpu, kp, op, src)
