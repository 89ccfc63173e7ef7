"""GPU parity, against the FP32 oracle, of the kernel paths and configurations the small-frame
parity tests leave out:

* the Wiener (final) pass at radius > 0, Gray and chroma (clip 0 = basic estimate gathered for the
  ref group, clip 1 = noisy, src/bm3d.cu:797-836), including the 4K radius-1 Wiener workload;
* the chroma basic pass at radius > 0, and a chroma plane whose scaled sigma is below FLT_EPSILON;
* the global-memory matcher (windows beyond the LDS budget), spatial and temporal, and the
  centre-frame search on it ahead of the temporal stage (planes of 65536+ blocks);
* the batched plan of uneven planes, per plane, at the batched workload's sizes.

Extractor 3 everywhere: the sums are exact integers, so match lists, distances, the num/den stack
and divided outputs must be bit-identical to the oracle.
"""
import numpy as np
import pytest

from gpu_util import assert_matches_equal, gpu_pass, kp_and_op
from vszipcu import device, synth

pytestmark = pytest.mark.gpu


def _basic_frames(orc, frames, sigma255, chroma=False, bs=8, rng=9):
    """Radius-0 basic estimate of every frame by the oracle. frames: [F, P, H, W] -> same shape."""
    F, P, H, W = frames.shape
    sig = device.scaled_sigma(list(sigma255) if chroma else [sigma255] * 3, False)
    op = orc.make_params(W, H, block_step=bs, bm_range=rng, chroma=chroma, sigma=tuple(sig), extractor_exp=3,
                         threads=0)
    out = np.empty_like(frames)
    for f in range(F):
        res, _, _ = orc.bm3d(op, frames[f][None, :, None])
        out[f] = orc.aggregate(res[:, 0], W)
    return out


def _check_stack(orc, torch, kp, op, src):
    """Radius > 0: the [P][TW][2][H][S] accumulator stack and the match lists, bit for bit."""
    ores, om, oe = orc.bm3d(op, src)
    gres, gm = gpu_pass(torch, kp, src, 1)
    assert_matches_equal(gm, om, oe)
    same = np.array_equal(gres, ores, equal_nan=True)
    assert same, f"stack differs: max {np.nanmax(np.abs(gres - ores))}"
    return ores


def _yuv_sequence(W, H, F, sig255, seed):
    """[F, 3, H, W] noisy YUV444 frames: moving luma, chroma planes drifting one pixel a frame."""
    _, y = synth.moving_sequence(W, H, F, sig255[0], seed=seed)
    u = synth.clean_chroma(W + F, H + F, seed + 1)
    v = synth.clean_chroma(W + F, H + F, seed + 2)
    out = np.empty((F, 3, H, W), np.float32)
    for f in range(F):
        out[f, 0] = y[f]
        out[f, 1] = synth.noisy(u[f:f + H, f:f + W], sig255[1], seed + 10 + f)
        out[f, 2] = synth.noisy(v[F - f - 1:F - f - 1 + H, f:f + W], sig255[2], seed + 20 + f)
    return out


@pytest.mark.parametrize("radius,bs,rng,ps_num,ps_range", [(1, 8, 9, 2, 4), (2, 8, 9, 2, 4), (1, 4, 16, 2, 4),
                                                           (2, 4, 8, 3, 3)])
def test_wiener_temporal_gray(orc, torch_gpu, radius, bs, rng, ps_num, ps_range):
    print(f"wiener temporal gray: r={radius} bs={bs} R={rng} ps={ps_num}/{ps_range}", flush=True)
    TW = 2 * radius + 1
    W, H = 96, 72
    _, nz = synth.moving_sequence(W, H, TW, 10.0, seed=50 + radius + bs)
    basic = _basic_frames(orc, nz[:, None], 10.0, bs=bs, rng=rng)[:, 0]
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, radius=radius, ps_num=ps_num, ps_range=ps_range, final=True,
                       sigma=device.scaled_sigma([10.0] * 3, True))
    _check_stack(orc, torch_gpu, kp, op, np.stack([basic, nz])[:, None])


@pytest.mark.parametrize("radius,bs,rng,skip", [(1, 8, 9, False), (2, 4, 12, False), (1, 4, 9, True)])
def test_wiener_temporal_chroma(orc, torch_gpu, radius, bs, rng, skip):
    """chroma shares plane 0's matches; with `skip` plane 2's raw sigma is processed but its
    scaled sigma is below FLT_EPSILON, so it is never accumulated (src/bm3d.cu:804-810)."""
    print(f"wiener temporal chroma: r={radius} bs={bs} R={rng} skip={skip}", flush=True)
    TW = 2 * radius + 1
    W, H = 80, 64
    sig255 = [6.0, 4.0, 1.5e-7 if skip else 3.0]
    nz = _yuv_sequence(W, H, TW, sig255, seed=60 + radius)
    basic = _basic_frames(orc, nz, [6.0, 4.0, 3.0], chroma=True, bs=bs, rng=rng)
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, radius=radius, chroma=True, final=True,
                       sigma=device.scaled_sigma(sig255, True))
    src = np.stack([basic.transpose(1, 0, 2, 3), nz.transpose(1, 0, 2, 3)])   # [2, 3, TW, H, W]
    ores = _check_stack(orc, torch_gpu, kp, op, src)
    if skip:
        assert np.all(ores[2] == 0)
    else:
        assert np.all(ores[:, radius, 1] > 0)


@pytest.mark.parametrize("radius,bs,rng", [(1, 4, 8), (1, 8, 9), (2, 6, 7)])
def test_basic_temporal_chroma(orc, torch_gpu, radius, bs, rng):
    print(f"basic temporal chroma: r={radius} bs={bs} R={rng}", flush=True)
    TW = 2 * radius + 1
    W, H = 88, 56
    sig255 = [8.0, 5.0, 4.0]
    nz = _yuv_sequence(W, H, TW, sig255, seed=70 + radius)
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, radius=radius, chroma=True,
                       sigma=device.scaled_sigma(sig255, False))
    _check_stack(orc, torch_gpu, kp, op, nz.transpose(1, 0, 2, 3)[None])


@pytest.mark.parametrize("metric,final", [("ssd", False), ("zssd", False), ("ssd", True)])
def test_global_matcher_spatial(orc, torch_gpu, metric, final):
    """block_step 8 with bm_range 48: the staged region exceeds the LDS budget, so matching runs
    on the global-memory matcher (raw buffer loads)."""
    print(f"global matcher spatial: 200x150 bs=8 R=48 {metric} final={final}", flush=True)
    W, H = 200, 150
    _, nz = synth.gray_frame(W, H, 10.0, seed=80)
    if final:
        basic = _basic_frames(orc, nz[None, None], 10.0)[0, 0]
        src = np.stack([basic, nz])[:, None, None]
    else:
        src = nz[None, None, None]
    kp, op = kp_and_op(orc, W, H, bs=8, rng=48, metric=metric, final=final,
                       sigma=device.scaled_sigma([10.0] * 3, final))
    assert device.match_plan(kp) == "global"
    ores = _check_stack(orc, torch_gpu, kp, op, src)
    gdst, _ = gpu_pass(torch_gpu, kp, src, 0)
    assert np.array_equal(gdst, orc.aggregate(ores[:, 0], W))


@pytest.mark.parametrize("radius,ps_num,ps_range,final", [(1, 2, 4, False), (2, 3, 5, False), (1, 2, 4, True)])
def test_global_matcher_temporal(orc, torch_gpu, radius, ps_num, ps_range, final):
    """radius > 0 with a window no LDS matcher takes: the global matcher does the centre-frame
    search, then the temporal stage the predictive search, the merges and the self-inclusion."""
    print(f"global matcher temporal: 120x96 bs=8 R=48 r={radius} ps={ps_num}/{ps_range} final={final}", flush=True)
    TW = 2 * radius + 1
    W, H = 120, 96
    _, nz = synth.moving_sequence(W, H, TW, 8.0, seed=90 + radius)
    if final:
        basic = _basic_frames(orc, nz[:, None], 8.0)[:, 0]
        src = np.stack([basic, nz])[:, None]
    else:
        src = nz[None, None]
    kp, op = kp_and_op(orc, W, H, bs=8, rng=48, radius=radius, ps_num=ps_num, ps_range=ps_range, final=final,
                       sigma=device.scaled_sigma([8.0] * 3, final))
    assert device.match_plan(kp) == "global+temporal"
    _check_stack(orc, torch_gpu, kp, op, src)


@pytest.mark.parametrize("radius,ps_range", [(0, 4), (1, 3)])
def test_window_scan_left_of_first_row(orc, torch_gpu, radius, ps_range):
    """A noise-free flat top-left corner: its blocks' best candidates are exact ties at image row
    0, columns 0..7, which the raw-buffer window scan reads from groups that start left of the
    plane (global matcher; radius 1 with ps_range 3 also scans predictive windows that way)."""
    print(f"left-of-row window scan: 200x150 bs=8 R=48 r={radius} ps_range={ps_range}", flush=True)
    W, H = 200, 150
    TW = 2 * radius + 1
    _, nz = synth.moving_sequence(W, H, TW, 10.0, seed=97)
    nz[:, :24, :40] = 0.5
    kp, op = kp_and_op(orc, W, H, bs=8, rng=48, radius=radius, ps_range=ps_range,
                       sigma=device.scaled_sigma([10.0] * 3, False))
    assert device.match_plan(kp).startswith("global")
    ores = _check_stack(orc, torch_gpu, kp, op, nz[None, None])
    om, _ = orc.match(op, nz)
    assert np.all(om[1, 1, 1:, 1] == 0) and np.all(om[1, 1, 1:, 0] < 8)   # ties at row 0, columns 0..6


def test_global_matcher_centre_frame_before_temporal_stage(orc, torch_gpu):
    """264x264 at block_step 1, bm_range 57, radius 1: 66049 blocks, so the centre-frame search
    would use the 2-band LDS matcher, whose region does not fit; it runs on the global matcher,
    which must leave the self-inclusion to the temporal stage (it comes after the merges).
    Matches of every border block and of a seeded sample of the others are compared."""
    W = H = 264
    print(f"global matcher centre frame: {W}x{H} bs=1 R=57 r=1", flush=True)
    torch = torch_gpu
    _, nz = synth.moving_sequence(W, H, 3, 8.0, seed=95)
    kp, op = kp_and_op(orc, W, H, bs=1, rng=57, radius=1, sigma=device.scaled_sigma([8.0] * 3, False))
    nbx, nby = device.num_blocks(kp)
    assert nbx * nby >= 65536 and device.match_plan(kp) == "global+temporal"
    m = torch.zeros((nby, nbx, 8, 4), dtype=torch.int32, device="cuda")
    device.match(kp, torch.from_numpy(nz).cuda(), m)
    torch.cuda.synchronize()
    gm = m.cpu().numpy().reshape(nby * nbx, 8, 4)
    ids = _sample_ids(nbx, nby, 2500, seed=7)
    om, oe = orc.match_blocks(op, nz, ids)
    assert_matches_equal(gm[ids], om, oe)
    # the self-inclusion happened after the merges: the centre block is in every list
    sel = gm[ids]
    xs, ys = (ids % nbx), (ids // nbx)
    has_self = np.any((sel[..., 0] == xs[:, None]) & (sel[..., 1] == ys[:, None]) & (sel[..., 2] == 1), axis=1)
    assert np.all(has_self)


@pytest.mark.parametrize("W,H,bs,rng,metric,plan", [(3840, 2160, 7, 9, "ssd", "lds_short"),
                                                     (1920, 1080, 3, 9, "zsad", "lds_rows1"),
                                                     (1920, 1080, 8, 12, "ssd/norm", "lds_rows4"),
                                                     (3840, 2160, 8, 9, "sad", "lds_rows2")])
def test_large_plane_matchers(orc, torch_gpu, W, H, bs, rng, metric, plan):
    """Large planes pick the other LDS-staged variants (no row split, the 16 x 8-block tile);
    every border block and a seeded sample of the others are compared."""
    print(f"large-plane matcher: {W}x{H} bs={bs} R={rng} {metric} ({plan})", flush=True)
    torch = torch_gpu
    _, nz = synth.gray_frame(W, H, 10.0, seed=W + bs)
    kp, op = kp_and_op(orc, W, H, bs=bs, rng=rng, metric=metric, sigma=device.scaled_sigma([10.0] * 3, False))
    assert device.match_plan(kp) == plan
    nbx, nby = device.num_blocks(kp)
    m = torch.zeros((nby, nbx, 8, 4), dtype=torch.int32, device="cuda")
    device.match(kp, torch.from_numpy(nz[None]).cuda(), m)
    torch.cuda.synchronize()
    gm = m.cpu().numpy().reshape(nby * nbx, 8, 4)
    ids = _sample_ids(nbx, nby, 3000, seed=bs)
    om, oe = orc.match_blocks(op, nz[None], ids)
    assert_matches_equal(gm[ids], om, oe)


def _sample_ids(nbx, nby, n, seed):
    by, bx = np.mgrid[0:nby, 0:nbx]
    border = ((by < 2) | (by >= nby - 2) | (bx < 2) | (bx >= nbx - 2)).ravel()
    rs = np.random.default_rng(seed)
    extra = rs.choice(nbx * nby, min(n, nbx * nby), replace=False)
    return np.unique(np.concatenate([np.flatnonzero(border), extra])).astype(np.int32)


def test_wiener_4k_radius1_window(orc, torch_gpu):
    """The 4K Wiener workload at radius 1 (3840x2160 Gray, block_step 8, bm_range 9, ps 2/4),
    one window of three frames with the oracle's basic estimates as the ref clip."""
    W, H = 3840, 2160
    print(f"wiener 4k window: {W}x{H} r=1 bs=8 R=9", flush=True)
    _, nz = synth.moving_sequence(W, H, 3, 10.0, seed=100)
    basic = _basic_frames(orc, nz[:, None], 10.0)[:, 0]
    kp, op = kp_and_op(orc, W, H, bs=8, rng=9, radius=1, final=True, sigma=device.scaled_sigma([10.0] * 3, True))
    ores = _check_stack(orc, torch_gpu, kp, op, np.stack([basic, nz])[:, None])
    assert np.all(ores[0, 1, 1] > 0)


# Batched plan at the batched workload's plane sizes (a 3840x2160 and a 1920x1080 plane among
# smaller ones), two passes at block_step 4, bm_range 12, each plane against the oracle's
# basic-then-Wiener chain.
BATCH_PLANES = [(3840, 2160, 3840, 7.5), (1920, 1080, 1920, 15.0), (640, 360, 640, 2.5), (320, 720, 320, 19.0),
                (1280, 240, 1280, 11.0)]


_ORACLE_BATCH = {}


def _oracle_two_pass(orc, i, nz):
    """The oracle's basic-then-Wiener output of batch plane i (cached across parametrisations)."""
    if i not in _ORACLE_BATCH:
        w, h, st, sg = BATCH_PLANES[i]
        sb = device.scaled_sigma([sg] * 3, False)
        sf = device.scaled_sigma([sg] * 3, True)
        pb = orc.make_params(w, h, block_step=4, bm_range=12, sigma=tuple(sb), extractor_exp=3, threads=0)
        rb, _, _ = orc.bm3d(pb, nz[None, None, None])
        basic = orc.aggregate(rb[:, 0], w)[0]
        pf = orc.make_params(w, h, block_step=4, bm_range=12, sigma=tuple(sf), final=True, extractor_exp=3,
                             threads=0)
        rf, _, _ = orc.bm3d(pf, np.stack([basic, nz])[:, None, None])
        _ORACLE_BATCH[i] = orc.aggregate(rf[:, 0], w)[0]
    return _ORACLE_BATCH[i]


@pytest.mark.parametrize("autotune,schedule", [(True, "auto"), (False, "staged"), (False, "persistent")])
def test_batch_planes_against_oracle(orc, torch_gpu, autotune, schedule):
    torch = torch_gpu
    print(f"batch vs oracle: {len(BATCH_PLANES)} planes autotune={autotune} schedule={schedule}", flush=True)
    b = device.Batch(BATCH_PLANES, block_step=4, bm_range=12, extractor_exp=3, autotune=autotune, schedule=schedule)
    noisy, outs = [], []
    for i, (w, h, st, sg) in enumerate(BATCH_PLANES):
        _, nz = synth.gray_frame(w, h, sg, seed=200 + i)
        noisy.append(torch.from_numpy(nz).cuda())
        outs.append(torch.full((h, st), -1.0, dtype=torch.float32, device="cuda"))
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    b.run(noisy, outs, work)
    torch.cuda.synchronize()
    b.check()
    for i, (w, h, st, sg) in enumerate(BATCH_PLANES):
        want = _oracle_two_pass(orc, i, noisy[i].cpu().numpy())
        got = outs[i].cpu().numpy()
        assert np.array_equal(got, want), f"plane {i} {w}x{h}: max {np.max(np.abs(got - want))}"
        print(f"  plane {i} {w}x{h} sigma {sg}: bit-identical, tile {b.tile(i, 0)} / {b.tile(i, 1)}", flush=True)
