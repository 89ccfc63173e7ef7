"""Batched uneven planes (vsz_bm3d_batch_*): one launch per stage over all planes, or the whole
batch in one persistent launch, replayed as a hipGraph; tile shapes, schedule and matcher from the
autotuner or the cost model.

With a non-zero extractor the accumulator sums are exact integers, so the batch must give the
same bits as running each plane alone through vsz_bm3d_run (basic, then Wiener with ref =
basic), whatever tiles either path picks. Graph replays are checked against eager runs, with
inputs changed in place between replays and a second pointer set forcing a new capture.
"""
import numpy as np
import pytest

from vszipcu import device, synth

# (width, height, stride, raw sigma): uneven sizes, one with a padded stride
PLANES = [(64, 48, 64, 5.0), (40, 32, 48, 12.0), (96, 40, 96, 3.5), (37, 29, 40, 18.0), (128, 72, 128, 9.0),
          (200, 120, 208, 1.5)]
KW = dict(block_step=4, bm_range=12, extractor_exp=3)
# every (width - 8) a multiple of 4: the shared-column matcher is allowed too
PLANES_SC = [(64, 48, 64, 5.0), (40, 32, 48, 12.0), (96, 40, 96, 3.5), (128, 72, 128, 9.0), (200, 120, 208, 1.5)]


def test_batch_rejects_bad_requests():
    with pytest.raises(RuntimeError, match="need 1..256 planes"):
        device.Batch([], **KW)
    with pytest.raises(RuntimeError, match="sigma must be >= FLT_EPSILON"):
        device.Batch([(64, 48, 64, 0.0)], **KW)
    with pytest.raises(RuntimeError, match="invalid geometry"):
        device.Batch([(64, 48, 60, 5.0)], **KW)          # stride < width
    with pytest.raises(RuntimeError, match="persistent schedule needs extractor_exp > 0, the ssd or sad"):
        device.Batch([(64, 48, 64, 5.0)], schedule="persistent", metric="zssd", **KW)
    with pytest.raises(RuntimeError, match="windows within the LDS budgets"):
        device.Batch([(640, 480, 640, 5.0)], schedule="persistent", block_step=8, bm_range=120, extractor_exp=3)
    with pytest.raises(RuntimeError, match="split-BF16 batches need LDS tiles"):
        device.Batch([(640, 480, 640, 5.0)], precision="split-bf16", block_step=8, bm_range=120, extractor_exp=3)
    with pytest.raises(RuntimeError, match="persistent schedule needs extractor_exp > 0, the ssd or sad"):
        device.Batch([(64, 48, 64, 5.0)], schedule="persistent", block_step=4, bm_range=12, extractor_exp=0)
    with pytest.raises(RuntimeError, match="persistent schedule needs .* the dct transform"):
        device.Batch([(64, 48, 64, 5.0)], schedule="persistent", transform="haar", **KW)
    with pytest.raises(RuntimeError, match="one transform for both stages"):
        device.Batch([(64, 48, 64, 5.0)], transform="wht", precision="split-bf16", **KW)


def _inputs(torch, planes, seed):
    out = []
    for i, (w, h, st, sg) in enumerate(planes):
        img = np.zeros((h, st), np.float32)
        img[:, :w] = synth.noisy(synth.clean_luma(w, h, seed + i), sg, seed + 100 + i)
        out.append(torch.from_numpy(img).cuda())
    return out


def _single(torch, plane, noisy, precision="fp32"):
    w, h, st, sg = plane
    kb = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, False), precision=precision, **KW)
    kf = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, True), final=True, precision=precision, **KW)
    work = torch.zeros(max(device.workspace_size(kb), device.workspace_size(kf)), dtype=torch.uint8, device="cuda")
    basic = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
    out = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
    device.run(kb, noisy[None, None], None, work, basic, 0)
    device.run(kf, basic[:, None], noisy[None, None], work, out, 0)
    torch.cuda.synchronize()
    return out[0].cpu().numpy()[:, :w]


@pytest.mark.gpu
@pytest.mark.parametrize("planes", ["uneven", "shared_cols"])
@pytest.mark.parametrize("autotune,precision,schedule", [(False, "fp32", "staged"), (True, "fp32", "staged"),
                                                         (True, "split-bf16", "staged"), (False, "fp32", "persistent"),
                                                         (True, "split-bf16", "persistent"), (True, "fp32", "auto")])
def test_batch_equals_single_plane_runs(torch_gpu, planes, autotune, precision, schedule):
    torch = torch_gpu
    P = PLANES if planes == "uneven" else PLANES_SC
    print(f"batch: {len(P)} {planes} planes autotune={autotune} precision={precision} schedule={schedule}", flush=True)
    b = device.Batch(P, autotune=autotune, precision=precision, schedule=schedule, **KW)
    noisy = _inputs(torch, P, 40)
    outs = [torch.full((h, st), -1.0, dtype=torch.float32, device="cuda") for (w, h, st, _) in P]
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    b.run(noisy, outs, work)
    torch.cuda.synchronize()
    b.check()                         # no dependency wait of a persistent launch ran out
    for i, p in enumerate(P):
        want = _single(torch, p, noisy[i], precision)
        got = outs[i].cpu().numpy()[:, :p[0]]
        assert np.array_equal(got, want), f"plane {i} {p}"
        tx, ty = b.tile(i, 0)
        assert tx * ty >= 1
    sched, matcher = b.schedule()
    if schedule != "auto":
        assert sched == schedule
    if planes == "uneven":
        assert matcher == "rows"      # a width with (W - 8) % 4 != 0 rules the shared-column matcher out
    elif not autotune:
        assert matcher == "shared-column"   # preferred where allowed


@pytest.mark.gpu
@pytest.mark.parametrize("bm_range", [48, 120])
def test_batch_beyond_lds_budgets_equals_single_plane_runs(torch_gpu, bm_range):
    """Windows whose staged matcher region exceeds the LDS budget take the global-memory matcher;
    at R 120 the 640 x 480 plane's filter region exceeds the tile budget too (direct atomics)."""
    torch = torch_gpu
    P = [(640, 480, 640, 6.0), (96, 64, 96, 9.0), (200, 120, 208, 3.0)]
    kw = dict(block_step=8, bm_range=bm_range, extractor_exp=3)
    print(f"batch beyond LDS: R={bm_range}", flush=True)
    b = device.Batch(P, autotune=True, **kw)
    noisy = _inputs(torch, P, 90)
    outs = [torch.full((h, st), -1.0, dtype=torch.float32, device="cuda") for (w, h, st, _) in P]
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    b.run(noisy, outs, work)
    torch.cuda.synchronize()
    assert b.schedule() == ("staged", "global")
    for i, (w, h, st, sg) in enumerate(P):
        kb = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, False), **kw)
        kf = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, True), final=True, **kw)
        wk = torch.zeros(max(device.workspace_size(kb), device.workspace_size(kf)), dtype=torch.uint8, device="cuda")
        basic = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
        want = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
        device.run(kb, noisy[i][None, None], None, wk, basic, 0)
        device.run(kf, basic[:, None], noisy[i][None, None], wk, want, 0)
        torch.cuda.synchronize()
        assert np.array_equal(outs[i].cpu().numpy()[:, :w], want[0].cpu().numpy()[:, :w]), f"plane {i}"


@pytest.mark.gpu
@pytest.mark.parametrize("transform", ["haar", "wht", "bior1.5"])
@pytest.mark.parametrize("exact", [True, False])
def test_batch_other_transforms_equal_single_plane_runs(torch_gpu, transform, exact):
    """A batch with one transform for both stages against each plane run alone with that transform
    (extractor 3: bit-identical; extractor 0: FP32 atomics in any order, so within rounding)."""
    torch = torch_gpu
    kw = dict(block_step=4, bm_range=12, extractor_exp=3 if exact else 0)
    print(f"batch: transform={transform} exact={exact}", flush=True)
    b = device.Batch(PLANES, autotune=True, transform=transform, **kw)
    noisy = _inputs(torch, PLANES, 70)
    outs = [torch.full((h, st), -1.0, dtype=torch.float32, device="cuda") for (w, h, st, _) in PLANES]
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    b.run(noisy, outs, work)
    torch.cuda.synchronize()
    for i, (w, h, st, sg) in enumerate(PLANES):
        kb = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, False), t2d=transform, t1d=transform, **kw)
        kf = device.kparams(w, h, st, sigma=device.scaled_sigma([sg] * 3, True), final=True, t2d=transform,
                            t1d=transform, **kw)
        wk = torch.zeros(max(device.workspace_size(kb), device.workspace_size(kf)), dtype=torch.uint8, device="cuda")
        basic = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
        want = torch.zeros((1, h, st), dtype=torch.float32, device="cuda")
        device.run(kb, noisy[i][None, None], None, wk, basic, 0)
        device.run(kf, basic[:, None], noisy[i][None, None], wk, want, 0)
        torch.cuda.synchronize()
        got, ref = outs[i].cpu().numpy()[:, :w], want[0].cpu().numpy()[:, :w]
        if exact:
            assert np.array_equal(got, ref), f"plane {i}"
        else:
            assert np.abs(got - ref).max() <= 1e-5, f"plane {i}: {np.abs(got - ref).max()}"


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["staged", "persistent"])
def test_batch_graph_replays_follow_inputs(torch_gpu, schedule):
    """Replays with inputs changed in place; for the persistent launch this also checks that every
    launch leaves its work queue empty for the next."""
    torch = torch_gpu
    print(f"batch replays: schedule={schedule}", flush=True)
    b = device.Batch(PLANES, autotune=True, schedule=schedule, **KW)
    noisy = _inputs(torch, PLANES, 60)
    outs = [torch.zeros((h, st), dtype=torch.float32, device="cuda") for (w, h, st, _) in PLANES]
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    for rep in range(3):
        # change the inputs in place between replays (eager work between the graph launches)
        fresh = _inputs(torch, PLANES, 60 + rep)
        for a, f in zip(noisy, fresh):
            a.copy_(f)
        torch.cuda.synchronize()
        b.run(noisy, outs, work)
        torch.cuda.synchronize()
        for i, p in enumerate(PLANES):
            assert np.array_equal(outs[i].cpu().numpy()[:, :p[0]], _single(torch, p, noisy[i])), f"replay {rep} plane {i}"
    # new output pointers: the plan re-uploads its tables and captures again
    outs2 = [torch.zeros_like(o) for o in outs]
    b.run(noisy, outs2, work)
    torch.cuda.synchronize()
    b.check()
    for o, o2 in zip(outs, outs2):
        assert torch.equal(o, o2)
