"""BM3D 1080p benchmark (single GPU).

Workload: plain two-pass BM3Dv2 (basic hard-threshold pass, then Wiener with ref = basic),
chroma=True on one 1920x1080 YUV444PS frame; sigma=[5,3,3], block_step=4, bm_range=16,
ps_num=2, ps_range=4, ssd, dct/dct, extractor_exp=3, float32. Inputs are synthetic (seeded
generator) and resident in HBM before the timed region. One step = both passes (matching,
filtering, aggregation, division) over the frame. value = Mpix/s = 1920*1080 / step time.

Roofline: the dominant kernel's algorithmic work is counted from the reference geometry:
SSD_FLOP = 192 per candidate block evaluation, summed over the clipped windows of every
reference block of both passes. Its duration comes from HIP events on the launch stream.
step_roofline: the whole step's lower bound T_min = (SSD_FLOP + TF_FLOP) / P_fp32, with
TF_FLOP = 128 per 8-point transform (384 per group and plane for the hard pass, 576 for Wiener),
against the measured step time.
end_to_end: the same two passes through the filter layer on host frames (upload, kernels,
download, 4 streams), in Mpix/s; reported beside `value`, which is device-resident.
CPU baseline: the FP32 oracle (multithreaded C++) on the full frame, plus one thread on a strip.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "vapoursynth-zipcu_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "BM3D 1080p Mpix/s: plain FP32 vs learned+BF16 form at ≤0.05 dB PSNR loss"
W, H = 1920, 1080
SIGMA = [5.0, 3.0, 3.0]
PARAMS = dict(block_step=4, bm_range=16, ps_num=2, ps_range=4, chroma=True, extractor_exp=3,
              metric="ssd", t2d="dct", t1d="dct")
PEAK_FP32_TFLOPS = 157.3


def origins(n, step):
    b = np.arange(0, (n - 8 + step - 1) // step + 1)
    return np.minimum(b * step, n - 8)


def ssd_candidates(width, height, step, rng):
    """Candidate block evaluations of one spatial pass (clipped windows, src/bm3d.cu:618-660)."""
    xs, ys = origins(width, step), origins(height, step)
    nx = np.minimum(xs + rng, width - 8) - np.maximum(xs - rng, 0) + 1
    ny = np.minimum(ys + rng, height - 8) - np.maximum(ys - rng, 0) + 1
    return int(nx.sum()) * int(ny.sum())


PEAK_BF16_TFLOPS = 2500.0
# dense FLOPs per pixel of the CNN as inference runs it: 2->32 and 32->1 end layers, four 32->32
# middle layers, 3x3 taps, 2 FLOPs per multiply-add
CNN_FLOP_PER_PX = 2 * 9 * (2 * 32 + 4 * 32 * 32 + 32 * 1)


def num_blocks(width, height, step):
    return len(origins(width, step)) * len(origins(height, step))


def temporal_candidates(width, height, step, radius, ps_num, ps_range):
    """Predictive-search candidates of one temporal pass: 2r frames x ps_num windows of
    (2 ps_range + 1)^2 per reference block, counted unclipped (windows that cross the plane edge
    scan fewer; an upper bound within a few percent at 1080p)."""
    return num_blocks(width, height, step) * 2 * radius * ps_num * (2 * ps_range + 1) ** 2


def step_roofline(ssd_flop, tf_flop, dt, *, split_bf16=False, mfma_flop=0.0):
    """T_min = SSD_FLOP / P_fp32 + TF_FLOP / P_tf (+ other MFMA work / P_bf16), against the step
    time dt: P_tf = P_fp32, or P_bf16 / 3 for split-compensated BF16 transforms."""
    p_tf = PEAK_BF16_TFLOPS / 3 if split_bf16 else PEAK_FP32_TFLOPS
    t_min = ssd_flop / (PEAK_FP32_TFLOPS * 1e12) + tf_flop / (p_tf * 1e12) + mfma_flop / (PEAK_BF16_TFLOPS * 1e12)
    out = {"ssd_gflop": round(ssd_flop / 1e9, 3), "tf_gflop": round(tf_flop / 1e9, 3),
           "t_min_ms": round(t_min * 1e3, 4), "frac": round(t_min / dt, 4)}
    if mfma_flop:
        out["mfma_gflop"] = round(mfma_flop / 1e9, 3)
    return out


def _lib_cnn_ws(n, h, w):
    from vszipcu import _lib
    return _lib.load().vsz_cnn_workspace_size(n, h, w)


def _oracle_two_pass(nz, rows, threads):
    import oracle
    from vszipcu import device
    strip = np.ascontiguousarray(nz[:, :rows])
    common = {k: v for k, v in PARAMS.items()}
    pb = oracle.make_params(W, rows, sigma=tuple(device.scaled_sigma(SIGMA, False)), threads=threads, **common)
    pf = oracle.make_params(W, rows, sigma=tuple(device.scaled_sigma(SIGMA, True)), final=True, threads=threads,
                            **common)
    t0 = time.perf_counter()
    res, _, _ = oracle.bm3d(pb, strip[None, :, None])
    basic = oracle.aggregate(res[:, 0], W)
    res2, _, _ = oracle.bm3d(pf, np.stack([basic, strip])[:, :, None])
    oracle.aggregate(res2[:, 0], W)
    return time.perf_counter() - t0


def cpu_baseline(nz, strip_rows):
    threads = min(os.cpu_count() or 1, 64)
    dt = _oracle_two_pass(nz, H, threads)
    dt1 = _oracle_two_pass(nz, strip_rows, 1)
    return {"value": round(W * H / dt / 1e6, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle, both passes over the whole {W}x{H} frame ({dt:.1f} s wall, {threads} threads)",
            "single_thread": {"value": round(W * strip_rows / dt1 / 1e6, 4), "unit": "Mpix/s",
                              "sample": f"1 thread, both passes on a {W}x{strip_rows} strip ({dt1:.1f} s)"}}


def end_to_end(nz, frames=24, streams=4):
    """Both passes through the filter layer on host frames: BM3D (basic) over `frames` frames,
    then BM3D with ref = those outputs (Wiener), `streams` frames in flight on as many host
    threads (num_streams = streams). The frame source hands over existing host planes and the
    outputs land in preallocated host planes, as a VapourSynth core hands frames to getFrame:
    the time is the engine's (pinned staging copies, H2D, kernels, D2H or zero-copy output,
    copy into the caller's frame), not a Python frame server's. Returns two-pass Mpix/s."""
    import ctypes as C
    import importlib
    import threading

    from vszipcu import YUV444PS, Clip, _lib
    vcore = importlib.import_module("vszipcu.core")
    L = _lib.load()
    src = [[np.ascontiguousarray(nz[p]) for p in range(3)] for _ in range(frames)]
    clip = Clip.from_frames(src, YUV444PS)
    kw = dict(sigma=SIGMA, block_step=4, bm_range=16, ps_num=2, ps_range=4, chroma=1, extractor_exp=3,
              num_streams=streams)

    def create(ref):
        args = vcore._bm3d_args(clip, ref, kw)
        h, err = C.c_void_p(), C.create_string_buffer(512)
        if L.vsz_bm3d_create(C.byref(args), C.byref(h), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return h

    def run(h, ref_planes, outs):
        def get_plane(_user, which, n, p, stride):
            a = (src if which == 0 else ref_planes)[n][p]
            stride[0] = a.strides[0]
            return a.ctypes.data
        cb = _lib.GET_PLANE_FN(get_plane)
        todo, lock, errors = list(range(frames)), threading.Lock(), []

        def worker():
            err = C.create_string_buffer(512)
            while True:
                with lock:
                    if not todo:
                        return
                    n = todo.pop(0)
                ptrs = (C.c_void_p * 3)(*[outs[n][p].ctypes.data for p in range(3)])
                st = (C.c_ssize_t * 3)(*[outs[n][p].strides[0] for p in range(3)])
                if L.vsz_filter_get_frame(h, n, cb, None, ptrs, st, err, 512) != 0:
                    errors.append(err.value.decode())
                    return

        ts = [threading.Thread(target=worker) for _ in range(streams)]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        if errors:
            raise RuntimeError(errors[0])
        return time.perf_counter() - t0

    basic = [[np.ones((H, W), np.float32) for _ in range(3)] for _ in range(frames)]
    final = [[np.ones((H, W), np.float32) for _ in range(3)] for _ in range(frames)]
    hb, hf = create(None), create(clip)
    try:
        run(hb, None, basic)                      # warm: streams, staging, first-touch of outputs
        run(hf, basic, final)
        dt = run(hb, None, basic) + run(hf, basic, final)
    finally:
        L.vsz_filter_free(hb)
        L.vsz_filter_free(hf)
    return round(W * H * frames / dt / 1e6, 3)


def profile_target(args):
    import torch
    from vszipcu import device, synth
    if args.profile_target == "copy":
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("bench: copied 3 x 256 MiB", flush=True)
        return
    _, nz = synth.yuv444_frame(W, H, SIGMA, seed=2)
    kp = device.kparams(W, H, sigma=device.scaled_sigma(SIGMA, False), **PARAMS)
    img = torch.from_numpy(nz[0][None]).cuda()
    nbx, nby = device.num_blocks(kp)
    m = torch.zeros((nby, nbx, 8, 4), dtype=torch.int32, device="cuda")
    for _ in range(args.steps):
        device.match(kp, img, m)
    torch.cuda.synchronize()
    print(f"bench: {args.steps} matching launches", flush=True)


def bench_batch(args):
    """64 independent Gray planes of seeded uneven sizes and sigmas, two-pass BM3D (bs 4, R 12,
    extractor 3) as one batched plan replayed as one hipGraph: filter tiles, schedule (one
    persistent launch, or one launch per stage) and matcher from the autotuner. Both schedules
    are also timed on their own, and the planes run one after another."""
    import torch
    from vszipcu import device, synth
    rng = np.random.default_rng(4)
    sizes = [(int(rng.choice([320, 640, 1280, 1920, 3840])), int(rng.choice([240, 360, 720, 1080, 2160])))
             for _ in range(64)]
    sigmas = [float(rng.uniform(1.0, 20.0)) for _ in range(64)]
    print(f"bench: batch of 64 planes, {sum(w * h for w, h in sizes) / 1e6:.1f} Mpix", file=sys.stderr, flush=True)
    clean = {}
    for (w, h) in sorted(set(sizes)):
        clean[(w, h)] = torch.from_numpy(synth.clean_luma(w, h, w * 7 + h)).cuda()
    gen = torch.Generator(device="cuda").manual_seed(5)
    noisy = [clean[sz] + torch.randn(sz[1], sz[0], generator=gen, device="cuda") * (sg / 255.0)
             for sz, sg in zip(sizes, sigmas)]
    outs = [torch.zeros_like(t) for t in noisy]
    planes = [(w, h, w, sg) for (w, h), sg in zip(sizes, sigmas)]
    t0 = time.perf_counter()
    b = device.Batch(planes, block_step=4, bm_range=12, extractor_exp=3, autotune=True)
    work = torch.zeros(b.workspace_size, dtype=torch.uint8, device="cuda")
    b.run(noisy, outs, work)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    for _ in range(args.warmup):
        b.run(noisy, outs, work)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run(noisy, outs, work)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    px = sum(w * h for w, h in sizes)
    # the same planes one after another through the single-plane entry point
    kw = dict(block_step=4, bm_range=12, extractor_exp=3)
    seq = []
    for (w, h), sg in zip(sizes, sigmas):
        kb = device.kparams(w, h, sigma=device.scaled_sigma([sg] * 3, False), **kw)
        kf = device.kparams(w, h, sigma=device.scaled_sigma([sg] * 3, True), final=True, **kw)
        seq.append((kb, kf))
    wmax = max(device.workspace_size(kf) for _, kf in seq)
    w1 = torch.zeros(wmax, dtype=torch.uint8, device="cuda")
    basic = torch.zeros(max(w * h for w, h in sizes), dtype=torch.float32, device="cuda")

    def sequential():
        for (kb, kf), t, o in zip(seq, noisy, outs):
            bv = basic[: kb.height * kb.width].view(1, kb.height, kb.width)
            device.run(kb, t[None, None], None, w1, bv, 0)
            device.run(kf, bv[:, None], t[None, None], w1, o[None], 0)

    sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        sequential()
    torch.cuda.synchronize()
    dt_seq = (time.perf_counter() - t0) / max(1, args.steps // 2)
    tiles = sorted({b.tile(i, 0) for i in range(64)} | {b.tile(i, 1) for i in range(64)})
    chosen, matcher = b.schedule()
    sched_ms = {}
    for sched in ("staged", "persistent"):
        bs_ = device.Batch(planes, block_step=4, bm_range=12, extractor_exp=3, autotune=True, schedule=sched)
        ws_ = torch.zeros(bs_.workspace_size, dtype=torch.uint8, device="cuda")
        os_ = [torch.zeros_like(t) for t in noisy]
        for _ in range(1 + args.warmup):
            bs_.run(noisy, os_, ws_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bs_.run(noisy, os_, ws_)
        torch.cuda.synchronize()
        sched_ms[sched] = round((time.perf_counter() - t0) / args.steps * 1e3, 3)
        del bs_, ws_, os_
    ssd = sum(192.0 * 2 * ssd_candidates(w, h, 4, 12) for w, h in sizes)
    tf = sum(128.0 * num_blocks(w, h, 4) * (384 + 576) for w, h in sizes)
    return {"metric": "BM3D batched uneven planes Mpix/s", "value": round(px / dt / 1e6, 3),
                      "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "dtype": "f32",
                      "data": "synthetic",
                      "config": {"workload": "64 Gray planes, sizes {320..3840}x{240..2160} (seed 4), sigma U[1,20], "
                                             "two-pass, bs 4, R 12, extractor 3, one hipGraph per run",
                                 "mpix": round(px / 1e6, 2), "schedule": chosen,
                                 "matcher": matcher,
                                 "staged_ms": sched_ms["staged"], "persistent_ms": sched_ms["persistent"],
                                 "sequential_ms": round(dt_seq * 1e3, 3),
                                 "sequential_mpix_s": round(px / dt_seq / 1e6, 3),
                                 "setup_s": round(t_setup, 2), "autotuned_tiles": tiles},
            "step_roofline": step_roofline(ssd, tf, dt)}


_WINDOW_IDX = {}


def _window_index(nf, radius, m):
    """Device tensor of the clamped window m-r .. m+r (built once, outside the timed loops)."""
    import torch
    key = (nf, radius, m)
    if key not in _WINDOW_IDX:
        _WINDOW_IDX[key] = torch.tensor([min(max(m - radius + t, 0), nf - 1) for t in range(2 * radius + 1)],
                                        device="cuda")
    return _WINDOW_IDX[key]


def _fused_chain(kp, frames, ref, work, acc, out, radius, stream):
    """BM3Dv2 fused chain on device-resident frames [F, H, W]: per centre frame m, the window
    clamp(m-r .. m+r) is gathered device-to-device (of `ref` too in a final pass) and one pass
    writes that centre's [TW][2][H][W] accumulator stack; then VAggregate combines the 2r+1
    stacks around each output frame (src/bm3d.zig:481-633, src/bm3d_vagg.cu)."""
    import torch
    from vszipcu import device
    nf, TW = frames.shape[0], 2 * radius + 1
    win0 = torch.empty((1, TW) + tuple(frames.shape[1:]), dtype=torch.float32, device="cuda")
    win1 = torch.empty_like(win0) if ref is not None else None
    for m in range(nf):
        idx = _window_index(nf, radius, m)
        if ref is not None:   # final pass: clip 0 = ref (basic estimate), clip 1 = noisy
            torch.index_select(ref, 0, idx, out=win0[0])
            torch.index_select(frames, 0, idx, out=win1[0])
        else:
            torch.index_select(frames, 0, idx, out=win0[0])
        device.run(kp, win0, win1, work, acc[m], 1, stream)
    for n in range(nf):
        slots = [acc[min(max(n - radius + i, 0), nf - 1)][0] for i in range(TW)]
        device.vaggregate(frames.shape[2], radius, n, nf, slots, out[n], stream)


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_temporal(args):
    """16 frames of 1920x1080 Gray with motion, sigma 8: BM3Dv2 basic pass at radius 2 (block_step
    8, bm_range 9, ps 2/4, extractor 3) through the fused on-device accumulator chain, as plain FP32
    and with split-BF16 DCT stages. One step = all 16 output frames."""
    import torch
    from vszipcu import device, synth
    nf, r, sig = 16, 2, 8.0
    clean, nz = synth.moving_sequence(W, H, nf, sig, seed=7)
    frames = torch.from_numpy(nz).cuda()
    kw = dict(radius=r, block_step=8, bm_range=9, ps_num=2, ps_range=4, extractor_exp=3)
    acc = torch.empty((nf, 1, 2 * r + 1, 2, H, W), dtype=torch.float32, device="cuda")
    out = torch.empty((nf, H, W), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    res = {}
    for name, prec in (("plain_fp32", "fp32"), ("split_bf16_dct", "split-bf16")):
        kp = device.kparams(W, H, sigma=device.scaled_sigma([sig] * 3, False), precision=prec, **kw)
        work = torch.zeros(device.workspace_size(kp), dtype=torch.uint8, device="cuda")
        dt = _timed(lambda: _fused_chain(kp, frames, None, work, acc, out, r, stream), args.steps, args.warmup)
        got = out.cpu().numpy()
        psnr = float(np.mean([synth.psnr(got[i], clean[i]) for i in range(nf)]))
        ssd = nf * 192.0 * (ssd_candidates(W, H, 8, 9) + temporal_candidates(W, H, 8, r, 2, 4))
        tf = nf * 128.0 * 384 * num_blocks(W, H, 8)
        res[name] = {"value": round(nf * W * H / dt / 1e6, 3), "ms_per_step": round(dt * 1e3, 3),
                     "psnr_db": round(psnr, 4), "step_roofline": step_roofline(ssd, tf, dt, split_bf16=prec != "fp32")}
        print(f"bench: temporal {name} {dt * 1e3:.2f} ms / 16 frames, {nf * W * H / dt / 1e6:.1f} Mpix/s, "
              f"PSNR {psnr:.3f} dB", file=sys.stderr, flush=True)
    res["split_bf16_dct"]["psnr_delta_db"] = round(res["split_bf16_dct"]["psnr_db"] - res["plain_fp32"]["psnr_db"], 4)
    return {"metric": "BM3Dv2 temporal (radius 2) 1080p Gray Mpix/s", "value": res["plain_fp32"]["value"],
                      "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": res["plain_fp32"]["ms_per_step"], "higher_is_better": True, "dtype": "f32",
                      "data": "synthetic",
                      "config": {"workload": "16 frames 1920x1080 Gray with motion (seed 7), sigma 8, basic pass, "
                                             "radius 2, bs 8, R 9, ps 2/4, extractor 3, fused accumulators + "
                                             "VAggregate, noisy input " f"PSNR {np.mean([synth.psnr(nz[i], clean[i]) for i in range(nf)]):.2f} dB"},
            "forms": res}


def bench_learned4k(args):
    """8 frames of 3840x2160 Gray with motion, sigma 10: the Wiener pass at radius 1 (fused chain,
    block_step 8, bm_range 9, extractor 3) with its reference from (a) the plain basic pass at
    radius 1 or (b) the learned CNN basic estimate (BF16 MFMA, one call per frame). One step = all
    8 output frames of both passes."""
    import torch
    from vszipcu import cnn, device, synth
    W4, H4, nf, r, sig = 3840, 2160, 8, 1, 10.0
    clean, nz = synth.moving_sequence(W4, H4, nf, sig, seed=8)
    frames = torch.from_numpy(nz).cuda()
    kw = dict(radius=r, block_step=8, bm_range=9, ps_num=2, ps_range=4, extractor_exp=3)
    kb = device.kparams(W4, H4, sigma=device.scaled_sigma([sig] * 3, False), **kw)
    kf = device.kparams(W4, H4, sigma=device.scaled_sigma([sig] * 3, True), final=True, **kw)
    work = torch.zeros(max(device.workspace_size(kb), device.workspace_size(kf)), dtype=torch.uint8, device="cuda")
    acc = torch.empty((nf, 1, 2 * r + 1, 2, H4, W4), dtype=torch.float32, device="cuda")
    basic = torch.empty((nf, H4, W4), dtype=torch.float32, device="cuda")
    out = torch.empty_like(basic)
    stream = torch.cuda.current_stream()
    model = cnn.Model.load()
    cm = model.cmodel()
    cwork = torch.empty(_lib_cnn_ws(1, H4, W4), dtype=torch.uint8, device="cuda")
    sig_map = torch.full((1,), sig / 255.0, dtype=torch.float32, device="cuda")

    def plain():
        _fused_chain(kb, frames, None, work, acc, basic, r, stream)
        _fused_chain(kf, frames, basic, work, acc, out, r, stream)

    def learned():
        for i in range(nf):   # one CNN call per frame, noise map already on the device
            cnn._check(cnn._lib.load().vsz_cnn_run(cnn.C.byref(cm), cnn._p(frames[i]), 1, H4, W4, W4, cnn._p(sig_map),
                                                   cnn._p(basic[i]), cnn._p(cwork), cnn._s(stream)))
        _fused_chain(kf, frames, basic, work, acc, out, r, stream)

    res = {}
    for name, fn in (("plain", plain), ("learned_cnn_basic", learned)):
        dt = _timed(fn, args.steps, args.warmup)
        got = out.cpu().numpy()
        psnr = float(np.mean([synth.psnr(got[i], clean[i]) for i in range(nf)]))
        ssd_pass = nf * 192.0 * (ssd_candidates(W4, H4, 8, 9) + temporal_candidates(W4, H4, 8, r, 2, 4))
        nb = num_blocks(W4, H4, 8)
        if name == "plain":
            roof = step_roofline(2 * ssd_pass, nf * 128.0 * nb * (384 + 576), dt)
        else:
            roof = step_roofline(ssd_pass, nf * 128.0 * nb * 576, dt, mfma_flop=nf * CNN_FLOP_PER_PX * W4 * H4)
        res[name] = {"value": round(nf * W4 * H4 / dt / 1e6, 3), "ms_per_step": round(dt * 1e3, 3),
                     "psnr_db": round(psnr, 4), "step_roofline": roof}
        print(f"bench: 4K {name} {dt * 1e3:.2f} ms / {nf} frames, {nf * W4 * H4 / dt / 1e6:.1f} Mpix/s, "
              f"PSNR {psnr:.3f} dB", file=sys.stderr, flush=True)
    res["learned_cnn_basic"]["psnr_delta_db"] = round(res["learned_cnn_basic"]["psnr_db"] - res["plain"]["psnr_db"], 4)
    return {"metric": "BM3Dv2 4K Gray Wiener radius 1 Mpix/s (plain vs learned basic pass)",
                      "value": res["plain"]["value"], "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": res["plain"]["ms_per_step"], "higher_is_better": True,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "8 frames 3840x2160 Gray with motion (seed 8), sigma 10, basic + Wiener "
                                             "at radius 1, bs 8, R 9, ps 2/4, extractor 3, fused accumulators + "
                                             "VAggregate; learned form: CNN basic estimate per frame"},
            "forms": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-strip-rows", type=int, default=136, help="rows of the single-thread CPU sample")
    ap.add_argument("--no-end-to-end", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the temporal16 / batch64 / learned4k lines of the default run")
    ap.add_argument("--workload", choices=["frame1080p", "batch64", "temporal16", "learned4k"], default="frame1080p",
                    help="batch64: 64 uneven Gray planes in one batched plan; temporal16: 16 frames 1080p Gray at "
                         "radius 2 through the fused chain; learned4k: 4K Gray Wiener at radius 1 after the plain or "
                         "the learned basic pass (none of them is the headline line)")
    ap.add_argument("--profile-target", choices=["step", "match", "copy"], default="step",
                    help="match: only the basic-pass matching of the bench frame, --steps launches (for "
                         "rocprofv3 --pmc runs); copy: a 256 MiB device copy, 3 times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the BM3D path would shard by frame, plane or row band)", file=sys.stderr)
        sys.exit(2)

    import torch
    from vszipcu import device, synth

    assert torch.cuda.is_available(), "bench needs a GPU"
    if args.profile_target != "step":
        return profile_target(args)
    side = {"batch64": bench_batch, "temporal16": bench_temporal, "learned4k": bench_learned4k}
    if args.workload in side:
        print(json.dumps(side[args.workload](args)), flush=True)
        return
    print(f"bench: workload 1920x1080 YUV444 two-pass BM3D, steps={args.steps} warmup={args.warmup}", file=sys.stderr,
          flush=True)
    clean, nz = synth.yuv444_frame(W, H, SIGMA, seed=2)
    kp_b = device.kparams(W, H, sigma=device.scaled_sigma(SIGMA, False), **PARAMS)
    kp_f = device.kparams(W, H, sigma=device.scaled_sigma(SIGMA, True), final=True, **PARAMS)
    d_noisy = torch.from_numpy(nz[:, None]).cuda()            # [3, 1, H, W]
    d_basic = torch.zeros((3, 1, H, W), dtype=torch.float32, device="cuda")
    d_final = torch.zeros((3, H, W), dtype=torch.float32, device="cuda")
    work_b = torch.zeros(device.workspace_size(kp_b), dtype=torch.uint8, device="cuda")
    work_f = torch.zeros(device.workspace_size(kp_f), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        device.run(kp_b, d_noisy, None, work_b, d_basic.view(3, H, W), 0, stream)
        device.run(kp_f, d_basic, d_noisy, work_f, d_final, 0, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    mpix = W * H / dt / 1e6
    print(f"bench: {dt * 1e3:.3f} ms/step, {mpix:.2f} Mpix/s", file=sys.stderr, flush=True)
    t_basic = _timed(lambda: device.run(kp_b, d_noisy, None, work_b, d_basic.view(3, H, W), 0, stream),
                     args.steps, args.warmup)

    # per-kernel timing with HIP events on the launch stream: the matching kernels of both
    # passes (re-launched alone on the same inputs) and the remainder of each pass
    nbx, nby = device.num_blocks(kp_b)
    mb = work_b[: nbx * nby * 128].view(torch.int32).view(nby, nbx, 8, 4)
    mf = work_f[: nbx * nby * 128].view(torch.int32).view(nby, nbx, 8, 4)
    reps = max(3, args.steps // 2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    for _ in range(reps):
        device.match(kp_b, d_noisy[0], mb, stream)
    ev[1].record(stream)
    for _ in range(reps):
        device.match(kp_f, d_basic[0], mf, stream)
    ev[2].record(stream)
    torch.cuda.synchronize()
    t_match_b = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    t_match_f = ev[1].elapsed_time(ev[2]) / reps * 1e-3
    cand = ssd_candidates(W, H, PARAMS["block_step"], PARAMS["bm_range"])
    flop = 192.0 * cand * 2
    achieved = flop / (t_match_b + t_match_f) / 1e12
    t_match = t_match_b + t_match_f
    print(f"bench: match kernels {t_match_b * 1e3:.3f} + {t_match_f * 1e3:.3f} ms, "
          f"{cand} candidates/pass, {achieved:.2f} TFLOP/s", file=sys.stderr, flush=True)

    traffic = None
    tpath = os.path.join(REPO, "profiles", "match_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    out = d_final.cpu().numpy()
    psnr_y = synth.psnr(out[0], clean[0])
    psnr_planes = [synth.psnr(out[p], clean[p]) for p in range(3)]

    # the same two passes with the DCT stages as compensated split-BF16 MFMA products
    kp_b16 = device.kparams(W, H, sigma=device.scaled_sigma(SIGMA, False), precision="split-bf16", **PARAMS)
    kp_f16 = device.kparams(W, H, sigma=device.scaled_sigma(SIGMA, True), final=True, precision="split-bf16",
                            **PARAMS)

    def step16():
        device.run(kp_b16, d_noisy, None, work_b, d_basic.view(3, H, W), 0, stream)
        device.run(kp_f16, d_basic, d_noisy, work_f, d_final, 0, stream)

    for _ in range(args.warmup):
        step16()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step16()
    torch.cuda.synchronize()
    dt16 = (time.perf_counter() - t0) / args.steps
    psnr16 = synth.psnr(d_final.cpu().numpy()[0], clean[0])
    print(f"bench: split-bf16 transforms {dt16 * 1e3:.3f} ms/step, {W * H / dt16 / 1e6:.2f} Mpix/s, "
          f"PSNR {psnr16:.4f} dB (plain {psnr_y:.4f})", file=sys.stderr, flush=True)

    # learned + BF16 form: the CNN basic estimate (BF16 MFMA convolutions, one call for the three
    # planes with their own noise levels), then the Wiener pass with split-BF16 DCT stages
    learned = None
    from vszipcu import cnn
    if os.path.exists(cnn.DEFAULT_WEIGHTS):
        model = cnn.Model.load()
        sig_map = torch.tensor([s / 255.0 for s in SIGMA], dtype=torch.float32, device="cuda")
        cwork = torch.empty(_lib_cnn_ws(3, H, W), dtype=torch.uint8, device="cuda")
        cm = model.cmodel()

        def step_learned(kpf):
            cnn._check(cnn._lib.load().vsz_cnn_run(cnn.C.byref(cm), cnn._p(d_noisy), 3, H, W, W, cnn._p(sig_map),
                                                   cnn._p(d_basic), cnn._p(cwork), cnn._s(stream)))
            device.run(kpf, d_basic, d_noisy, work_f, d_final, 0, stream)

        learned = {}
        for name, kpf in (("learned_bf16", kp_f16), ("learned_fp32_wiener", kp_f)):
            for _ in range(args.warmup):
                step_learned(kpf)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step_learned(kpf)
            torch.cuda.synchronize()
            dtl = (time.perf_counter() - t0) / args.steps
            fl = d_final.cpu().numpy()
            pl = synth.psnr(fl[0], clean[0])
            deltas = [round(synth.psnr(fl[p], clean[p]) - psnr_planes[p], 4) for p in range(3)]
            learned[name] = {"value": round(W * H / dtl / 1e6, 3), "ms_per_step": round(dtl * 1e3, 4),
                             "psnr_y_db": round(pl, 4), "psnr_delta_db": round(pl - psnr_y, 4),
                             "psnr_delta_db_yuv": deltas}
            print(f"bench: {name} {dtl * 1e3:.3f} ms/step, {W * H / dtl / 1e6:.2f} Mpix/s, PSNR {pl:.4f} dB "
                  f"(delta {pl - psnr_y:+.4f})", file=sys.stderr, flush=True)
        # the CNN alone, per step (3 planes)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            cnn._check(cnn._lib.load().vsz_cnn_run(cnn.C.byref(cm), cnn._p(d_noisy), 3, H, W, W, cnn._p(sig_map),
                                                   cnn._p(d_basic), cnn._p(cwork), cnn._s(stream)))
        ev1.record(stream)
        torch.cuda.synchronize()
        t_cnn = ev0.elapsed_time(ev1) / args.steps * 1e-3
        cnn_flop = float(CNN_FLOP_PER_PX) * W * H * 3   # the FLOPs inference executes (end layers unpadded)
        learned["cnn_ms"] = round(t_cnn * 1e3, 4)
        learned["cnn_tflops"] = round(cnn_flop / t_cnn / 1e12, 2)
        learned["weights"] = {k: model.meta.get(k) for k in ("pairs", "patch", "epochs", "steps", "train_s",
                                                             "chroma_frac", "log_sigma")}
        # the judge: the learned basic pass runs only if every plane's (kind, sigma) has a measured
        # PSNR delta within the bound; otherwise the plain basic pass is the fallback for the frame
        table = cnn.load_judge()
        kinds = ["luma", "chroma", "chroma"]
        allowed = [cnn.learned_allowed(SIGMA[p], kinds[p], table) for p in range(3)]
        use_learned = all(allowed)
        chosen = learned["learned_bf16"] if use_learned else {"value": round(mpix, 3), "ms_per_step": round(dt * 1e3, 4),
                                                              "psnr_delta_db_yuv": [0.0, 0.0, 0.0]}
        learned["judged"] = {"basic_pass": "learned" if use_learned else "plain (fallback)",
                             "planes_allowed": allowed, "bound_db": cnn.JUDGE_BOUND_DB, **chosen}
        # what even a free basic pass could give: the Wiener pass alone bounds the learned form
        t_wiener = dt - t_basic
        learned["speedup_bound"] = {"wiener_pass_ms": round(t_wiener * 1e3, 4),
                                    "max_speedup_vs_plain": round(dt / t_wiener, 3)}
        print(f"bench: judge -> {learned['judged']['basic_pass']} (planes allowed {allowed}); a free basic pass "
              f"would give at most {dt / t_wiener:.2f}x", file=sys.stderr, flush=True)

    nbx, nby = device.num_blocks(kp_b)
    tf_flop = 128.0 * nbx * nby * 3 * (384 + 576)
    t_min = (flop + tf_flop) / (PEAK_FP32_TFLOPS * 1e12)
    e2e = e2e8 = None
    if not args.no_end_to_end:
        print("bench: end-to-end through the filter layer ...", file=sys.stderr, flush=True)
        e2e = end_to_end(nz)
        e2e8 = end_to_end(nz, streams=8)
        print(f"bench: end-to-end {e2e} Mpix/s (4 streams), {e2e8} Mpix/s (8 streams)", file=sys.stderr, flush=True)

    result = {
        "metric": METRIC,
        "value": round(mpix, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "BM3Dv2 two-pass (basic + Wiener ref=basic), chroma, 1920x1080 YUV444PS",
                   "sigma": SIGMA, "path": "plain FP32", **{k: v for k, v in PARAMS.items()},
                   "psnr_y_db": round(psnr_y, 3), "psnr_yuv_db": [round(v, 3) for v in psnr_planes],
                   "match_ms_per_step": round(t_match * 1e3, 4), "basic_pass_ms": round(t_basic * 1e3, 4),
                   "matcher": device.match_plan(kp_b),
                   "filter_tiles": {"basic": list(device.filter_plan(kp_b)), "wiener": list(device.filter_plan(kp_f))}},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": "k_match_lds4 (SSD block matching, both passes)",
                     "work": "dense-equivalent SSD FLOP, 192 per candidate block; the kernel executes about half "
                             "(each lane takes half of its distance tree from its neighbour) and issues VALU work "
                             "on every SIMD cycle (profiles/k_match_lds4_1080p_valu_counters.md)"},
        "step_roofline": {"ssd_gflop": round(flop / 1e9, 3), "tf_gflop": round(tf_flop / 1e9, 3),
                          "t_min_ms": round(t_min * 1e3, 4), "frac": round(t_min / dt, 4)},
        "end_to_end_mpix_s": e2e,
        "end_to_end_8_streams_mpix_s": e2e8,
        "forms": {"learned": learned, "split_bf16_dct": {"value": round(W * H / dt16 / 1e6, 3), "ms_per_step": round(dt16 * 1e3, 4),
                                     "psnr_y_db": round(psnr16, 4), "psnr_delta_db": round(psnr16 - psnr_y, 4),
                                     "note": "DCT stages as compensated split-BF16 MFMA, FP32 accumulation; "
                                             "matching FP32"}},
    }
    if not args.no_workloads:
        # the temporal, batched and 4K workloads, fewer steps each; not the headline `value`
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = max(2, min(args.steps, 8)), max(1, min(args.warmup, 2))
        result["workloads"] = {}
        for name, fn in (("temporal16", bench_temporal), ("batch64", bench_batch), ("learned4k", bench_learned4k)):
            print(f"bench: workload {name} ...", file=sys.stderr, flush=True)
            r = fn(sub)
            result["workloads"][name] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "config")
                                         if k in r}
            for k in ("forms", "step_roofline"):
                if k in r:
                    result["workloads"][name][k] = r[k]
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(nz, args.cpu_strip_rows)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
