"""Failure and recovery of the filter layer on the GPU (src/bm3d.zig:285-293, 509-521, 674-680,
726-731; src/framecache.zig:101-107; the accumulator-cache sizing of src/bm3d.zig:898-949,
1235-1246):

* a frame source that returns no data for one frame of a window: the request fails with the
  reference's message, the cache claims it made are forgotten, and later requests (the same frame
  again, concurrent ones) succeed without hanging and equal a fresh instance's frames;
* the fused accumulator cache sized against free device memory: below the minimum, creation fails
  with the reference's message, with the numbers of the sizing rule; at the minimum it succeeds.
  VSZIPCU_FREE_MEM_CAP_MIB caps the free memory the sizing sees, so the path is reachable.
"""
import threading

import numpy as np
import pytest

from vszipcu import GRAYS, Clip, Error, core, synth

pytestmark = pytest.mark.gpu


class FlakySource:
    """Frames of a clip; requests for frame `bad` fail while `armed`."""

    def __init__(self, frames, bad):
        self.frames, self.bad, self.armed, self.failed = frames, bad, True, 0

    def __call__(self, n):
        if self.armed and n == self.bad:
            self.failed += 1
            raise RuntimeError("source unavailable")
        from vszipcu.core import Frame
        return Frame([self.frames[n]])


def _clip(frames, bad):
    src = FlakySource(frames, bad)
    H, W = frames[0].shape
    return Clip(GRAYS, W, H, len(frames), src), src


def _frames(clip, nf):
    return [clip.get_frame(n).planes for n in range(nf)]


def _get_with_deadline(clip, n, seconds=60):
    box = {}

    def run():
        try:
            box["frame"] = clip.get_frame(n)
        except Exception as e:  # noqa: BLE001
            box["error"] = e

    t = threading.Thread(target=run, daemon=True)
    t.start()
    t.join(seconds)
    assert not t.is_alive(), f"get_frame({n}) did not return within {seconds} s"
    if "error" in box:
        raise box["error"]
    return box["frame"]


@pytest.mark.parametrize("fn,kw,msg", [
    ("BM3Dv2", dict(radius=1, fast_fused=1), "BM3Dv2: process frame failed."),
    ("BM3Dv2", dict(radius=2, fast_fused=1, num_streams=1), "BM3Dv2: process frame failed."),
    ("BM3D", dict(radius=1), "BM3D: process frame failed."),
    ("BM3D", dict(radius=0), "BM3D: process frame failed."),
])
def test_frame_source_failure_then_recovery(torch_gpu, fn, kw, msg):
    print(f"recovery: {fn} {kw}", flush=True)
    W, H, nf, bad = 64, 48, 8, 4
    _, nz = synth.moving_sequence(W, H, nf, 8.0, seed=31)
    frames = [nz[i] for i in range(nf)]
    args = dict(sigma=8.0, block_step=4, bm_range=6, extractor_exp=3, **kw)
    want = _frames(getattr(core.vszipcu, fn)(Clip.from_frames(frames, GRAYS), **args), nf)

    clip, src = _clip(frames, bad)
    out = getattr(core.vszipcu, fn)(clip, **args)
    n_fail = bad - 1 if kw.get("radius", 0) else bad
    with pytest.raises(Error, match=msg.replace(".", r"\.")):
        _get_with_deadline(out, n_fail)
    assert src.failed >= 1
    src.armed = False
    # the same frame again, then every frame from concurrent threads
    assert np.array_equal(_get_with_deadline(out, n_fail).planes[0], want[n_fail][0])
    got = [None] * nf

    def work(n):
        got[n] = out.get_frame(n).planes

    ts = [threading.Thread(target=work, args=(n,), daemon=True) for n in range(nf)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(60)
        assert not t.is_alive(), "concurrent get_frame did not return"
    for n in range(nf):
        for p in range(len(want[n])):
            assert np.array_equal(got[n][p], want[n][p]), (n, p)


def test_failures_across_concurrent_requests(torch_gpu):
    """Several frames of a fused chain fail while others are requested concurrently on four
    streams; afterwards every frame equals a fresh instance's."""
    W, H, nf = 64, 48, 10
    _, nz = synth.moving_sequence(W, H, nf, 8.0, seed=33)
    frames = [nz[i] for i in range(nf)]
    args = dict(sigma=8.0, block_step=4, bm_range=6, extractor_exp=3, radius=1, fast_fused=1, num_streams=4)
    want = _frames(core.vszipcu.BM3Dv2(Clip.from_frames(frames, GRAYS), **args), nf)
    clip, src = _clip(frames, 6)
    out = core.vszipcu.BM3Dv2(clip, **args)
    results = [None] * nf

    def work(n):
        try:
            results[n] = out.get_frame(n).planes
        except Error as e:
            results[n] = e

    ts = [threading.Thread(target=work, args=(n,), daemon=True) for n in range(nf)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(60)
        assert not t.is_alive()
    for n in range(nf):
        if 4 <= n <= 8:   # windows n-2r..n+2r that hold frame 6
            assert isinstance(results[n], Error) and "BM3Dv2: process frame failed." in str(results[n])
        else:
            assert np.array_equal(results[n][0], want[n][0]), n
    src.armed = False
    for n in range(nf):
        assert np.array_equal(_get_with_deadline(out, n).planes[0], want[n][0]), n


def test_fused_cache_sizing_error(torch_gpu, monkeypatch):
    """1920x1080 Gray, radius 1: one accumulator slot is [TW=3][num, den][1080][1920] float32
    (rows padded to 64 floats), and the chain needs 2r+1 = 3 of them plus a 256 MiB margin."""
    W, H, r = 1920, 1080, 1
    clip = Clip.from_frames([np.zeros((H, W), np.float32)] * 4, GRAYS)
    S = (W + 63) // 64 * 64
    slot = (2 * r + 1) * 2 * H * S * 4
    min_slots = 2 * r + 1
    need_mib = (min_slots * slot) >> 20
    enough_mib = -(-(min_slots * slot + (256 << 20)) // (1 << 20))    # ceil
    args = dict(sigma=5.0, radius=r, fast_fused=1, num_streams=1)

    monkeypatch.setenv("VSZIPCU_FREE_MEM_CAP_MIB", str(enough_mib - 1))
    with pytest.raises(Error) as ei:
        core.vszipcu.BM3Dv2(clip, **args)
    assert str(ei.value) == (
        f"BM3Dv2: fast_fused needs at least {need_mib} MiB of accumulator cache ({min_slots} slots x {slot >> 20} "
        f"MiB) but only {enough_mib - 1} MiB of device memory is free. Use fast_fused=False (composed BM3D + "
        f"VAggregate, much less VRAM), or lower num_streams/radius.")

    monkeypatch.setenv("VSZIPCU_FREE_MEM_CAP_MIB", str(enough_mib))
    out = core.vszipcu.BM3Dv2(clip, **args)
    assert out.get_frame(0).planes[0].shape == (H, W)

    monkeypatch.setenv("VSZIPCU_FREE_MEM_CAP_MIB", "0")
    with pytest.raises(Error, match="but only 0 MiB of device memory is free"):
        core.vszipcu.BM3Dv2(clip, **args)
    # the composed path does not size an accumulator cache
    out = core.vszipcu.BM3Dv2(clip, sigma=5.0, radius=r, fast_fused=0, num_streams=1)
    assert out.num_frames == 4
