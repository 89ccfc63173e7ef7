"""vsz_bm3d_run_noalloc (the pass without a workspace argument, on the library's static device
scratch) against vsz_bm3d_run with a caller workspace, through the C ABI: bit-identical outputs on
two calls in a row, the second smaller than the first (the scratch keeps the first call's
contents), and a clear rejection of a pass whose workspace exceeds the scratch.
"""
import ctypes as C

import numpy as np
import pytest

from vszipcu import _lib, device, synth

pytestmark = pytest.mark.gpu

VSZ_EINVAL = -1


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _both(torch, kp, src, out_kind):
    """(output of vsz_bm3d_run with a workspace, output of vsz_bm3d_run_noalloc)."""
    L = _lib.load()
    P = 3 if kp.chroma else 1
    TW = 2 * kp.radius + 1
    c0 = torch.from_numpy(np.ascontiguousarray(src[0])).cuda()
    c1 = torch.from_numpy(np.ascontiguousarray(src[1])).cuda() if kp.final_pass else None
    shape = (P, kp.height, kp.stride) if out_kind == 0 else (P, TW, 2, kp.height, kp.stride)
    outs = []
    for noalloc in (False, True):
        out = torch.full(shape, -7.0, dtype=torch.float32, device="cuda")
        c1p = C.c_void_p(c1.data_ptr() if c1 is not None else 0)
        if noalloc:
            rc = L.vsz_bm3d_run_noalloc(C.byref(kp), C.c_void_p(c0.data_ptr()), c1p, C.c_void_p(out.data_ptr()),
                                        out_kind, _stream(torch))
        else:
            work = torch.zeros(device.workspace_size(kp), dtype=torch.uint8, device="cuda")
            rc = L.vsz_bm3d_run(C.byref(kp), C.c_void_p(c0.data_ptr()), c1p, C.c_void_p(work.data_ptr()),
                                C.c_void_p(out.data_ptr()), out_kind, _stream(torch))
        assert rc == 0, _lib.last_error()
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    return outs


def test_noalloc_pass_equals_workspace_pass(torch_gpu):
    torch = torch_gpu
    # call 1: the 1920x1080 YUV444 chroma basic pass (block_step 4, bm_range 16), divided output
    W, H = 1920, 1080
    print(f"noalloc: {W}x{H} YUV444 chroma basic pass", flush=True)
    _, nz = synth.yuv444_frame(W, H, (5.0, 3.0, 3.0), seed=2)
    kp = device.kparams(W, H, block_step=4, bm_range=16, chroma=True, extractor_exp=3,
                        sigma=device.scaled_sigma([5.0, 3.0, 3.0], False))
    assert device.workspace_size(kp) <= 128 << 20
    want, got = _both(torch, kp, nz[None, :, None], 0)
    assert np.array_equal(want, got)
    # call 2, smaller: a 200x150 Gray Wiener pass at radius 1, accumulator-stack output
    W, H = 200, 150
    print(f"noalloc: {W}x{H} Gray Wiener pass radius 1", flush=True)
    _, seq = synth.moving_sequence(W, H, 3, 10.0, seed=5)
    basic = seq * np.float32(0.5) + np.float32(0.25)
    kp2 = device.kparams(W, H, block_step=4, bm_range=9, radius=1, final=True, extractor_exp=3,
                         sigma=device.scaled_sigma([10.0] * 3, True))
    want2, got2 = _both(torch, kp2, np.stack([basic, seq])[:, None], 1)
    assert np.array_equal(want2, got2)
    assert np.all(got2[0, 1, 1] > 0)


def test_noalloc_rejects_a_pass_beyond_the_scratch(torch_gpu):
    L = _lib.load()
    kp = device.kparams(3840, 2160, block_step=4, bm_range=16, chroma=True, extractor_exp=3,
                        sigma=device.scaled_sigma([5.0, 3.0, 3.0], False))
    assert device.workspace_size(kp) > 128 << 20
    rc = L.vsz_bm3d_run_noalloc(C.byref(kp), C.c_void_p(16), None, C.c_void_p(16), 0, None)
    assert rc == VSZ_EINVAL
    assert "static scratch" in _lib.last_error()
