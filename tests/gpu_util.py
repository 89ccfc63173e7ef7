"""Shared helpers of the GPU tests: run one kernel-level pass through the C ABI on device
buffers (torch allocates them; the native kernels compute) and the matching oracle pass."""
import numpy as np

from vszipcu import device


def kp_and_op(orc, W, H, *, bs=8, rng=9, radius=0, ps_num=2, ps_range=4, chroma=False, final=False,
              sigma=(1.0, 1.0, 1.0), ext=3, metric="ssd", t2d="dct", t1d="dct", proc_mask=0b111, stride=None):
    S = stride or W
    kw = dict(block_step=bs, bm_range=rng, radius=radius, ps_num=ps_num, ps_range=ps_range, chroma=chroma,
              final=final, sigma=tuple(sigma), extractor_exp=ext, metric=metric, t2d=t2d, t1d=t1d)
    kp = device.kparams(W, H, S, proc_mask=proc_mask, **kw)
    op = orc.make_params(W, H, S, threads=0, **kw)
    return kp, op


def gpu_pass(torch, kp, src, out_kind):
    """src: [(final?2:1), P, TW, H, S] float32 numpy. Returns (out, matches [nby,nbx,8,4])."""
    P = 3 if kp.chroma else 1
    TW = 2 * kp.radius + 1
    c0 = torch.from_numpy(np.ascontiguousarray(src[0])).cuda()
    c1 = torch.from_numpy(np.ascontiguousarray(src[1])).cuda() if kp.final_pass else None
    work = torch.zeros(device.workspace_size(kp), dtype=torch.uint8, device="cuda")
    shape = (P, kp.height, kp.stride) if out_kind == 0 else (P, TW, 2, kp.height, kp.stride)
    out = torch.zeros(shape, dtype=torch.float32, device="cuda")
    device.run(kp, c0, c1, work, out, out_kind)
    torch.cuda.synchronize()
    nbx, nby = device.num_blocks(kp)
    m = work[: nbx * nby * 8 * 16].view(torch.int32).reshape(nby, nbx, 8, 4).cpu().numpy()
    return out.cpu().numpy(), m


def assert_matches_equal(gm, om, oe):
    """GPU match lists (x, y, z, error bits) against the oracle's (x, y, z) and errors."""
    assert np.array_equal(gm[..., :3], om), f"{np.argwhere(np.any(gm[..., :3] != om, axis=-1))[:5]}"
    ge = gm[..., 3].copy().view(np.float32)
    assert np.array_equal(ge, oe.astype(np.float32)), "block distances differ"
