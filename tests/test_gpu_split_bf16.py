"""The compensated split-BF16 MFMA form of the DCT stages against the FP64 oracle.

Matching is untouched (FP32 VALU), so the match lists must equal the plain path's. The
transform products carry about 2^-16 relative error (bf16 hi*hi + hi*lo + lo*hi with FP32
accumulation), which can flip hard-threshold decisions of coefficients within that margin of
sigma. Tolerances, against the FP64 oracle run on the same match lists:
  max-abs <= 5e-3 and mean-abs <= 5e-5 per output pixel (pixels in [0, 1]);
  PSNR against the clean image within 0.05 dB of the plain FP32 path.
"""
import numpy as np
import pytest

from vszipcu import device, synth

MAX_ABS = 5e-3
MEAN_ABS = 5e-5
PSNR_DB = 0.05


def _run(torch, W, H, sig, final, src, precision):
    kp = device.kparams(W, H, block_step=4, bm_range=9, sigma=tuple(sig), extractor_exp=3, final=final,
                        precision=precision)
    c0 = torch.from_numpy(np.ascontiguousarray(src[0, 0])).cuda()[None]
    c1 = torch.from_numpy(np.ascontiguousarray(src[-1, 0])).cuda()[None] if final else None
    work = torch.zeros(device.workspace_size(kp), dtype=torch.uint8, device="cuda")
    out = torch.zeros((1, H, W), dtype=torch.float32, device="cuda")
    device.run(kp, c0, c1, work, out, 0)
    torch.cuda.synchronize()
    nbx, nby = device.num_blocks(kp)
    m = work[: nbx * nby * 128].view(torch.int32).reshape(nby, nbx, 8, 4).cpu().numpy()
    return out.cpu().numpy()[0], m


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,final,seed", [(96, 64, False, 7), (128, 96, False, 8), (96, 64, True, 9)])
def test_split_bf16_against_fp64(orc, torch_gpu, W, H, final, seed):
    print(f"split-bf16: {W}x{H} final={final} seed={seed}", flush=True)
    clean, nz = synth.gray_frame(W, H, 10.0, seed=seed)
    sb = device.scaled_sigma([10.0] * 3, False)
    pb = orc.make_params(W, H, block_step=4, bm_range=9, sigma=tuple(sb), extractor_exp=3)
    res, _, _ = orc.bm3d(pb, nz[None, None, None])
    basic = orc.aggregate(res[:, 0], W)[0]
    sig = device.scaled_sigma([10.0] * 3, True) if final else sb
    src = np.stack([basic, nz])[:, None, None] if final else nz[None, None, None]
    plain, m32 = _run(torch_gpu, W, H, sig, final, src, "fp32")
    split, m16 = _run(torch_gpu, W, H, sig, final, src, "split-bf16")
    assert np.array_equal(m32, m16)
    po = orc.make_params(W, H, block_step=4, bm_range=9, sigma=tuple(sig), extractor_exp=3, final=final)
    r64, _, _ = orc.bm3d(po, src.astype(np.float64), dtype=np.float64, given_matches=np.ascontiguousarray(m32[..., :3]))
    want = orc.aggregate(r64[:, 0], W, dtype=np.float64)[0]
    err = np.abs(split.astype(np.float64) - want)
    print(f"  max-abs {err.max():.3g} mean-abs {err.mean():.3g} (plain: {np.abs(plain - want).max():.3g})", flush=True)
    assert err.max() <= MAX_ABS and err.mean() <= MEAN_ABS
    assert abs(synth.psnr(split, clean) - synth.psnr(plain, clean)) <= PSNR_DB


def test_split_bf16_needs_the_dct_pair():
    with pytest.raises(RuntimeError):
        device.num_blocks(device.kparams(64, 48, t2d="haar", precision="split-bf16"))
