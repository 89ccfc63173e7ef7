"""CNN kernels (BF16 MFMA, FP32 accumulation) against a plain PyTorch FP32 reference on the CPU.

Inputs and weights are rounded to bf16 first and the reference convolves those exact values in
FP32, so the only differences are the kernels' summation order and, for bf16 outputs, the final
rounding. Tolerances: bf16 outputs |got - ref| <= 8e-3 |ref| + 1e-3 max|ref| (one bf16 rounding
is 2^-9 relative); fp32 outputs and gradients |got - ref| <= 1e-4 max|ref|.
"""
import numpy as np
import pytest

from vszipcu import _lib, cnn, synth

pytestmark = pytest.mark.gpu

CH = 32


def _bf16(torch, a):
    return torch.as_tensor(a, dtype=torch.float32).to(torch.bfloat16)


def _weights(torch, seed, scale=0.08):
    rng = np.random.default_rng(seed)
    return _bf16(torch, rng.normal(0, scale, (9, CH, CH)).astype(np.float32))


def _torch_w(torch, wb):  # [9][o][i] -> conv2d weight [o][i][3][3]
    return wb.float().view(3, 3, CH, CH).permute(2, 3, 0, 1).contiguous()


def _run_conv(torch, x_bf, w_bf, bias, mode, aux_bf=None, base=None, out_stride=0):
    n, h, w, _ = x_bf.shape
    L = _lib.load()
    xd = x_bf.view(torch.int16).cuda().contiguous()
    wd = w_bf.view(torch.int16).cuda().contiguous()
    bd = torch.as_tensor(bias, dtype=torch.float32).cuda() if bias is not None else None
    ad = aux_bf.view(torch.int16).cuda().contiguous() if aux_bf is not None else None
    bsd = base.cuda().contiguous() if base is not None else None
    if mode == 1:
        y = torch.zeros((n, h, out_stride), dtype=torch.float32, device="cuda")
    else:
        y = torch.zeros((n, h, w, CH), dtype=torch.int16, device="cuda")
    cnn._check(L.vsz_cnn_conv3x3(cnn._p(xd), cnn._p(wd), cnn._p(bd), n, h, w, mode, cnn._p(ad), cnn._p(bsd),
                                 cnn._p(y), out_stride, cnn._s(None)))
    torch.cuda.synchronize()
    return y.cpu() if mode == 1 else y.cpu().view(torch.bfloat16).float()


def _close(got, ref, rel, absfrac):
    scale = float(ref.abs().max())
    err = (got - ref).abs()
    bound = rel * ref.abs() + absfrac * scale
    assert bool((err <= bound).all()), f"max err {float(err.max()):.3g} (scale {scale:.3g})"


@pytest.mark.parametrize("n,h,w", [(1, 16, 16), (2, 20, 37), (1, 33, 18)])
def test_conv_forward_relu(torch_gpu, n, h, w):
    torch = torch_gpu
    rng = np.random.default_rng(n * 100 + h)
    x = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))
    wb = _weights(torch, h)
    bias = rng.normal(0, 0.1, CH).astype(np.float32)
    got = _run_conv(torch, x, wb, bias, 0)
    ref = torch.relu(torch.nn.functional.conv2d(x.float().permute(0, 3, 1, 2), _torch_w(torch, wb), torch.tensor(bias),
                                                padding=1)).permute(0, 2, 3, 1)
    _close(got, ref, 8e-3, 1e-3)


def test_conv_output_layer_fp32(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(5)
    n, h, w, stride = 2, 24, 30, 32
    x = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))
    wb = _weights(torch, 6)
    bias = rng.normal(0, 0.1, CH).astype(np.float32)
    base = torch.as_tensor(rng.random((n, h, w), dtype=np.float32))
    got = _run_conv(torch, x, wb, bias, 1, base=base, out_stride=stride)[:, :, :w]
    conv = torch.nn.functional.conv2d(x.float().permute(0, 3, 1, 2), _torch_w(torch, wb), torch.tensor(bias), padding=1)
    _close(got, base + conv[:, 0], 0.0, 1e-4)


def test_conv_backward_data_through_relu(torch_gpu):
    torch = torch_gpu
    L = _lib.load()
    rng = np.random.default_rng(7)
    n, h, w = 2, 20, 27
    dz = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))
    aux = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))   # previous activation; >0 keeps
    w32 = torch.as_tensor(rng.normal(0, 0.08, (9, CH, CH)).astype(np.float32)).cuda()
    wf = torch.zeros(9 * CH * CH, dtype=torch.int16, device="cuda")
    wbk = torch.zeros_like(wf)
    cnn._check(L.vsz_cnn_pack_weights(cnn._p(w32), CH, CH, cnn._p(wf), cnn._p(wbk), cnn._s(None)))
    torch.cuda.synchronize()
    got = _run_conv(torch, dz, wbk.cpu().view(torch.bfloat16).view(9, CH, CH), None, 2, aux_bf=aux)
    W = _torch_w(torch, wf.cpu().view(torch.bfloat16).view(9, CH, CH))
    xin = torch.zeros((n, CH, h, w), requires_grad=True)
    torch.nn.functional.conv2d(xin, W, padding=1).backward(dz.float().permute(0, 3, 1, 2))
    ref = xin.grad.permute(0, 2, 3, 1) * (aux.float() > 0)
    _close(got, ref, 8e-3, 1e-3)


def test_weight_and_bias_gradients(torch_gpu):
    torch = torch_gpu
    L = _lib.load()
    rng = np.random.default_rng(9)
    n, h, w = 3, 40, 35          # 3 x 3 x 3 = 27 tiles: several workgroups, a partial last one
    x = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))
    dz = _bf16(torch, rng.normal(0, 1, (n, h, w, CH)).astype(np.float32))
    dw = torch.zeros(9 * CH * CH, dtype=torch.float32, device="cuda")
    db = torch.zeros(CH, dtype=torch.float32, device="cuda")
    xd, dzd = x.view(torch.int16).cuda(), dz.view(torch.int16).cuda()
    cnn._check(L.vsz_cnn_conv3x3_dw(cnn._p(xd), cnn._p(dzd), n, h, w, cnn._p(dw), cnn._s(None)))
    cnn._check(L.vsz_cnn_bias_grad(cnn._p(dzd), n * h * w, cnn._p(db), cnn._s(None)))
    torch.cuda.synchronize()
    Wv = torch.zeros((CH, CH, 3, 3), requires_grad=True)
    bv = torch.zeros(CH, requires_grad=True)
    torch.nn.functional.conv2d(x.float().permute(0, 3, 1, 2), Wv, bv, padding=1).backward(dz.float().permute(0, 3, 1, 2))
    ref_w = Wv.grad.permute(2, 3, 0, 1).reshape(9, CH, CH)
    _close(dw.cpu().view(9, CH, CH), ref_w, 0.0, 1e-4)
    _close(db.cpu(), bv.grad, 0.0, 1e-4)


def test_training_reduces_the_loss(torch_gpu):
    torch = torch_gpu
    print("cnn: 120 training steps on 256 generated 64x64 pairs", flush=True)
    noisy, basic, sig = cnn.make_pairs(256, patch=64, frame=256, draws_per_frame=4, seed=3)
    m = cnn.Model.init(3)
    tr = cnn.Trainer(m, 16, 64, 64, lr=2e-3)
    losses = []
    for st in range(120):
        idx = torch.arange(16 * (st % 16), 16 * (st % 16 + 1), device="cuda")
        losses.append(float(tr.step(noisy[idx].contiguous(), basic[idx].contiguous(), sig[idx].contiguous()).item()))
    first, last = np.mean(losses[:8]), np.mean(losses[-8:])
    print(f"  loss {first:.3e} -> {last:.3e}", flush=True)
    assert np.isfinite(last) and last < 0.5 * first


def test_committed_weights_denoise(torch_gpu):
    torch = torch_gpu
    import os
    if not os.path.exists(cnn.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the package yet")
    m = cnn.Model.load()
    clean, nz = synth.gray_frame(256, 192, 12.0, seed=77)
    out = m.run(torch.from_numpy(nz).cuda(), 12.0).cpu().numpy()
    p_noisy, p_out = synth.psnr(nz, clean), synth.psnr(out, clean)
    print(f"  PSNR noisy {p_noisy:.2f} dB, learned basic {p_out:.2f} dB", flush=True)
    assert p_out > p_noisy + 3.0


def test_inference_path_matches_training_forward(torch_gpu):
    """vsz_cnn_run (first/last layers on FP32 VALU, middle layers on MFMA) against the forward pass
    of the training kernels (every layer on the padded MFMA kernel) on the same weights: the
    same bf16 operands, different summation order, so bf16 activations can round the other way
    and the outputs agree to that rounding: max-abs <= 1e-2 and mean-abs <= 1e-3 of the output
    range (the output layer is scaled up here so that it carries real weight; a layout error
    gives differences of the order of the range itself)."""
    torch = torch_gpu
    L = _lib.load()
    m = cnn.Model.init(11)
    for li in range(6):   # non-trivial biases and a real output layer
        m.b[li].copy_(torch.linspace(-0.05, 0.05, CH, device="cuda"))
        m.w[li].mul_(1.0 if li < 5 else 100.0)
    m.pack()
    n, h, w = 2, 40, 52
    clean = np.stack([synth.clean_luma(w, h, 5 + i) for i in range(n)])
    nz = torch.from_numpy(np.stack([synth.noisy(clean[i], 10.0 + 5 * i, 9 + i) for i in range(n)])).cuda()
    got = m.run(nz, [10.0, 15.0])
    tr = cnn.Trainer(m, n, h, w)
    sig = torch.tensor([10.0 / 255, 15.0 / 255], device="cuda")
    s = cnn._s(None)
    cnn._check(L.vsz_cnn_pack_input(cnn._p(nz), n, h, w, w, cnn._p(sig), cnn._p(tr.a[0]), cnn._p(tr.base), s))
    for li in range(5):
        cnn._check(L.vsz_cnn_conv3x3(cnn._p(tr.a[li]), cnn._p(m.wf[li]), cnn._p(m.b[li]), n, h, w, 0, None, None,
                                     cnn._p(tr.a[li + 1]), 0, s))
    cnn._check(L.vsz_cnn_conv3x3(cnn._p(tr.a[5]), cnn._p(m.wf[5]), cnn._p(m.b[5]), n, h, w, 1, None, cnn._p(tr.base),
                                 cnn._p(tr.out), w, s))
    torch.cuda.synchronize()
    ref = tr.out.view(n, h, w)
    d = (got - ref).abs()
    scale = float(ref.abs().max())
    print(f"  inference vs training forward: max-abs {float(d.max()):.3g}, mean-abs {float(d.mean()):.3g}, "
          f"output range {scale:.3g}", flush=True)
    assert float(d.max()) <= 1e-2 * scale and float(d.mean()) <= 1e-3 * scale
