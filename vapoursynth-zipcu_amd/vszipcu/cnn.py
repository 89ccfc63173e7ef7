"""Learned basic pass: a residual CNN that stands in for BM3D's hard-threshold pass.

Six 3x3 convolutions (2 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU between them); the
inputs are the noisy plane and a constant noise-level map sigma/255, the output is
noisy + f(noisy, sigma). It is trained to reproduce the engine's own BM3D basic estimate, on
pairs the engine generates from the synthetic generator, and its output is the reference of the
unchanged Wiener pass. Every convolution, gradient, loss and optimizer step runs in the
library's HIP kernels (BF16 weights and activations, FP32 accumulation and FP32 master
weights); torch only allocates device memory and provides the stream.

    m = Model.load(DEFAULT_WEIGHTS)              # or Model.init(seed) + train(...)
    out = m.run(noisy, sigma255)                 # noisy: [n, h, stride] f32 device tensor
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from typing import Optional, Sequence

import numpy as np

from . import _lib, device, synth

CH = 32
LAYERS = [(2, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 1)]   # (cin, cout)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "cnn_basic.npz")


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"vszipcu: {_lib.last_error()} (code {rc})")


def _s(stream):
    import torch
    return C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


class Model:
    """FP32 master parameters (zero-padded to [9][32][32] / [32]) plus their BF16 packings."""

    def __init__(self, w32: Sequence, b32: Sequence):
        import torch
        self.w = [torch.as_tensor(w, dtype=torch.float32).cuda().contiguous() for w in w32]   # [9, 32, 32]
        self.b = [torch.as_tensor(b, dtype=torch.float32).cuda().contiguous() for b in b32]   # [32]
        self.wf = [torch.zeros(9 * CH * CH, dtype=torch.int16, device="cuda") for _ in LAYERS]
        self.wb = [torch.zeros(9 * CH * CH, dtype=torch.int16, device="cuda") for _ in LAYERS]
        self.meta = {}
        self.pack()

    @classmethod
    def init(cls, seed: int = 0) -> "Model":
        rng = np.random.default_rng(seed)
        ws, bs = [], []
        for li, (cin, cout) in enumerate(LAYERS):
            w = np.zeros((9, CH, CH), np.float32)
            std = math.sqrt(2.0 / (9 * cin)) if li < len(LAYERS) - 1 else 1e-3
            w[:, :cout, :cin] = rng.normal(0.0, std, (9, cout, cin))
            ws.append(w)
            bs.append(np.zeros(CH, np.float32))
        return cls(ws, bs)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "Model":
        z = np.load(path, allow_pickle=False)
        ws, bs = [], []
        for li, (cin, cout) in enumerate(LAYERS):
            w = np.zeros((9, CH, CH), np.float32)
            w[:, :cout, :cin] = z[f"w{li}"]
            b = np.zeros(CH, np.float32)
            b[:cout] = z[f"b{li}"]
            ws.append(w)
            bs.append(b)
        m = cls(ws, bs)
        m.meta = json.loads(str(z["meta"])) if "meta" in z else {}
        return m

    def save(self, path: str, meta: Optional[dict] = None):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        arrs = {}
        for li, (cin, cout) in enumerate(LAYERS):
            arrs[f"w{li}"] = self.w[li].cpu().numpy()[:, :cout, :cin].copy()
            arrs[f"b{li}"] = self.b[li].cpu().numpy()[:cout].copy()
        arrs["meta"] = np.array(json.dumps(meta or self.meta))
        np.savez_compressed(path, **arrs)

    def pack(self, stream=None):
        L = _lib.load()
        for li in range(len(LAYERS)):
            _check(L.vsz_cnn_pack_weights(_p(self.w[li]), CH, CH, _p(self.wf[li]), _p(self.wb[li]), _s(stream)))

    def cmodel(self) -> _lib.CnnModel:
        m = _lib.CnnModel()
        for li in range(len(LAYERS)):
            m.w[li] = self.wf[li].data_ptr()
            m.b[li] = self.b[li].data_ptr()
        return m

    # ---- inference ----------------------------------------------------------------------------
    def run(self, noisy, sigma255, out=None, work=None, stream=None):
        """noisy: [n, h, stride] (or [h, stride]) f32 device tensor of planes of width `w` = stride
        unless given via out's shape; sigma255: per-image raw sigma(s). Returns out (same shape)."""
        import torch
        x = noisy if noisy.dim() == 3 else noisy[None]
        n, h, stride = x.shape
        wd = stride
        sig = torch.tensor(np.broadcast_to(np.asarray(sigma255, np.float32) / 255.0, (n,)).copy(), device="cuda")
        y = out if out is not None else torch.empty_like(x)
        y = y if y.dim() == 3 else y[None]
        L = _lib.load()
        if work is None:
            work = torch.empty(L.vsz_cnn_workspace_size(n, h, wd), dtype=torch.uint8, device="cuda")
        cm = self.cmodel()
        _check(L.vsz_cnn_run(C.byref(cm), _p(x), n, h, wd, stride, _p(sig), _p(y), _p(work), _s(stream)))
        return y if noisy.dim() == 3 else y[0]


class Trainer:
    """Adam on the FP32 masters; forward keeps every activation for the backward pass."""

    def __init__(self, model: Model, batch: int, h: int, w: int, lr: float = 1e-3):
        import torch
        self.m, self.B, self.h, self.wd = model, batch, h, w
        px = batch * h * w
        act = lambda: torch.zeros(px * CH, dtype=torch.int16, device="cuda")  # noqa: E731
        self.a = [act() for _ in range(len(LAYERS))]   # a0 = packed input, a1..a5 = hidden
        self.g = [act(), act()]
        self.base = torch.zeros(px, dtype=torch.float32, device="cuda")
        self.out = torch.zeros(px, dtype=torch.float32, device="cuda")
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.dw = [torch.zeros(9 * CH * CH, dtype=torch.float32, device="cuda") for _ in LAYERS]
        self.db = [torch.zeros(CH, dtype=torch.float32, device="cuda") for _ in LAYERS]
        self.mw = [torch.zeros_like(d) for d in self.dw]
        self.vw = [torch.zeros_like(d) for d in self.dw]
        self.mb = [torch.zeros_like(d) for d in self.db]
        self.vb = [torch.zeros_like(d) for d in self.db]
        self.lr, self.t = lr, 0

    def step(self, noisy, target, sigma_map, lr=None):
        """noisy, target: [B, h, w] f32 device tensors; sigma_map: [B] f32 (sigma/255)."""
        L, s = _lib.load(), _s(None)
        B, h, w = self.B, self.h, self.wd
        m = self.m
        _check(L.vsz_cnn_pack_input(_p(noisy), B, h, w, w, _p(sigma_map), _p(self.a[0]), _p(self.base), s))
        for li in range(5):
            _check(L.vsz_cnn_conv3x3(_p(self.a[li]), _p(m.wf[li]), _p(m.b[li]), B, h, w, 0, None, None,
                                     _p(self.a[li + 1]), 0, s))
        _check(L.vsz_cnn_conv3x3(_p(self.a[5]), _p(m.wf[5]), _p(m.b[5]), B, h, w, 1, None, _p(self.base),
                                 _p(self.out), w, s))
        self.loss.zero_()
        for d in self.dw + self.db:
            d.zero_()
        _check(L.vsz_cnn_loss_grad(_p(self.out), _p(target), B * h * w, _p(self.g[0]), _p(self.loss), s))
        cur = 0
        for li in range(5, -1, -1):   # g[cur] = gradient w.r.t. layer li's pre-activation
            _check(L.vsz_cnn_conv3x3_dw(_p(self.a[li]), _p(self.g[cur]), B, h, w, _p(self.dw[li]), s))
            _check(L.vsz_cnn_bias_grad(_p(self.g[cur]), B * h * w, _p(self.db[li]), s))
            if li > 0:
                _check(L.vsz_cnn_conv3x3(_p(self.g[cur]), _p(m.wb[li]), None, B, h, w, 2, _p(self.a[li]), None,
                                         _p(self.g[1 - cur]), 0, s))
                cur = 1 - cur
        self.t += 1
        lr = self.lr if lr is None else lr
        for li in range(len(LAYERS)):
            _check(L.vsz_cnn_adam(_p(m.w[li]), _p(self.dw[li]), _p(self.mw[li]), _p(self.vw[li]), 9 * CH * CH,
                                  lr, 0.9, 0.999, 1e-8, self.t, s))
            _check(L.vsz_cnn_adam(_p(m.b[li]), _p(self.db[li]), _p(self.mb[li]), _p(self.vb[li]), CH,
                                  lr, 0.9, 0.999, 1e-8, self.t, s))
        m.pack()
        return self.loss


# ---- training data: the engine's own basic estimates ------------------------------------------
BASIC_PARAMS = dict(block_step=4, bm_range=16, extractor_exp=3)


def make_pairs(count: int, patch: int = 128, frame: int = 1024, draws_per_frame: int = 8, seed: int = 0,
               sigma_range=(2.0, 25.0), chroma_frac: float = 0.0, log_sigma: bool = False):
    """(noisy, basic, sigma/255) patches: synthetic clean frames, Gaussian noise with sigma drawn
    from sigma_range (per frame draw; log-uniform with log_sigma), BM3D basic pass of the engine on
    the whole frame, then the frame cut into patch x patch tiles. A fraction chroma_frac of the
    clean frames comes from the low-frequency chroma generator ([-0.5, 0.5]) instead of the luma
    one. Returns device tensors [count, patch, patch] x 2, [count]."""
    import torch
    per = (frame // patch) ** 2
    draws = (count + per - 1) // per
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    kp_work = None
    noisy_p, basic_p, sig_p = [], [], []
    clean = None
    for d in range(draws):
        if d % draws_per_frame == 0:
            gen_fn = synth.clean_chroma if rng.random() < chroma_frac else synth.clean_luma
            clean = torch.from_numpy(gen_fn(frame, frame, seed * 1000 + d)).cuda()
        if log_sigma:
            s255 = float(np.exp(rng.uniform(np.log(sigma_range[0]), np.log(sigma_range[1]))))
        else:
            s255 = float(rng.uniform(*sigma_range))
        nz = clean + torch.randn(frame, frame, generator=gen, device="cuda") * (s255 / 255.0)
        kp = device.kparams(frame, frame, sigma=device.scaled_sigma([s255] * 3, False), **BASIC_PARAMS)
        if kp_work is None:
            kp_work = torch.zeros(device.workspace_size(kp), dtype=torch.uint8, device="cuda")
        basic = torch.zeros((1, frame, frame), dtype=torch.float32, device="cuda")
        device.run(kp, nz[None, None], None, kp_work, basic, 0)
        k = frame // patch
        noisy_p.append(nz.view(k, patch, k, patch).permute(0, 2, 1, 3).reshape(-1, patch, patch))
        basic_p.append(basic[0].view(k, patch, k, patch).permute(0, 2, 1, 3).reshape(-1, patch, patch))
        sig_p.append(torch.full((per,), s255 / 255.0, device="cuda"))
    return (torch.cat(noisy_p)[:count].contiguous(), torch.cat(basic_p)[:count].contiguous(),
            torch.cat(sig_p)[:count].contiguous())


def train(count=20480, patch=128, batch=32, epochs=8, lr=1e-3, seed=0, log=print, chroma_frac=0.0,
          log_sigma=False):
    """Generates the pairs, trains with Adam and a cosine learning-rate decay, returns the model."""
    import torch
    t0 = time.perf_counter()
    noisy, basic, sig = make_pairs(count, patch, seed=seed, chroma_frac=chroma_frac, log_sigma=log_sigma)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    m = Model.init(seed)
    tr = Trainer(m, batch, patch, patch, lr)
    steps = epochs * (count // batch)
    gen = torch.Generator(device="cuda").manual_seed(seed + 1)
    t0 = time.perf_counter()
    hist = []
    for st in range(steps):
        if st % (count // batch) == 0:
            perm = torch.randperm(count, generator=gen, device="cuda")
        idx = perm[(st % (count // batch)) * batch:(st % (count // batch) + 1) * batch]
        cur_lr = 1e-5 + 0.5 * (lr - 1e-5) * (1.0 + math.cos(math.pi * st / steps))
        loss = tr.step(noisy[idx].contiguous(), basic[idx].contiguous(), sig[idx].contiguous(), cur_lr)
        if st % 200 == 0 or st == steps - 1:
            v = float(loss.item())
            hist.append((st, v))
            log(f"cnn: step {st}/{steps} loss {v:.3e} (rmse {math.sqrt(v) * 255:.3f}/255) lr {cur_lr:.2e}")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    m.meta = {"pairs": count, "patch": patch, "batch": batch, "epochs": epochs, "steps": steps, "lr": lr,
              "seed": seed, "basic_params": BASIC_PARAMS, "sigma255": [2.0, 25.0], "chroma_frac": chroma_frac,
              "log_sigma": log_sigma, "data_s": round(t_data, 2),
              "train_s": round(t_train, 2), "loss_history": hist}
    return m


# ---- the judge: where the learned basic pass may replace the plain one --------------------------
JUDGE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "cnn_judge.json")
JUDGE_BOUND_DB = 0.05
JUDGE_PARAMS = dict(block_step=4, bm_range=16, extractor_exp=3)


def _two_pass_psnr(noisy_np, clean_np, s255, model, work_cache):
    """(PSNR of plain basic + Wiener, PSNR of learned basic + Wiener) of one Gray plane."""
    import torch
    h, w = noisy_np.shape
    nz = torch.from_numpy(noisy_np).cuda()
    kb = device.kparams(w, h, sigma=device.scaled_sigma([s255] * 3, False), **JUDGE_PARAMS)
    kf = device.kparams(w, h, sigma=device.scaled_sigma([s255] * 3, True), final=True, **JUDGE_PARAMS)
    need = max(device.workspace_size(kb), device.workspace_size(kf))
    if work_cache.get("n", 0) < need:
        work_cache["t"] = torch.zeros(need, dtype=torch.uint8, device="cuda")
        work_cache["n"] = need
    work = work_cache["t"]
    basic = torch.zeros((1, h, w), dtype=torch.float32, device="cuda")
    out = torch.zeros((1, h, w), dtype=torch.float32, device="cuda")
    device.run(kb, nz[None, None], None, work, basic, 0)
    device.run(kf, basic[:, None], nz[None, None], work, out, 0)
    plain = synth.psnr(out[0].cpu().numpy(), clean_np)
    model.run(nz[None], s255, out=basic)
    device.run(kf, basic[:, None], nz[None, None], work, out, 0)
    learned = synth.psnr(out[0].cpu().numpy(), clean_np)
    return plain, learned


def measure_judge(model=None, sigmas=(2.0, 3.0, 5.0, 8.0, 12.0, 16.0, 20.0, 25.0), size=768, seeds=(11, 12, 13)):
    """PSNR delta (learned basic + Wiener minus plain basic + Wiener, against the clean plane) per
    sigma and plane kind, averaged over seeded synthetic planes: "luma" from the luma generator,
    "chroma" from the low-frequency chroma generator ([-0.5, 0.5]). Runs on the GPU."""
    model = model or Model.load()
    cache = {}
    table = {"bound_db": JUDGE_BOUND_DB, "params": JUDGE_PARAMS, "size": size, "seeds": list(seeds), "rows": []}
    for kind in ("luma", "chroma"):
        for s255 in sigmas:
            d = []
            for sd in seeds:
                clean = synth.clean_luma(size, size, sd) if kind == "luma" else synth.clean_chroma(size, size, sd)
                nz = synth.noisy(clean, s255, sd + 500)
                pp, pl = _two_pass_psnr(nz, clean, s255, model, cache)
                d.append(pl - pp)
            table["rows"].append({"kind": kind, "sigma": s255, "delta_db": round(float(np.mean(d)), 4)})
    return table


def load_judge(path: str = JUDGE_PATH) -> Optional[dict]:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def learned_allowed(sigma255: float, kind: str, table: Optional[dict]) -> bool:
    """The judge's rule: the learned basic pass replaces the plain one for a plane of this kind
    and sigma only if the measured PSNR delta at both measured sigmas around it (or at that sigma)
    is within the bound; outside the measured range, or without a table, the plain pass runs."""
    if not table:
        return False
    rows = sorted((r["sigma"], r["delta_db"]) for r in table["rows"] if r["kind"] == kind)
    if not rows or sigma255 < rows[0][0] or sigma255 > rows[-1][0]:
        return False
    lo = max(r for r in rows if r[0] <= sigma255)
    hi = min(r for r in rows if r[0] >= sigma255)
    return lo[1] >= -table["bound_db"] and hi[1] >= -table["bound_db"]


def main(argv=None):
    """python -m vszipcu.cnn --epochs 24 --out vszipcu/data/cnn_basic.npz (needs the GPU)."""
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--pairs", type=int, default=20480)
    ap.add_argument("--patch", type=int, default=128)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--chroma-frac", type=float, default=0.0, help="share of chroma-generator training frames")
    ap.add_argument("--log-sigma", action="store_true", help="draw training sigmas log-uniformly")
    ap.add_argument("--judge", action="store_true",
                    help="measure the PSNR-delta table of the saved weights into vszipcu/data/cnn_judge.json")
    a = ap.parse_args(argv)
    if a.judge:
        t = measure_judge(Model.load(a.out))
        with open(JUDGE_PATH, "w") as f:
            json.dump(t, f, indent=1)
        for r in t["rows"]:
            print(f"cnn judge: {r['kind']:6s} sigma {r['sigma']:5.1f}: delta {r['delta_db']:+.4f} dB", flush=True)
        return
    m = train(a.pairs, a.patch, a.batch, a.epochs, a.lr, a.seed, log=lambda s: print(s, flush=True),
              chroma_frac=a.chroma_frac, log_sigma=a.log_sigma)
    m.save(a.out)
    print(f"cnn: saved {a.out} ({m.meta['train_s']} s training, {m.meta['data_s']} s data)", flush=True)


if __name__ == "__main__":
    main()
