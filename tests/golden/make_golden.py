"""Writes the golden fixtures tests/golden/*.npz from the FP32 oracle.

The reference ships no BM3D vectors (its GPU path cannot run here), so these fixtures are the
oracle's outputs on small seeded inputs. They pin the oracle against drift and give the GPU
tests a fixed target: inputs, kernel-level parameters, block-match lists (x, y, z), block
distances, the num/den stack and the divided output.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (REPO, os.path.join(REPO, "vapoursynth-zipcu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from vszipcu import device, synth  # noqa: E402

PARAM_KEYS = ("block_step", "bm_range", "radius", "ps_num", "ps_range", "chroma", "final", "extractor_exp")


def cases():
    # basic pass, Gray 64x48, block_step 4, bm_range 9
    c, nz = synth.gray_frame(64, 48, 10.0, seed=21)
    yield "gray_basic_64x48", dict(block_step=4, bm_range=9, radius=0, ps_num=2, ps_range=4, chroma=False,
                                   final=False, extractor_exp=3, sigma255=[10.0] * 3), nz[None, None, None]
    # Wiener pass on the oracle's basic estimate
    p = oracle.make_params(64, 48, block_step=4, bm_range=9, sigma=tuple(device.scaled_sigma([10.0] * 3, False)),
                           extractor_exp=3)
    res, _, _ = oracle.bm3d(p, nz[None, None, None])
    basic = oracle.aggregate(res[:, 0], 64)[0]
    yield "gray_final_64x48", dict(block_step=4, bm_range=9, radius=0, ps_num=2, ps_range=4, chroma=False,
                                   final=True, extractor_exp=3, sigma255=[10.0] * 3), np.stack([basic, nz])[:, None, None]
    # ragged 37x29, block_step 3, extractor 0
    c, nz = synth.gray_frame(37, 29, 15.0, seed=22)
    yield "gray_basic_37x29_ext0", dict(block_step=3, bm_range=5, radius=0, ps_num=2, ps_range=4, chroma=False,
                                        final=False, extractor_exp=0, sigma255=[15.0] * 3), nz[None, None, None]
    # temporal radius 1
    c, nz = synth.moving_sequence(40, 32, 3, 8.0, seed=23)
    yield "gray_temporal_r1_40x32", dict(block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=4, chroma=False,
                                         final=False, extractor_exp=3, sigma255=[8.0] * 3), nz[None, None]
    # chroma (CBM3D) 48x40
    c, nz = synth.yuv444_frame(48, 40, (5.0, 3.0, 3.0), seed=24)
    yield "yuv444_chroma_48x40", dict(block_step=4, bm_range=8, radius=0, ps_num=2, ps_range=4, chroma=True,
                                      final=False, extractor_exp=3, sigma255=[5.0, 3.0, 3.0]), nz[None, :, None]


def main():
    for name, prm, src in cases():
        H, W = src.shape[-2:]
        sig = device.scaled_sigma(prm["sigma255"], prm["final"])
        p = oracle.make_params(W, H, sigma=tuple(sig), **{k: prm[k] for k in PARAM_KEYS})
        res, m, e = oracle.bm3d(p, src)
        out = {"src": src, "res": res, "matches": m, "errors": e, "sigma": np.array(sig, np.float32),
               "params": np.array([prm[k] for k in PARAM_KEYS], np.int32)}
        if prm["radius"] == 0:
            out["dst"] = oracle.aggregate(res[:, 0], W)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
