"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

CPU tier: the oracle still reproduces every fixture bit for bit (guards the checker against
drift). GPU tier: the kernel layer reproduces the fixtures bit for bit (match lists, block
distances, num/den stack, output), with the extractor-0 case within FP32-atomic tolerance.
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
KEYS = ("block_step", "bm_range", "radius", "ps_num", "ps_range", "chroma", "final", "extractor_exp")


def _load(path):
    z = np.load(path)  # allow_pickle defaults to False
    prm = dict(zip(KEYS, [int(v) for v in z["params"]]))
    return z, prm


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_fixture(orc, path):
    z, prm = _load(path)
    src = z["src"]
    H, W = src.shape[-2:]
    p = orc.make_params(W, H, sigma=tuple(float(s) for s in z["sigma"]), **prm)
    res, m, e = orc.bm3d(p, src)
    assert np.array_equal(m, z["matches"])
    assert np.array_equal(e, z["errors"])
    assert np.array_equal(res, z["res"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_gpu_reproduces_fixture(torch_gpu, path):
    from gpu_util import gpu_pass
    from vszipcu import device
    z, prm = _load(path)
    src = z["src"]
    H, W = src.shape[-2:]
    kp = device.kparams(W, H, block_step=prm["block_step"], bm_range=prm["bm_range"], radius=prm["radius"],
                        ps_num=prm["ps_num"], ps_range=prm["ps_range"], chroma=bool(prm["chroma"]),
                        final=bool(prm["final"]), sigma=[float(s) for s in z["sigma"]],
                        extractor_exp=prm["extractor_exp"])
    res, gm = gpu_pass(torch_gpu, kp, src, 1)
    assert np.array_equal(gm[..., :3], z["matches"])
    assert np.array_equal(gm[..., 3].copy().view(np.float32), z["errors"])
    if prm["extractor_exp"]:
        assert np.array_equal(res, z["res"])
    else:
        np.testing.assert_allclose(res, z["res"], rtol=2e-6, atol=1e-6)
    if "dst" in z:
        dst, _ = gpu_pass(torch_gpu, kp, src, 0)
        tol = 0.0 if prm["extractor_exp"] else 1e-5
        assert np.nanmax(np.abs(dst - z["dst"])) <= tol
