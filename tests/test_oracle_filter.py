"""Pins the oracle's collaborative filtering and aggregation (src/bm3d.cu:416-546, 790-922).

Single groups are checked against scipy's orthonormal 3-D DCT: the reference's forward 3-D
transform is 64 * orthonormal, its inverse 64 * orthonormal inverse, and the 1/4096 factor in
the shrinkage undoes both, so
  hard:   out = idctn(keep * dctn(g)),                weight = 1 / #kept   (DC always kept)
  Wiener: out = idctn(c * dctn(g)), c = R^2/(R^2+s^2), weight = 1 / sum c^2 (DC c = 1)
with R = 64 * dctn(ref) and the threshold compared against 64 * dctn(g).
Whole frames are checked through invariants: sigma -> 0 reproduces the input, a constant frame
stays constant, the Wiener pass with a fixed reference is linear in the noisy input, and the
exact (extractor) accumulation is independent of the thread count.
"""
import numpy as np
import pytest
import scipy.fft


def dctn(x):
    return scipy.fft.dctn(x.astype(np.float64), norm="ortho")


def idctn(x):
    return scipy.fft.idctn(x, norm="ortho")


def test_hard_group_closed_form(orc):
    rng = np.random.default_rng(0)
    for trial in range(20):
        g = rng.random((8, 8, 8)).astype(np.float32)
        coef = 64 * dctn(g)
        sigma = float(np.quantile(np.abs(coef), 0.8))
        if np.min(np.abs(np.abs(coef) - sigma)) < 1e-3:  # keep decisions away from rounding
            continue
        keep = np.abs(coef) >= sigma
        keep[0, 0, 0] = True
        want = idctn(np.where(keep, dctn(g), 0.0))
        got, w = orc.filter_group(g, sigma)
        np.testing.assert_allclose(got, want, atol=2e-5)
        assert w == np.float32(1.0 / keep.sum())


def test_hard_group_huge_sigma_gives_mean(orc):
    g = np.random.default_rng(1).random((8, 8, 8)).astype(np.float32)
    got, w = orc.filter_group(g, 1e30)
    assert w == 1.0
    np.testing.assert_allclose(got, np.full_like(g, g.astype(np.float64).mean()), atol=2e-6)


def test_wiener_group_closed_form(orc):
    rng = np.random.default_rng(2)
    for trial in range(10):
        g = rng.random((8, 8, 8)).astype(np.float32)
        r = (g + 0.1 * rng.standard_normal((8, 8, 8))).astype(np.float32)
        sigma = 3.0
        R = 64 * dctn(r)
        c = R * R / (R * R + sigma * sigma)
        c[0, 0, 0] = 1.0
        want = idctn(c * dctn(g))
        got, w = orc.filter_group(g, sigma, ref=r)
        np.testing.assert_allclose(got, want, atol=2e-5)
        assert abs(w - 1.0 / np.sum(c * c)) <= 1e-6 * (1.0 / np.sum(c * c))


def _frame_run(orc, img, *, sigma, bs=4, rng=6, final=False, ref=None, extractor_exp=0, threads=4, t="dct",
               dtype=np.float32, given=None):
    H, W = img.shape
    p = orc.make_params(W, H, block_step=bs, bm_range=rng, sigma=(sigma,) * 3, final=final,
                        extractor_exp=extractor_exp, threads=threads, t2d=t, t1d=t)
    src = np.stack([ref, img])[:, None, None] if final else img[None, None, None]
    res, m, e = orc.bm3d(p, src, dtype=dtype, given_matches=given)
    return orc.aggregate(res[:, 0], W, dtype=dtype)[0], res, m


@pytest.mark.parametrize("t", ["dct", "wht", "bior1.5"])
def test_sigma_to_zero_reproduces_input(orc, t):
    img = np.random.default_rng(3).random((40, 37)).astype(np.float32)
    out, _, _ = _frame_run(orc, img, sigma=1e-30, t=t)
    np.testing.assert_allclose(out, img, atol=5e-6)


def test_haar_defect_does_not_reproduce(orc):
    img = np.random.default_rng(3).random((24, 24)).astype(np.float32)
    out, _, _ = _frame_run(orc, img, sigma=1e-30, t="haar")
    assert np.max(np.abs(out - img)) > 1e-2


def test_constant_frame_stays_constant(orc):
    img = np.full((33, 41), 0.375, np.float32)
    out, _, _ = _frame_run(orc, img, sigma=5.0)
    np.testing.assert_allclose(out, 0.375, atol=1e-6)


def test_wiener_linear_in_noisy_input(orc):
    rng = np.random.default_rng(5)
    ref = rng.random((32, 32)).astype(np.float32)
    x = rng.random((32, 32)).astype(np.float32)
    y = rng.random((32, 32)).astype(np.float32)
    a, b = 0.75, -0.5
    ox, _, _ = _frame_run(orc, x, sigma=0.05, final=True, ref=ref)
    oy, _, _ = _frame_run(orc, y, sigma=0.05, final=True, ref=ref)
    oxy, _, _ = _frame_run(orc, (a * x + b * y).astype(np.float32), sigma=0.05, final=True, ref=ref)
    np.testing.assert_allclose(oxy, a * ox + b * oy, atol=2e-5)


def test_exact_accumulation_thread_independent(orc):
    img = np.random.default_rng(6).random((48, 40)).astype(np.float32)
    o1, r1, _ = _frame_run(orc, img, sigma=2.0, extractor_exp=3, threads=1)
    o8, r8, _ = _frame_run(orc, img, sigma=2.0, extractor_exp=3, threads=8)
    assert np.array_equal(r1, r8) and np.array_equal(o1, o8)
    # every accumulated value lies on the 2^(e-24) grid
    assert np.all(np.mod(r1.astype(np.float64) * 2.0 ** 21, 1.0) == 0.0)


def test_float_accumulation_raster_order_thread_independent(orc):
    img = np.random.default_rng(7).random((48, 40)).astype(np.float32)
    _, r1, _ = _frame_run(orc, img, sigma=2.0, threads=1)
    _, r8, _ = _frame_run(orc, img, sigma=2.0, threads=8)
    assert np.array_equal(r1, r8)


def test_f32_tracks_f64_with_same_matches(orc):
    rng = np.random.default_rng(8)
    clean = rng.random((40, 40)).astype(np.float32)
    noisy = (clean + 0.05 * rng.standard_normal((40, 40))).astype(np.float32)
    o32, _, m = _frame_run(orc, noisy, sigma=5.0)
    o64, _, _ = _frame_run(orc, noisy.astype(np.float64), sigma=5.0, dtype=np.float64, given=m)
    # hard-threshold decisions can flip within rounding; bound the typical and worst error
    d = np.abs(o32 - o64)
    assert np.median(d) < 1e-6 and d.max() < 5e-3


def test_extractor_quantises_contributions(orc):
    # E = 2^3: a value u contributes (u + 8) - 8, so tiny weights vanish entirely
    img = np.random.default_rng(9).random((16, 16)).astype(np.float32)
    out0, r0, _ = _frame_run(orc, img, sigma=2.0, extractor_exp=0)
    out3, r3, _ = _frame_run(orc, img, sigma=2.0, extractor_exp=3)
    assert not np.array_equal(r0, r3)
    # hard-pass weights 1/k reach 1/512; on the 2^-21 grid that is ~2.5e-4 relative
    np.testing.assert_allclose(out3, out0, atol=1e-3)
