"""Pins the oracle's 8-point transforms (src/bm3d.cu:32-280) against closed forms.

DCT: forward = 4 * orthonormal DCT-II, inverse = 4 * its transpose (scipy reference).
WHT: sqrt(2) * the sequency-ordered Walsh matrix; self-inverse up to 16.
bior1.5: synthesis * analysis = 16 * I.
Haar: forward = 4 * orthonormal Haar; the reference inverse reads only v0, v1, v2, v4 (rank 4),
reproduced literally.
The float32 instances follow the same op sequence; they must agree with float64 to ~1 ulp of
the operands' scale.
"""
import numpy as np
import pytest
import scipy.fft


def test_dct_forward_is_4x_orthonormal(orc):
    M = orc.transform_matrix("dct", True)
    C = scipy.fft.dct(np.eye(8), norm="ortho", axis=0)
    np.testing.assert_allclose(M, 4 * C, atol=1e-13)


def test_dct_inverse_is_4x_transpose(orc):
    M = orc.transform_matrix("dct", False)
    C = scipy.fft.dct(np.eye(8), norm="ortho", axis=0)
    np.testing.assert_allclose(M, 4 * C.T, atol=1e-13)


def test_dct_round_trip_16(orc):
    F, I = orc.transform_matrix("dct", True), orc.transform_matrix("dct", False)
    np.testing.assert_allclose(I @ F, 16 * np.eye(8), atol=1e-12)


def _sylvester(n):
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    return H


def test_wht_is_scaled_hadamard_rows(orc):
    W = orc.transform_matrix("wht", True)
    np.testing.assert_allclose(np.abs(W), np.sqrt(2.0), atol=1e-14)
    H = _sylvester(8) * np.sqrt(2.0)
    # every row of W is a row of sqrt(2) * H, each used once
    used = set()
    for r in W:
        hits = [i for i in range(8) if np.allclose(r, H[i], atol=1e-13)]
        assert len(hits) == 1
        used.add(hits[0])
    assert used == set(range(8))
    np.testing.assert_allclose(W @ W, 16 * np.eye(8), atol=1e-12)
    np.testing.assert_allclose(orc.transform_matrix("wht", False), W, atol=0)


def test_wht_sequency_order(orc):
    # sequency-ordered Walsh matrix: row k has exactly k sign changes
    W = orc.transform_matrix("wht", True)
    changes = [int(np.sum(np.diff(np.sign(r)) != 0)) for r in W]
    assert changes == list(range(8))


def test_bior15_perfect_reconstruction(orc):
    F, I = orc.transform_matrix("bior1.5", True), orc.transform_matrix("bior1.5", False)
    np.testing.assert_allclose(I @ F, 16 * np.eye(8), atol=1e-12)
    # analysis row 0 is the scaled average, rows 4..7 the Haar details of pairs
    np.testing.assert_allclose(F[0], np.sqrt(2.0) * np.ones(8), atol=1e-14)
    for k in range(4):
        e = np.zeros(8)
        e[2 * k], e[2 * k + 1] = 2 * np.sqrt(2.0), -2 * np.sqrt(2.0)
        np.testing.assert_allclose(F[4 + k], e, atol=1e-14)


def test_haar_forward_is_4x_orthonormal(orc):
    F = orc.transform_matrix("haar", True)
    np.testing.assert_allclose(F @ F.T, 16 * np.eye(8), atol=1e-12)
    np.testing.assert_allclose(F[0], np.sqrt(2.0) * np.ones(8), atol=1e-14)


def test_haar_inverse_reference_defect(orc):
    I = orc.transform_matrix("haar", False)
    # reads only v0, v1, v2, v4 (src/bm3d.cu:158-177): columns 3, 5, 6, 7 are zero
    for col in (3, 5, 6, 7):
        assert np.all(I[:, col] == 0.0)
    assert np.linalg.matrix_rank(I) == 4
    F = orc.transform_matrix("haar", True)
    assert not np.allclose(I @ F, 16 * np.eye(8), atol=1e-6)
    # literal closed form of the inverse
    s2, q = np.sqrt(2.0), 2.0
    expect = np.zeros((8, 8))
    for o in range(8):
        expect[o, 0] = s2
        expect[o, 1] = s2 if o < 4 else -s2
        sgn2 = [1, 1, -1, -1][o % 4]
        sgn4 = [1, -1, 1, -1][o % 4]
        expect[o, 2] = s2 * s2 * sgn2
        expect[o, 4] = s2 * q * sgn4
    np.testing.assert_allclose(I, expect, atol=1e-13)


@pytest.mark.parametrize("kind", ["dct", "haar", "wht", "bior1.5"])
@pytest.mark.parametrize("fwd", [True, False])
def test_f32_instance_tracks_f64(orc, kind, fwd):
    rng = np.random.default_rng(7)
    for _ in range(200):
        v = rng.standard_normal(8).astype(np.float32)
        a = orc.transform(kind, fwd, v, np.float32).astype(np.float64)
        b = orc.transform(kind, fwd, v.astype(np.float64), np.float64)
        np.testing.assert_allclose(a, b, atol=2e-6 * (1 + np.abs(v).sum()))


def test_codelet_constants_closed_forms():
    # the decimal literals used by both the oracle and the kernels are these closed forms
    import math
    pairs = {
        0.414213562373095048801688724209698078569671875: math.tan(math.pi / 8),
        1.847759065022573512256366378793576573644833252: 2 * math.cos(math.pi / 8),
        0.198912367379658006911597622644676228597850501: math.tan(math.pi / 16),
        1.961570560806460898252364472268478073947867462: 2 * math.cos(math.pi / 16),
        0.668178637919298919997757686523080761552472251: math.tan(3 * math.pi / 16),
        1.662939224605090474157576755235811513477121624: 2 * math.cos(3 * math.pi / 16),
        0.707106781186547524400844362104849039284835938: 1 / math.sqrt(2),
        1.414213562373095048801688724209698078569671875: math.sqrt(2),
        2.828427124746190097603377448419396157139343751: 2 * math.sqrt(2),
    }
    for lit, closed in pairs.items():
        assert abs(lit - closed) < 1e-15
        assert np.float32(lit) == np.float32(closed)
