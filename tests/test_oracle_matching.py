"""Pins the oracle's block matching (src/bm3d.cu:579-788) by brute force on tiny frames.

The brute force below is written from the reference's rules, not from the oracle: float64
distances of every candidate of the clipped window, a stable sort by (distance, scan
position), the first 8 kept (padded with (FLT_MAX, 0, 0)), then self-inclusion. The temporal
predictive search is restated the same way. Exact ties are planted with periodic frames.
"""
import numpy as np
import pytest

FLT_MAX = float(np.finfo(np.float32).max)


def dist64(metric, a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    if metric == "ssd":
        return float(np.sum((a - b) ** 2))
    if metric == "sad":
        return float(np.sum(np.abs(a - b)))
    if metric in ("zssd", "zsad"):
        d = (a - b) - (a.mean() - b.mean())
        return float(np.sum(d * d)) if metric == "zssd" else float(np.sum(np.abs(d)))
    na, nb = np.sqrt(np.sum(a * a)), np.sqrt(np.sum(b * b))
    return float(2.0 - 2.0 * np.sum(a * b) / (na * nb + np.finfo(np.float32).eps))


def origins(n, step):
    # reference-block origins b*step while b*step < n-8+step, clamped to n-8 (src/bm3d.cu:588-595)
    out, b = [], 0
    while b * step < n - 8 + step:
        out.append(min(b * step, n - 8))
        b += 1
    return out


def top8(cands):
    """cands: list of (err, x, y, z) in scan order -> stable sort, first 8, padded."""
    order = sorted(range(len(cands)), key=lambda i: (cands[i][0], i))
    lst = [cands[i] for i in order if cands[i][0] < FLT_MAX][:8]
    return lst


def brute_match(frames, x, y, bs_range, R=0, ps_num=2, ps_range=4, metric="ssd"):
    TW, H, W = frames.shape
    ref = frames[R, y:y + 8, x:x + 8]

    def window(cx, cy, rng, z):
        c = []
        for yy in range(max(cy - rng, 0), min(cy + rng, H - 8) + 1):
            for xx in range(max(cx - rng, 0), min(cx + rng, W - 8) + 1):
                c.append((dist64(metric, ref, frames[z, yy:yy + 8, xx:xx + 8]), xx, yy, z))
        return c

    lst = top8(window(x, y, bs_range, R))
    pad = lambda l, z: l + [(FLT_MAX, 0, 0, z)] * (8 - len(l))  # noqa: E731
    lst = pad(lst, R)
    if R > 0:
        centre = list(lst)
        for d in (-1, 1):
            last = centre
            for t in range(1, R + 1):
                tz = R + d * t
                fc = []
                for i in range(ps_num):
                    fc += window(last[i][1], last[i][2], ps_range, tz)
                f = pad(top8(fc), tz)
                # merge the first ps_num of the frame list (strict <: after equal errors)
                merged = lst + [(e, fx, fy, tz) for (e, fx, fy, _) in f[:ps_num]]
                lst = sorted(range(len(merged)), key=lambda i: (merged[i][0], i))
                lst = [merged[i] for i in lst][:8]
                last = f
    if not any(e[1] == x and e[2] == y and (R == 0 or e[3] == R) for e in lst):
        lst = [(0.0, x, y, R)] + lst[:7]
    return lst


def run_oracle(orc, frames, bs, rng, R=0, ps_num=2, ps_range=4, metric="ssd"):
    TW, H, W = frames.shape
    p = orc.make_params(W, H, block_step=bs, bm_range=rng, radius=R, ps_num=ps_num, ps_range=ps_range,
                        metric=metric, threads=2)
    return orc.match(p, frames)


def compare(orc, frames, bs, rng, R=0, metric="ssd", ps_num=2, ps_range=4, check_err=True):
    TW, H, W = frames.shape
    m, e = run_oracle(orc, frames, bs, rng, R, ps_num, ps_range, metric)
    ys, xs = origins(H, bs), origins(W, bs)
    assert m.shape[:2] == (len(ys), len(xs))
    for by, y in enumerate(ys):
        for bx, x in enumerate(xs):
            want = brute_match(frames, x, y, rng, R, ps_num, ps_range, metric)
            got = [tuple(v) for v in m[by, bx]]
            assert got == [(w[1], w[2], w[3]) for w in want], (x, y, got, want)
            if check_err:
                for k, w in enumerate(want):
                    if w[0] < FLT_MAX:
                        assert abs(float(e[by, bx, k]) - w[0]) <= 1e-5 * (1 + abs(w[0])), (x, y, k)


@pytest.mark.parametrize("bs,rng", [(1, 3), (3, 5), (4, 9), (8, 30)])
def test_spatial_ssd_brute_force(orc, bs, rng):
    frames = np.random.default_rng(bs * 100 + rng).random((1, 21, 26), dtype=np.float32)
    compare(orc, frames, bs, rng)


@pytest.mark.parametrize("metric", ["sad", "zssd", "zsad", "ssd/norm"])
def test_spatial_metrics_brute_force(orc, metric):
    frames = np.random.default_rng(3).random((1, 18, 20), dtype=np.float32) + 0.25
    compare(orc, frames, 2, 4, metric=metric)


def test_planted_ties_keep_scan_order(orc):
    # period-4 texture: every shift by a multiple of 4 matches exactly (distance 0)
    tile = np.random.default_rng(5).random((4, 4), dtype=np.float32)
    frames = np.tile(tile, (6, 7))[None, :22, :26].copy()
    compare(orc, frames, 3, 9)
    m, e = run_oracle(orc, frames, 3, 9)
    assert np.all(e[..., :3] == 0.0)


def test_fewer_than_eight_candidates(orc):
    frames = np.random.default_rng(9).random((1, 9, 8), dtype=np.float32)
    m, e = run_oracle(orc, frames, 8, 9)
    assert m.shape == (2, 1, 8, 3)
    # two positions exist; the other six entries are (FLT_MAX, 0, 0) fill
    assert np.all(e[:, 0, 2:] == np.float32(FLT_MAX))
    assert np.all(m[:, 0, 2:, :2] == 0)
    compare(orc, frames, 8, 9)


@pytest.mark.parametrize("R,ps_num,ps_range", [(1, 2, 2), (2, 3, 1), (1, 8, 3)])
def test_temporal_predictive_search_brute_force(orc, R, ps_num, ps_range):
    rng = np.random.default_rng(R * 10 + ps_num)
    base = rng.random((20, 24), dtype=np.float32)
    frames = np.stack([np.roll(base, (t, -t), axis=(0, 1)) + 0.05 * rng.random((20, 24), dtype=np.float32)
                       for t in range(2 * R + 1)]).astype(np.float32)
    compare(orc, frames, 4, 4, R=R, ps_num=ps_num, ps_range=ps_range)


def test_temporal_duplicate_frames(orc):
    # clamped windows repeat frames: zero-distance ties across z resolve by merge order
    base = np.random.default_rng(11).random((16, 17), dtype=np.float32)
    frames = np.stack([base, base, base]).astype(np.float32)
    compare(orc, frames, 4, 3, R=1, ps_num=2, ps_range=2)


def test_block_distance_matches_float64(orc):
    rng = np.random.default_rng(1)
    for metric in ["ssd", "sad", "zssd", "zsad", "ssd/norm"]:
        for _ in range(50):
            a = rng.random((8, 8), dtype=np.float32)
            b = rng.random((8, 8), dtype=np.float32)
            got = orc.block_distance(metric, a, b)
            want = dist64(metric, a, b)
            assert abs(got - want) <= 2e-6 * (1 + abs(want)), metric


def test_chroma_matching_ignores_uv(orc):
    rng = np.random.default_rng(4)
    y = rng.random((1, 24, 24), dtype=np.float32)
    src1 = np.stack([y, rng.random((1, 24, 24), dtype=np.float32), rng.random((1, 24, 24), dtype=np.float32)])[None]
    src2 = np.stack([y, rng.random((1, 24, 24), dtype=np.float32), rng.random((1, 24, 24), dtype=np.float32)])[None]
    p = orc.make_params(24, 24, block_step=4, bm_range=5, chroma=True, sigma=(1.0, 1.0, 1.0), threads=2)
    _, m1, _ = orc.bm3d(p, src1)
    _, m2, _ = orc.bm3d(p, src2)
    assert np.array_equal(m1, m2)
