"""Pins aggZ (src/bm3d.zig:46-48) and the temporal aggregation (src/bm3d_vagg.cu:18-48,
src/bm3d.zig:1284-1317) of the oracle.

aggZ: hand table for small (n, nf, r) including clamped borders, plus the defining property
that slot z of centre m = clamp(n-r+i) holds frame n (window[m][z] == n).
VAggregate: ascending float32 sums over the 2r+1 slots, then one IEEE division (numpy float32
elementwise ops in the same order give the same bits).
"""
import numpy as np
import pytest


def test_agg_z_hand_table(orc):
    # r = 1, nf = 10
    assert [orc.agg_z(i, 0, 10, 1) for i in range(3)] == [1, 1, 0]
    assert [orc.agg_z(i, 5, 10, 1) for i in range(3)] == [2, 1, 0]
    assert [orc.agg_z(i, 9, 10, 1) for i in range(3)] == [2, 1, 1]
    # r = 2, nf = 3 (every window clamps on both sides)
    assert [orc.agg_z(i, 1, 3, 2) for i in range(5)] == [3, 3, 2, 1, 1]


@pytest.mark.parametrize("r", [1, 2, 3, 5])
@pytest.mark.parametrize("nf", [1, 2, 4, 9, 20])
def test_agg_z_slot_holds_frame_n(orc, r, nf):
    clamp = lambda v: min(max(v, 0), nf - 1)  # noqa: E731
    for n in range(nf):
        for i in range(2 * r + 1):
            m = clamp(n - r + i)
            z = orc.agg_z(i, n, nf, r)
            assert 0 <= z <= 2 * r
            assert clamp(m - r + z) == n, (n, i, m, z)


def test_vaggregate_ascending_sums(orc):
    rng = np.random.default_rng(0)
    r, nf, n, H, W = 2, 7, 1, 12, 10
    TW = 2 * r + 1
    slots = [rng.random((TW, 2, H, W), dtype=np.float32) + 0.1 for _ in range(TW)]
    got = orc.vaggregate(slots, W, r, n, nf)
    sv = np.zeros((H, W), np.float32)
    sw = np.zeros((H, W), np.float32)
    for i in range(TW):
        z = orc.agg_z(i, n, nf, r)
        sv = sv + slots[i][z, 0]
        sw = sw + slots[i][z, 1]
    assert np.array_equal(got, sv / sw)
