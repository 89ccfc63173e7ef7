"""bench.py's host-side contract: the candidate enumeration behind `roofline.achieved` and the
single-GPU rule. The enumeration follows the clipped-window geometry of src/bm3d.cu:579-660;
the expected counts are the ones derived independently from that geometry for the 640x360
(bs 8, R 9) and 1920x1080 (bs 4, R 16) configurations."""
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def _brute(width, height, step, rng):
    total = 0
    for y in bench.origins(height, step):
        ny = min(y + rng, height - 8) - max(y - rng, 0) + 1
        for x in bench.origins(width, step):
            total += (min(x + rng, width - 8) - max(x - rng, 0) + 1) * ny
    return total


def test_candidate_counts_of_the_two_frames():
    assert bench.ssd_candidates(640, 360, 8, 9) == 1_252_500
    assert bench.ssd_candidates(1920, 1080, 4, 16) == 138_350_419


def test_candidate_count_matches_per_block_enumeration():
    for (w, h, bs, r) in [(64, 48, 4, 6), (72, 40, 8, 9), (99, 57, 3, 5), (16, 16, 8, 20)]:
        assert bench.ssd_candidates(w, h, bs, r) == _brute(w, h, bs, r)


def test_block_origins_clamp_to_the_last_full_block():
    o = bench.origins(21, 4)
    assert o.tolist() == [0, 4, 8, 12, 13] and np.all(o <= 21 - 8)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_workload_work_counts():
    """Blocks and temporal candidates of the temporal (1080p, bs 8, R 9, r 2, ps 2/4) and 4K
    (bs 8, R 9, r 1) workloads; the CNN's executed FLOPs per pixel (2->32, 4 x 32->32, 32->1)."""
    assert bench.num_blocks(1920, 1080, 8) == 240 * 135
    assert bench.num_blocks(3840, 2160, 8) == 480 * 270
    assert bench.temporal_candidates(1920, 1080, 8, 2, 2, 4) == 32400 * 648
    assert bench.temporal_candidates(3840, 2160, 8, 1, 2, 4) == 129600 * 324
    ssd = 192 * (bench.ssd_candidates(1920, 1080, 8, 9) + bench.temporal_candidates(1920, 1080, 8, 2, 2, 4))
    assert abs(ssd / 1e9 - 6.28) < 0.05
    assert bench.CNN_FLOP_PER_PX == 2 * 9 * (2 * 32 + 4 * 32 * 32 + 32)
    r = bench.step_roofline(157.3e12 * 1e-3, 0.0, 2e-3)
    assert abs(r["t_min_ms"] - 1.0) < 1e-9 and abs(r["frac"] - 0.5) < 1e-9
    r = bench.step_roofline(0.0, 2500e12 / 3 * 1e-3, 1e-3, split_bf16=True)
    assert abs(r["frac"] - 1.0) < 1e-9
