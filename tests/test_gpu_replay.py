"""Replays of captured jobs: every replay of a captured graph does the same
work.  Each replay starts from the targets as they were at capture time
(restored on the host and uploaded), so each one must equal the oracle,
not merely the previous replay.  Covers the config-2 feature set at fixture
size, scenes that take either clip kernel, the full 1080p benchmark frame
and the 256-context benchmark batch,
plus replays that run back to back without a restore, as the benchmark's
timed loop does."""
import numpy as np
import pytest

import bench
import pixelforge_amd as pf
import scenes
from conftest import render_oracle

pytestmark = pytest.mark.gpu

REPLAYS = 3


def _replays(cap, n=REPLAYS):
    lib = pf.lib()
    snap = cap.snapshot()
    out = []
    for k in range(1 + n):                              # sizing run, then n graph replays
        cap.restore(snap)
        assert lib.pfdJobLaunch(cap.job) == 0
        assert lib.pfdJobSync(cap.job) == 0
        assert lib.pfdJobDownload(cap.job) == 0
        out.append(cap.images())
    cap.restore(snap)
    return out


# clip_load_20k / clip_load_60k: the lane-group clipper on the main stream
# and the per-lane clipper on the side stream (fork and join inside the
# captured graph).
@pytest.mark.parametrize("name", ["c2_heightfield_bilinear", "c3_spheres_phong", "clip_3d", "odd_clear_pixel0_3d",
                                  "clip_load_20k", "clip_load_60k"])
def test_fixture_scene_replays(name):
    p = scenes.params(dict(scenes.ALL_SCENES)[name])
    ref = render_oracle(p)[0]
    cap = pf.capture(p)
    try:
        for k, imgs in enumerate(_replays(cap)):
            assert np.array_equal(imgs[0].reshape(ref.shape), ref), f"{name}: run {k} differs from the oracle"
    finally:
        cap.close()


def test_benchmark_frame_replays():
    """The 1080p, 200k-triangle benchmark frame: every replay equals the
    oracle; back-to-back replays without a restore (the timed loop) give
    the same image as the first."""
    p = bench.workload()
    ref = render_oracle(p)[0].ravel()
    cap = pf.capture(p)
    lib = pf.lib()
    try:
        for k, imgs in enumerate(_replays(cap)):
            print(f"1080p run {k}", flush=True)
            assert np.array_equal(imgs[0], ref), f"run {k}: {np.count_nonzero(imgs[0] != ref)} pixels differ"
        for _ in range(5):
            assert lib.pfdJobLaunch(cap.job) == 0
        assert lib.pfdJobSync(cap.job) == 0
        assert lib.pfdJobDownload(cap.job) == 0
        assert np.array_equal(cap.images()[0], ref), "back-to-back replays changed the frame"
    finally:
        cap.close()


def test_benchmark_batch_replays():
    """The 256-context benchmark batch: every replay equals the first run,
    and every eighth context equals the oracle."""
    params = pf.batch_workload(256)
    cap = pf.capture_batch(params)
    try:
        runs = _replays(cap)
    finally:
        cap.close()
    for k in range(1, len(runs)):
        for i in range(len(params)):
            assert np.array_equal(runs[k][i], runs[0][i]), f"run {k}: context {i} differs from the first run"
    for i in range(0, len(params), 8):
        ref = render_oracle(params[i])[0].ravel()
        assert np.array_equal(runs[-1][i], ref), f"context {i} differs from the oracle"
