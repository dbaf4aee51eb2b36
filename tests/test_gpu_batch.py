"""Batched contexts on the GPU: several contexts of uneven size, triangle
count and texture, half of them additively blended, rendered in the same
launches (pfSubmitBatch), must each equal their single-context render and
the oracle bit for bit, at every tile size and raster mode."""
import ctypes as C

import numpy as np
import pytest

import pixelforge_amd as pf
import scenes
from conftest import render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def test_batch_equals_single_context_renders():
    params = pf.batch_workload(24, seed=7, max_side=256, max_tris=3000, max_tex=128)
    imgs, err = pf.render_batch(params)
    assert err == 0
    for i, (p, img) in enumerate(zip(params, imgs)):
        single, _, e = render_gpu(p)
        assert e == 0
        assert np.array_equal(img, single), f"context {i} ({p.width}x{p.height}) differs from its single render"
        ref, _, e, frags = render_oracle(p)
        assert frags > 0
        assert np.array_equal(img, ref), f"context {i} differs from the oracle"


@pytest.mark.parametrize("tile_shift", [3, 4, 5, 6])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_batch_job_every_tile_size(tile_shift, mode):
    params = pf.batch_workload(10, seed=8, max_side=200, max_tris=2000, max_tex=96)
    ref = [render_oracle(p)[0] for p in params]
    cap = pf.capture_batch(params)
    try:
        runs = _run_capture(cap, tile_shift, mode, graph_runs=2)
        for k, imgs in enumerate(runs):
            for i, (p, r) in enumerate(zip(params, ref)):
                got = imgs[i].reshape(p.height, p.width)
                assert np.array_equal(got, r), f"run {k}: context {i} at tile {1 << tile_shift}, mode {mode}"
    finally:
        cap.close()


def _run_capture(cap, tile_shift, mode=0, graph_runs=1):
    """Sizing run, then `graph_runs` replays of the captured graph, each
    starting from the targets as they were when the job was captured
    (restored on the host and uploaded).  Returns the images of every run."""
    lib = pf.lib()
    cfg = pf.PFDconfig()
    cfg.tileShift, cfg.rasterMode = tile_shift, mode
    assert lib.pfdJobSetConfig(cap.job, C.byref(cfg)) == 0
    snap = cap.snapshot()
    runs = []
    for _ in range(1 + graph_runs):                     # sizing run, then the graph
        cap.restore(snap)
        assert lib.pfdJobLaunch(cap.job) == 0
        assert lib.pfdJobSync(cap.job) == 0
        assert lib.pfdJobDownload(cap.job) == 0
        runs.append(cap.images())
    cap.restore(snap)
    return runs


def test_benchmark_batch_wide_tile_keys():
    """The 256-context benchmark batch at 16x16 tiles has ~278k tiles, so
    its tile keys take 19 bits and the sort runs three 8-bit passes; at
    64x64 tiles the keys take 15 bits (two passes).  Tile
    size is a performance knob: both render the same images, and a sample
    of contexts equals the oracle."""
    params = pf.batch_workload(256)
    cap = pf.capture_batch(params)
    try:
        wide = _run_capture(cap, 4)[-1]
        narrow = _run_capture(cap, 6)[-1]
    finally:
        cap.close()
    for i, p in enumerate(params):
        n = p.width * p.height
        assert np.array_equal(wide[i][:n], narrow[i][:n]), f"context {i} differs between 16x16 and 64x64 tiles"
    for i in (0, 17, 101, 255):
        p = params[i]
        ref = render_oracle(p)[0]
        got = wide[i][: p.width * p.height].reshape(p.height, p.width)
        assert np.array_equal(got, ref), f"context {i} differs from the oracle"


def test_batch_with_tiny_and_clear_only_contexts():
    """Contexts smaller than a block and a context with nothing but a clear
    share the launches with ordinary ones."""
    kws = [dict(kind=pf.RANDOM_NDC, width=1, height=1, ntris=20, seed=94),
           dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=0, seed=95),
           dict(kind=pf.RANDOM_3D, width=200, height=150, ntris=500, seed=96, texture=1, texW=64, texH=64),
           dict(kind=pf.RANDOM_NDC, width=7, height=3, ntris=30, seed=97)]
    params = [pf.scene(**kw) for kw in kws]
    imgs, err = pf.render_batch(params)
    assert err == 0
    for i, (p, img) in enumerate(zip(params, imgs)):
        ref, _, e, _ = render_oracle(p)
        assert np.array_equal(img, ref), f"context {i} ({p.width}x{p.height}) differs from the oracle"


def test_multi_pass_frames_replay_in_graphs():
    """Frames with several passes (triangles, then points / lines / fog /
    pixel rectangles, then triangles again) and clipped 3D triangles,
    captured as one batch job and replayed as a graph several times: every
    replay, started from the targets as captured, equals the oracle (the
    clip-list counter is left at zero by each pass for the next one)."""
    names = ["points_discs_3d", "lines_3d", "fog_between_draws_3d", "pixels_3d_rgba_ubyte", "rects_3d",
             "odd_clear_pixel0_3d"]
    kws = dict(scenes.ALL_SCENES)
    params = [scenes.params(kws[n]) for n in names]
    refs = [render_oracle(p)[0] for p in params]
    cap = pf.capture_batch(params)
    lib = pf.lib()
    try:
        snap = cap.snapshot()
        for replay in range(4):
            print(f"multi-pass replay {replay}", flush=True)
            cap.restore(snap)
            assert lib.pfdJobLaunch(cap.job) == 0
            assert lib.pfdJobSync(cap.job) == 0
            assert lib.pfdJobDownload(cap.job) == 0
            for n, ref, img in zip(names, refs, cap.images()):
                got = img.reshape(ref.shape)
                assert np.array_equal(got, ref), f"replay {replay}: {n} differs from the oracle"
    finally:
        cap.close()
