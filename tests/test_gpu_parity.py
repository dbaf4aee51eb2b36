"""HIP path against the CPU oracle: same recorded frame, same pixels.

Each scene is recorded by the host library and executed twice: by the HIP
kernels (the default executor) and by the oracle.  Colour and depth must be
bit-identical.  Where a reference golden exists the GPU frame must also
equal it.
"""
import ctypes as C
import os

import numpy as np
import pytest

import pixelforge_amd as pf
import scenes
from conftest import GOLDEN, assert_same, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kw", scenes.ALL_SCENES, ids=[n for n, _ in scenes.ALL_SCENES])
def test_gpu_matches_oracle(name, kw):
    print(f"scene {name}", flush=True)
    p = scenes.params(kw)
    oc, od, oe, _ = render_oracle(p)
    gc, gd, ge = render_gpu(p)
    assert ge == 0, f"GPU path reported error {ge}: {pf.lib().pfdLastError()}"
    assert_same(oc, od, gc, gd, name)
    path = os.path.join(GOLDEN, f"{name}.npz")
    if os.path.exists(path):
        g = np.load(path, allow_pickle=False)
        assert_same(g["color"], g["depth"], gc, gd, name + " vs golden")


CONFIG_SWEEP = ["c1_ndc_less", "c2_heightfield_nearest", "c3_spheres_phong", "blend_1", "clip_3d",
                "odd_size_clear", "multi_frame", "bilinear_0", "target_rgb_ubyte_raster", "target_bgra_u4444_raster",
                "target_lumalpha_half_raster", "target_bgr_float_raster", "target_bgra_half_raster"]


def test_clip_load_scenes_take_both_clip_kernels():
    """The two clip-load scenes clip fewer and more than 16k triangles, so
    the oracle comparisons above ran both clip kernels."""
    counts = []
    for name, kw in scenes.CLIP_LOAD_SCENES:
        cap = pf.capture(scenes.params(kw))
        try:
            L = pf.lib()
            assert L.pfdJobUploadTargets(cap.job) == 0
            assert L.pfdJobLaunch(cap.job) == 0 and L.pfdJobSync(cap.job) == 0
            st = pf.PFDstats()
            assert L.pfdJobStats(cap.job, C.byref(st)) == 0
            counts.append(int(st.clippedTriangles))
        finally:
            cap.close()
    print("clipped", counts, flush=True)
    assert 4096 < counts[0] <= 16384 < counts[1]


@pytest.mark.parametrize("tile_shift", [3, 4, 5, 6])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_every_tile_size_and_raster_mode(tile_shift, mode):
    """Tile size and raster mode are performance knobs only: every
    configuration renders the oracle's frame bit for bit."""
    table = dict(scenes.ALL_SCENES)
    try:
        pf.set_default_config(tile_shift, mode)
        for name in CONFIG_SWEEP:
            print(f"config {tile_shift}/{mode} scene {name}", flush=True)
            p = scenes.params(table[name])
            oc, od, _, _ = render_oracle(p)
            gc, gd, ge = render_gpu(p)
            assert ge == 0
            assert_same(oc, od, gc, gd, f"{name} tile {1 << tile_shift} mode {mode}")
    finally:
        pf.set_default_config()
