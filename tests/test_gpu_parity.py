"""HIP path against the CPU oracle: same recorded frame, same pixels.

Each scene is recorded by the host library and executed twice: by the HIP
kernels (the default executor) and by the oracle.  Colour and depth must be
bit-identical.  Where a reference golden exists the GPU frame must also
equal it.
"""
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
