"""pfdExecuteFrameNoAlloc, the entry point that allocates no device memory,
against pfdExecuteFrame (the one-shot entry point it is a variant of).

Both run the same recorded frames through the host library's executor hook
(a ctypes callback that calls the C ABI).  Two different frames go through
the no-allocation entry point one after the other, the second smaller than
the first, so the second call runs in the static array the first one left
full; colour and depth must equal pfdExecuteFrame's bit for bit.  A frame
larger than the array is refused with an error, and the next call works.
"""
import ctypes as C

import numpy as np
import pytest

import pixelforge_amd as pf
from conftest import assert_same, render_gpu

pytestmark = pytest.mark.gpu

PFD_E_NOMEM = -2


def render_no_alloc(params):
    """Scene -> (color, depth, err, rc of the entry point)."""
    L = pf.lib()
    rcs = []

    def run(frame, user):
        rc = L.pfdExecuteFrameNoAlloc(frame)
        rcs.append(rc)
        return rc

    cb = pf.EXECUTOR(run)
    L.pfxSetExecutor(C.cast(cb, C.c_void_p), None)
    try:
        c, d, e = pf.render(params, pf.LIB_PATH)
    finally:
        L.pfxSetExecutor(None, None)
    return c, d, e, rcs


LARGE = pf.scene(kind=pf.HEIGHTFIELD, width=640, height=360, gridX=200, gridZ=40, seed=11, texture=1,
                 texW=256, texH=256, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR, depthFunc=pf.PF_LEQUAL,
                 cull=pf.PF_BACK, blend=1, blendMode=pf.PF_BLEND_ALPHA, blendFraction=0.3)
SMALL = pf.scene(kind=pf.RANDOM_3D, width=97, height=61, ntris=150, seed=5, texture=1, texW=33, texH=17,
                 wrap=pf.PF_MIRRORED_REPEAT, filter=pf.PF_NEAREST, blend=1, blendMode=pf.PF_BLEND_ADD)


def test_no_alloc_matches_execute_frame_twice():
    expected = [render_gpu(p) for p in (LARGE, SMALL)]
    for p, (ec, ed, ee), what in zip((LARGE, SMALL), expected, ("first call (640x360)", "second call (97x61)")):
        print(what, flush=True)
        assert ee == 0
        c, d, e, rcs = render_no_alloc(p)
        assert rcs and all(rc == 0 for rc in rcs), f"{what}: {rcs} {pf.lib().pfdLastError()}"
        assert e == 0
        assert np.count_nonzero(ec != ec[0, 0]) > 100, "scene draws nothing"
        assert_same(ec, ed, c, d, what)


def test_no_alloc_refuses_frames_beyond_the_array():
    big = pf.scene(kind=pf.RANDOM_NDC, width=6000, height=6000, ntris=10, seed=3)
    print("oversized frame", flush=True)
    _, _, e, rcs = render_no_alloc(big)
    assert rcs == [PFD_E_NOMEM]
    assert b"static scratch" in pf.lib().pfdLastError()
    assert e != 0
    print("next call", flush=True)
    ec, ed, _ = render_gpu(SMALL)
    c, d, e, rcs = render_no_alloc(SMALL)
    assert rcs == [0] and e == 0
    assert_same(ec, ed, c, d, "call after a refused frame")
