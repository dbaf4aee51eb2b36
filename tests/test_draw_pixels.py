"""pfDrawPixels (context.c:1987-2080): argument errors, and the recorded
frame carrying the call as a command between the draws around it.  The
pixels themselves are checked against reference goldens (scenes pixels_*,
test_oracle_golden.py on the oracle, test_gpu_parity.py on the GPU)."""
import ctypes as C

import numpy as np

from test_host_api import FrameLog, _ctx

INVALID_ENUM, INVALID_VALUE = 1, 2
PF_CMD_PIXELS = 3


def _bind(L):
    L.pfDrawPixels.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_void_p]
    L.pfDrawPixels.restype = None
    return L


def test_draw_pixels_argument_errors():
    L, ctx, _ = _ctx()
    _bind(L)
    img = np.zeros(64, np.uint32)
    try:
        L.pfDrawPixels(0, 4, 7, 0, img.ctypes.data)
        assert L.pfGetError() == INVALID_VALUE          # context.c:1990-1993
        L.pfDrawPixels(4, 4, 7, 1, img.ctypes.data)     # RGBA / UNSIGNED_SHORT: no getter
        assert L.pfGetError() == INVALID_ENUM
        L.pfDrawPixels(4, 4, 7, 0, None)                # the reference reads NULL
        assert L.pfGetError() == INVALID_VALUE
    finally:
        L.pfDeleteContext(ctx)


def test_draw_pixels_is_recorded_between_draws():
    L, ctx, _ = _ctx()
    _bind(L)
    log = FrameLog()
    L.pfxSetExecutor(C.cast(log.cb, C.c_void_p), None)
    img = np.arange(12, dtype=np.uint8)
    try:
        L.pfBegin(2)
        for v in [(-1, -1), (1, -1), (0, 1)]:
            L.pfVertex2f(C.c_float(v[0]), C.c_float(v[1]))
        L.pfEnd()
        L.pfDrawPixels(2, 2, 6, 0, img.ctypes.data)     # RGB / UNSIGNED_BYTE
        img[:] = 0                                      # the call copied the image
        L.pfBegin(2)
        for v in [(-1, -1), (1, -1), (0, 1)]:
            L.pfVertex2f(C.c_float(v[0]), C.c_float(v[1]))
        L.pfEnd()
        L.pfFinish()
        assert L.pfGetError() == 0
    finally:
        L.pfxSetExecutor(None, None)
        L.pfDeleteContext(ctx)
    f = log.frames[0]
    assert [k for k, _, _ in f["cmds"]] == [2, PF_CMD_PIXELS, 2]
    assert f["cmds"][1][1] == 0
