"""pfPostProcess (context.c:2397-2440): the caller's function runs over
every pixel with its depth after the frame has been drawn; compared with
the compiled reference running the same function on the same scene (the
scene's triangles drawn by the oracle here)."""
import ctypes as C

import numpy as np
import pytest

import pixelforge_amd as pf
from conftest import REF_SO, OracleExecutor, have_ref


# PFcolor (4 bytes) is passed and returned in a 32-bit integer register on
# x86-64 (SysV ABI), which ctypes can express as uint32 (r in the low byte).
POSTFN = C.CFUNCTYPE(C.c_uint32, C.c_int32, C.c_int32, C.c_float, C.c_uint32)


@POSTFN
def _invert_by_depth(x, y, z, c):
    r, g, b, a = c & 255, (c >> 8) & 255, (c >> 16) & 255, c >> 24
    k = 0 if z > 1e30 else int(abs(z) * 40) & 255
    return ((255 - r) ^ (x & 15)) | (((g + y) & 255) << 8) | (((b ^ k) & 255) << 16) | (a << 24)


def _render(path, use_oracle):
    L = C.CDLL(path)
    L.pfCreateContext.restype = C.c_void_p
    L.pfCreateContext.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.pfMakeCurrent.argtypes = [C.c_void_p]
    L.pfDeleteContext.argtypes = [C.c_void_p]
    L.pfPostProcess.argtypes = [POSTFN]
    L.pfBegin.argtypes = [C.c_int]
    L.pfColor4ub.argtypes = [C.c_ubyte] * 4
    L.pfVertex3f.argtypes = [C.c_float] * 3
    L.pfGetError.restype = C.c_int
    W, H = 40, 30
    buf = np.zeros(W * H + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, W, H, 7, 0)
    L.pfMakeCurrent(ctx)
    rng = np.random.default_rng(3)
    try:
        L.pfEnable(0x0008)                              # PF_DEPTH_TEST
        L.pfClear(3)
        L.pfBegin(2)
        for _ in range(30):
            for _ in range(3):
                L.pfColor4ub(*[int(v) for v in rng.integers(0, 256, 4)])
                L.pfVertex3f(*[float(v) for v in rng.uniform(-1, 1, 3)])
        L.pfEnd()
        if use_oracle:
            with OracleExecutor():
                L.pfPostProcess(_invert_by_depth)
        else:
            L.pfPostProcess(_invert_by_depth)
        assert L.pfGetError() == 0
        return buf[:W * H].copy()
    finally:
        L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built")
def test_post_process_matches_reference():
    ours = _render(pf.LIB_PATH, True)
    ref = _render(REF_SO, False)
    assert np.count_nonzero(ours != ref) == 0
    assert len(np.unique(ours)) > 100


def test_post_process_rejects_null():
    L = pf.lib()
    buf = np.zeros(8 * 8 + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, 8, 8, 7, 0)
    L.pfMakeCurrent(ctx)
    try:
        L.pfPostProcess(None)                           # the reference calls NULL
        assert L.pfGetError() == 2                      # PF_INVALID_VALUE
    finally:
        L.pfDeleteContext(ctx)
