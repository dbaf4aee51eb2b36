"""Writes tests/golden/readback.npz: pfReadPixels of one 16x16 RGBA8
framebuffer into every (format, type) pair with a pixel setter, as the CPU
build of the reference (oracle/_ref, `make -C oracle ref`) returns it.

The framebuffer is the context's main buffer, filled directly: every byte
value appears in every channel, alpha also takes 49, 50 and 51 (the
5_5_5_1 alpha threshold, config.h:45-47).  Two reads are stored per pair:
the whole buffer, and a 7x5 rectangle at (12, -2) clipped by the buffer's
edges (context.c:2371-2374).

    python tests/golden/make_readback.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")

PAIRS = ([(f, t) for f in range(10) for t in (0, 9, 10)] + [(6, 2), (8, 2)]
         + [(f, t) for f in (7, 9) for t in (3, 4)])
W = H = 16
RECT = (12, -2, 7, 5)


def framebuffer():
    rng = np.random.default_rng(7)
    px = np.zeros((W * H, 4), np.uint8)
    for c in range(4):
        px[:, c] = rng.permutation(256).astype(np.uint8)
    px[:6, 3] = [49, 50, 51, 0, 255, 128]
    return px.view(np.uint32).reshape(-1)


def texel_bytes(f, t):
    comps = 1 if f <= 4 else 2 if f == 5 else 3 if f in (6, 8) else 4
    if t in (2, 3, 4):
        return 2
    return comps * {0: 1, 9: 2, 10: 4}[t]


def read_all(lib_path):
    """{(f, t): (whole, rect)} byte arrays from the library at lib_path."""
    L = C.CDLL(lib_path)
    L.pfCreateContext.restype = C.c_void_p
    L.pfCreateContext.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.pfMakeCurrent.argtypes = [C.c_void_p]
    L.pfDeleteContext.argtypes = [C.c_void_p]
    L.pfReadPixels.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int, C.c_int, C.c_void_p]
    L.pfGetError.restype = C.c_int
    buf = np.zeros(W * H + 16, np.uint32)
    buf[:W * H] = framebuffer()
    ctx = L.pfCreateContext(buf.ctypes.data, W, H, 7, 0)
    assert ctx
    L.pfMakeCurrent(ctx)
    out = {}
    try:
        for f, t in PAIRS:
            bpp = texel_bytes(f, t)
            whole = np.full(W * H * bpp + 64, 0xA5, np.uint8)
            L.pfReadPixels(0, 0, W, H, f, t, whole.ctypes.data)
            rx, ry, rw, rh = RECT
            rect = np.full(rw * rh * bpp + 64, 0xA5, np.uint8)
            L.pfReadPixels(rx, ry, rw, rh, f, t, rect.ctypes.data)
            assert L.pfGetError() == 0, (f, t)
            out[(f, t)] = (whole, rect)
    finally:
        L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)
    return out


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f"{REF_SO} missing: run make -C oracle ref")
    res = read_all(REF_SO)
    arrays = {"framebuffer": framebuffer()}
    for (f, t), (whole, rect) in res.items():
        arrays[f"whole_{f}_{t}"] = whole
        arrays[f"rect_{f}_{t}"] = rect
    np.savez_compressed(os.path.join(HERE, "readback.npz"), **arrays)
    print(len(res), "pairs")


if __name__ == "__main__":
    main()
