"""Writes tests/golden/target_access.npz: the host-side pixel accessors of
render targets other than RGBA8, as the CPU build of the reference
(oracle/_ref, `make -C oracle ref`) performs them.

For each render-target pair (RGB/BGR bytes, RGB/BGR 5_6_5, RGBA/BGRA
5_5_5_1 and 4_4_4_4, and luminance-alpha, RGB, BGR, RGBA and BGRA in half
and float components), on 16x16 targets:
  - a context's main buffer of that pair filled with fixed random bytes,
    read back with pfReadPixels as RGBA / UNSIGNED_BYTE (the pair's scalar
    getter, pixel.h:549-707) and as RGB / UNSIGNED_SHORT_5_6_5;
  - pfPostProcess over that buffer with a function of (x, y, colour)
    (getter, then setter, context.c:2397-2441);
  - a framebuffer of that pair from pfGenFramebuffer: pfClearFramebuffer,
    pfSetFramebufferPixel / ...Depth / ...DepthTest at fixed pixels with
    colours whose alpha crosses the 5_5_5_1 threshold (pixel.h:206-400
    setters), and pfGetFramebufferPixel of every pixel (framebuffer.c:86-148).

    python tests/golden/make_target_access.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")

TARGETS = [(6, 0), (8, 0), (6, 2), (8, 2), (7, 3), (9, 3), (7, 4), (9, 4),
           (5, 9), (5, 10), (6, 9), (8, 9), (6, 10), (8, 10), (7, 9), (9, 9), (7, 10), (9, 10)]
W = H = 16
# (x, y, RGBA8 word, depth) of the pixel writes, in order.
WRITES = [(0, 0, 0x31FF8040, 0.5), (5, 3, 0x32102030, 0.25), (15, 15, 0x00FFFFFF, 0.75),
          (7, 9, 0x33C0A080, 0.125), (8, 1, 0xFF0A141E, 0.9), (2, 14, 0x807F7F7F, 0.3)]


def comps(f):
    return 2 if f == 5 else 3 if f in (6, 8) else 4


def bpp(f, t):
    if t in (9, 10):
        return comps(f) * (2 if t == 9 else 4)
    return 2 if t else 3 if f in (6, 8) else 4


# Component values of the half and float buffers besides uniform ones in
# [-0.5, 1.5): exact ends, values just under a byte step, a half subnormal
# (flushed to zero when read), infinities and a NaN (read as byte 0).
SPECIAL = [0.0, -0.0, 1.0, 254.99 / 255, 1 / 255, 6e-8, 0.5, float("inf"), float("-inf"), float("nan")]


def buffer_bytes(f, t, seed):
    rng = np.random.default_rng(seed)
    if t in (9, 10):
        v = rng.uniform(-0.5, 1.5, W * H * comps(f)).astype(np.float32)
        v[rng.permutation(v.size)[: 4 * len(SPECIAL)]] = np.repeat(np.array(SPECIAL, np.float32), 4)
        return v.astype(np.float16 if t == 9 else np.float32).view(np.uint8)
    return rng.integers(0, 256, W * H * bpp(f, t), dtype=np.uint8)


class FB(C.Structure):
    _fields_ = [("texture", C.c_void_p), ("zbuffer", C.c_void_p)]


# PFcolor (four bytes) is passed and returned as one 32-bit integer on
# x86-64 (SysV: a 4-byte aggregate is INTEGER class).
POST = C.CFUNCTYPE(C.c_uint32, C.c_int32, C.c_int32, C.c_float, C.c_uint32)


def _post(x, y, z, c):
    r, g, b, a = c & 255, (c >> 8) & 255, (c >> 16) & 255, c >> 24
    return ((r + 17 * x) & 255) | (((g ^ (9 * y)) & 255) << 8) | ((255 - b) << 16) | (((a + x * y) & 255) << 24)


_post_cb = POST(_post)


def access_all(lib_path):
    """{name: uint8 array} of every pair's results from the library."""
    L = C.CDLL(lib_path)
    L.pfCreateContext.restype = C.c_void_p
    L.pfCreateContext.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.pfMakeCurrent.argtypes = [C.c_void_p]
    L.pfDeleteContext.argtypes = [C.c_void_p]
    L.pfReadPixels.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int, C.c_int, C.c_void_p]
    L.pfGetError.restype = C.c_int
    L.pfPostProcess.argtypes = [POST]
    L.pfGenFramebuffer.restype = FB
    L.pfGenFramebuffer.argtypes = [C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.pfDeleteFramebuffer.argtypes = [C.POINTER(FB)]
    L.pfClearFramebuffer.argtypes = [C.POINTER(FB), C.c_uint32, C.c_float]
    L.pfSetFramebufferPixel.argtypes = [C.POINTER(FB), C.c_int32, C.c_int32, C.c_uint32]
    L.pfSetFramebufferPixelDepth.argtypes = [C.POINTER(FB), C.c_int32, C.c_int32, C.c_float, C.c_uint32]
    L.pfSetFramebufferPixelDepthTest.argtypes = [C.POINTER(FB), C.c_int32, C.c_int32, C.c_float, C.c_uint32, C.c_int]
    L.pfGetFramebufferPixel.restype = C.c_uint32
    L.pfGetFramebufferPixel.argtypes = [C.POINTER(FB), C.c_int32, C.c_int32]
    L.pfGetTexturePixels.restype = C.c_void_p
    L.pfGetTexturePixels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    out = {}
    for k, (f, t) in enumerate(TARGETS):
        n = W * H * bpp(f, t)
        buf = np.zeros(n + 64, np.uint8)
        buf[:n] = buffer_bytes(f, t, 100 + k)
        ctx = L.pfCreateContext(buf.ctypes.data, W, H, f, t)
        assert ctx, (f, t)
        L.pfMakeCurrent(ctx)
        try:
            rgba = np.zeros(W * H * 4, np.uint8)
            L.pfReadPixels(0, 0, W, H, 7, 0, rgba.ctypes.data)
            rgb565 = np.zeros(W * H * 2, np.uint8)
            L.pfReadPixels(0, 0, W, H, 6, 2, rgb565.ctypes.data)
            L.pfPostProcess(_post_cb)
            post = buf[:n].copy()
            fb = L.pfGenFramebuffer(W, H, f, t)
            assert fb.texture, (f, t)
            L.pfClearFramebuffer(C.byref(fb), 0x80C06030, 0.6)
            for i, (x, y, c, z) in enumerate(WRITES):
                if i % 3 == 0:
                    L.pfSetFramebufferPixel(C.byref(fb), x, y, c)
                elif i % 3 == 1:
                    L.pfSetFramebufferPixelDepth(C.byref(fb), x, y, z, c)
                else:
                    L.pfSetFramebufferPixelDepthTest(C.byref(fb), x, y, z, c, 2)     # PF_LESS
            got = np.array([L.pfGetFramebufferPixel(C.byref(fb), x, y) for y in range(H) for x in range(W)],
                           np.uint32)
            px = C.cast(L.pfGetTexturePixels(fb.texture, None, None, None, None), C.POINTER(C.c_uint8))
            fbbytes = np.ctypeslib.as_array(px, (n,)).copy()
            L.pfDeleteFramebuffer(C.byref(fb))
            assert L.pfGetError() == 0, (f, t)
        finally:
            L.pfDeleteContext(ctx)
            L.pfMakeCurrent(None)
        key = f"{f}_{t}"
        out[f"read_rgba_{key}"] = rgba
        out[f"read_565_{key}"] = rgb565
        out[f"post_{key}"] = post
        out[f"fb_{key}"] = fbbytes
        out[f"get_{key}"] = got.view(np.uint8)
    return out


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f"{REF_SO} missing: run make -C oracle ref")
    res = access_all(REF_SO)
    np.savez_compressed(os.path.join(HERE, "target_access.npz"), **res)
    print(len(res), "arrays")


if __name__ == "__main__":
    main()
