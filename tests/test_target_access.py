"""Host-side pixel access of render targets other than RGBA8 (pfReadPixels,
pfPostProcess, pfClearFramebuffer, pfSetFramebufferPixel*,
pfGetFramebufferPixel: the scalar getters and setters, pixel.h:206-707)
against the bytes the compiled reference produced for the same calls
(tests/golden/target_access.npz, written by
tests/golden/make_target_access.py), plus closed forms of the accessors."""
import os
import sys

import numpy as np
import pytest

import pixelforge_amd as pf

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_target_access as mt  # noqa: E402

FIXTURE = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "target_access.npz"))
OURS = None


def _ours():
    global OURS
    if OURS is None:
        OURS = mt.access_all(pf.LIB_PATH)
    return OURS


@pytest.mark.parametrize("pair", mt.TARGETS, ids=[f"{f}_{t}" for f, t in mt.TARGETS])
def test_target_access_matches_reference(pair):
    key = f"{pair[0]}_{pair[1]}"
    ours = _ours()
    for what in ("read_rgba", "read_565", "post", "fb", "get"):
        assert np.array_equal(ours[f"{what}_{key}"], FIXTURE[f"{what}_{key}"]), what


def test_target_getters_closed_forms():
    """Independent of the fixture: RGB / BGR bytes read back in channel
    order with alpha 255; 5_6_5 fields scale by 255/31 and 255/63 in float
    and truncate, so 31 -> 255 and 1 -> 8; 4_4_4_4 nibbles scale by 17."""
    ours = _ours()
    for k, (f, t) in enumerate(mt.TARGETS):
        src = mt.buffer_bytes(f, t, 100 + k)
        rgba = ours[f"read_rgba_{f}_{t}"].reshape(-1, 4)
        assert (rgba[:, 3] == 255).all() or t in (3, 4) or f in (5, 7, 9)
        if t in (9, 10):
            # (PFubyte)(v * 255) as x86-64 converts: truncation toward zero
            # to a 32-bit integer, low byte kept; infinities and NaN give the
            # integer 0x80000000, so byte 0.  Half subnormals read as zero.
            v = src.view(np.float16 if t == 9 else np.float32).astype(np.float32)
            if t == 9:
                v[np.abs(v) < np.float32(2.0 ** -14)] = 0.0
            with np.errstate(invalid="ignore", over="ignore"):
                p = v * np.float32(255)
                byte = np.where(np.isfinite(p), np.trunc(np.nan_to_num(p)), 0).astype(np.int64) & 255
            c = list(byte.reshape(-1, mt.comps(f)).T) + [255]
            want = {5: (0, 0, 0, 1), 6: (0, 1, 2, -1), 8: (2, 1, 0, -1), 7: (0, 1, 2, 3), 9: (2, 1, 0, 3)}[f]
            want = [c[j] for j in want]
            for ch in range(4):
                assert np.array_equal(rgba[:, ch], np.broadcast_to(want[ch], rgba[:, ch].shape)), (f, t, ch)
        elif t == 0:
            px = src.reshape(-1, 3)
            want = px if f == 6 else px[:, ::-1]
            assert np.array_equal(rgba[:, :3], want)
        elif t == 2:
            w = src.view(np.uint16).astype(np.uint32)
            hi = np.floor(((w >> 11) & 31).astype(np.float32) * np.float32(255.0 / 31)).astype(np.uint8)
            g = np.floor(((w >> 5) & 63).astype(np.float32) * np.float32(255.0 / 63)).astype(np.uint8)
            lo = np.floor((w & 31).astype(np.float32) * np.float32(255.0 / 31)).astype(np.uint8)
            r, b = (hi, lo) if f == 6 else (lo, hi)
            assert np.array_equal(rgba[:, 0], r) and np.array_equal(rgba[:, 1], g) and np.array_equal(rgba[:, 2], b)
        elif t == 4:
            w = src.view(np.uint16).astype(np.uint32)
            n = [((w >> s) & 15) * 17 for s in (12, 8, 4, 0)]
            r, b = (n[0], n[2]) if f == 7 else (n[2], n[0])
            assert np.array_equal(rgba[:, 0], r) and np.array_equal(rgba[:, 2], b)
            assert np.array_equal(rgba[:, 1], n[1]) and np.array_equal(rgba[:, 3], n[3])


def test_target_setters_closed_forms():
    """pfClearFramebuffer's colour (r 0x30, g 0x60, b 0xC0, a 0x80) through
    each pair's scalar setter, at a pixel no later write touches: roundf of
    c / 255 times the field maximum; the BGRA 5_5_5_1 / 4_4_4_4 setters
    keep RGBA order; the 5_5_5_1 alpha bit is a > 50."""
    ours = _ours()
    r, g, b, a = 0x30, 0x60, 0xC0, 0x80
    rnd = lambda c, m: int(np.floor(np.float32(c) * np.float32(1 / 255) * np.float32(m) + np.float32(0.5)))
    want = {
        (6, 0): bytes([r, g, b]), (8, 0): bytes([b, g, r]),
        (6, 2): (rnd(r, 31) << 11 | rnd(g, 63) << 5 | rnd(b, 31)).to_bytes(2, "little"),
        (8, 2): (rnd(b, 31) << 11 | rnd(g, 63) << 5 | rnd(r, 31)).to_bytes(2, "little"),
        (7, 3): (rnd(r, 31) << 11 | rnd(g, 31) << 6 | rnd(b, 31) << 1 | 1).to_bytes(2, "little"),
        (7, 4): (rnd(r, 15) << 12 | rnd(g, 15) << 8 | rnd(b, 15) << 4 | rnd(a, 15)).to_bytes(2, "little"),
    }
    want[(9, 3)], want[(9, 4)] = want[(7, 3)], want[(7, 4)]
    # Half and float pairs: the channels times (float)(1/255); luminance
    # 0.299 r + 0.587 g + 0.114 b of those in float; BGR reversed, BGRA in
    # RGBA order.  Halves round to nearest (none of these values is a tie,
    # where the reference's conversion rounds up).
    inv = np.float32(1 / 255)
    n = {ch: np.float32(v) * inv for ch, v in zip("rgba", (r, g, b, a))}
    grey = n["r"] * np.float32(0.299) + n["g"] * np.float32(0.587) + n["b"] * np.float32(0.114)
    order = {5: [grey, n["a"]], 6: [n["r"], n["g"], n["b"]], 8: [n["b"], n["g"], n["r"]],
             7: [n["r"], n["g"], n["b"], n["a"]], 9: [n["r"], n["g"], n["b"], n["a"]]}
    for f, vals in order.items():
        v = np.array(vals, np.float32)
        assert not ((v.view(np.uint32) & 0x1FFF) == 0x1000).any()
        want[(f, 9)] = v.astype(np.float16).tobytes()
        want[(f, 10)] = v.tobytes()
    for (f, t), w in want.items():
        bpp = mt.bpp(f, t)
        px = ours[f"fb_{f}_{t}"][3 * bpp:4 * bpp].tobytes()      # pixel (3, 0)
        assert px == w, (f, t)
