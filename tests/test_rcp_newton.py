"""The bilinear sampler takes 1/w of the texture width and height (sampler.h
step of one texel) from the hardware reciprocal refined by one fused Newton
step instead of a full IEEE division (pf_device.hip, rcp_int_rn).  This
checks, in exact rational arithmetic, that for every integer side 1..65535
and every starting value within 1 ulp of the correctly rounded reciprocal
(the hardware reciprocal's documented accuracy), the refined value equals
the correctly rounded 1/w: the result is then the same as __fdiv_rn(1, w).
"""
import math
import struct
from fractions import Fraction as F


def _f32(x):
    """x (a Fraction) rounded to the nearest float32, ties to even."""
    if x == 0:
        return 0.0
    s = -1 if x < 0 else 1
    x = abs(x)
    e = math.floor(math.log2(x))
    while F(2) ** e > x:
        e -= 1
    while F(2) ** (e + 1) <= x:
        e += 1
    ulp = F(2) ** (e - 23)
    q = x / ulp
    n = math.floor(q)
    r = q - n
    if r > F(1, 2) or (r == F(1, 2) and n % 2 == 1):
        n += 1
    return s * float(n * ulp)


def _step(x, d):
    b = struct.unpack("<I", struct.pack("<f", x))[0] + d
    return struct.unpack("<f", struct.pack("<I", b))[0]


def test_rounding_helper_matches_float32():
    for v in (F(1, 3), F(2, 3), F(1, 65535), F(-7, 9), F(10 ** 6, 7)):
        assert _f32(v) == struct.unpack("<f", struct.pack("<f", float(v)))[0] or abs(float(v)) > 1e30


def test_newton_reciprocal_is_correctly_rounded_for_every_texture_side():
    bad = []
    for w in range(1, 65536):
        exact = _f32(F(1, w))
        for d in (-1, 0, 1):
            r0 = _step(exact, d)
            e = _f32(F(-w) * F(r0) + 1)              # fma(-w, r0, 1)
            r1 = _f32(F(e) * F(r0) + F(r0))          # fma(e, r0, r0)
            if r1 != exact:
                bad.append((w, d))
    assert not bad, bad[:10]
