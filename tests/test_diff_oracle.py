"""The FP64 oracle of the differentiable path against FP64 central finite
differences of its own forward (h = 1e-4), on the CPU.

The loss is L = sum (image - target)^2 with the target rendered from a
texture perturbed by N(0, 0.1); the analytic gradient is the oracle's
adjoint with dL/dimage = 2 (image - target).  The image is piecewise linear
in every parameter (clamps aside), so the central differences are exact up
to rounding and the tolerance is tight.  Scenes cover bilinear and nearest
filtering, attenuation and spot lights, clipped triangles (the mesh extends
past the side planes), unlit and untextured frames.
"""
import numpy as np
import pytest

import pixelforge_amd as pf
from diffscene import Scene, oracle_job, params_of, target_of

H_FD = 1e-4

CASES = {
    "phong_bilinear": dict(),
    "phong_nearest_clamp": dict(filt=pf.PF_NEAREST, wrap=pf.PF_CLAMP_TO_EDGE, seed=2),
    "phong_att_spot_mirror": dict(attenuation=True, spot=True, wrap=pf.PF_MIRRORED_REPEAT, uv_scale=2.5, seed=3),
    "unlit_textured": dict(lit=False, seed=4),
    "lit_untextured": dict(textured=False, seed=5),
}


def _loss(job, target, *p):
    img = job.forward(*p)
    return float(((img - target) ** 2).sum()), img


def _check(name, fd, an, scale):
    err = np.abs(fd - an)
    bad = err > 1e-7 * scale + 1e-6 * np.abs(an)
    assert not bad.any(), f"{name}: fd {fd[bad][:5]} vs adjoint {an[bad][:5]}"


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_matches_central_differences(case):
    s = Scene(**CASES[case])
    try:
        job = oracle_job(s)
        p = list(params_of(s))
        target = target_of(s, job, color_sigma=0.0 if s.textured else 0.05)
        L0, img = _loss(job, target, *p)
        assert L0 > 0
        g = job.backward(*p, 2.0 * (img - target))
        rng = np.random.default_rng(7)
        scale = max(1.0, L0)
        for k, name in enumerate(("texels", "colors", "ldiff", "lspec")):
            if (name in ("ldiff", "lspec") and not s.lit) or (name == "texels" and not s.textured):
                assert not g[k].any()
                continue
            flat = g[k].ravel()
            nz = np.flatnonzero(flat)
            if name in ("texels", "colors"):
                assert len(nz) > 0, f"no gradient reaches the {name}"
                pick = np.concatenate([nz[np.argsort(-np.abs(flat[nz]))[:8]], rng.choice(flat.size, 8, replace=False)])
            else:
                pick = np.array([lid * 3 + c for lid in s.light_ids for c in range(3)])
            fd = []
            for i in pick:
                q = [a.copy() for a in p]
                q[k].ravel()[i] += H_FD
                Lp, _ = _loss(job, target, *q)
                q[k].ravel()[i] -= 2 * H_FD
                Lm, _ = _loss(job, target, *q)
                fd.append((Lp - Lm) / (2 * H_FD))
            _check(f"{case}/{name}", np.array(fd), flat[pick], scale)
    finally:
        s.close()


def test_coverage_is_the_integer_render():
    """Pixels without a winner keep the integer pass's RGBA8 (the clear
    colour, except pixels 0..7 that pfClear leaves alone: context.c:693-713)
    and the float forward returns them / 255."""
    s = Scene(seed=8)
    try:
        job = oracle_job(s)
        win = job.winners()
        col = job.color()
        assert (win >= 0).sum() > 0.5 * win.size
        bg = np.flatnonzero(win < 0)
        assert len(bg) > 0
        clear = 10 | 20 << 8 | 30 << 16 | 255 << 24
        assert np.all(col[bg[bg >= 8]] == clear)
        img = job.forward(*params_of(s)).reshape(-1, 4)
        assert np.array_equal(img[bg], pf.unpack(col[bg]).astype(np.float64) / 255.0)
    finally:
        s.close()


def test_alpha_and_clamp_have_no_gradient():
    """Phong writes alpha 0 (lighting.c:244-256): no gradient flows from the
    alpha channel; saturated channels pass none either."""
    s = Scene(seed=9)
    try:
        job = oracle_job(s)
        p = params_of(s)
        d = np.zeros((s.H, s.W, 4))
        d[..., 3] = 1.0
        g = job.backward(*p, d)
        assert all(not x.any() for x in g)
        # Saturate every channel: diffuse 20x -> acc > 1 nearly everywhere.
        t, c, ld, ls = p
        img = job.forward(t, c, ld * 20, ls * 20)
        sat = (img[..., :3] >= 1.0)
        d = np.zeros((s.H, s.W, 4))
        d[..., :3] = sat
        g = job.backward(t, c, ld * 20, ls * 20, d)
        assert all(not x.any() for x in g)
    finally:
        s.close()
