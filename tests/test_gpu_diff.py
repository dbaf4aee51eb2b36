"""Differentiable path on the GPU (pfDiffRender / pfDiffBackward) against
the FP64 oracle of the same recorded frame.

Bars: the RGBA8 image of the integer pass is bit-identical to the oracle's;
the float image agrees to 2e-5; every gradient has relative error <= 1e-3
against FP64 (||g - g64|| <= 1e-3 ||g64||, and per element within 1e-3 of
max |g64|).  A torch.autograd.Function wraps the C entry points through
ctypes (test harness only) and must give the same gradients.
"""
import numpy as np
import pytest

import pixelforge_amd as pf
from diffscene import Scene, oracle_job, params_of, target_of
from test_diff_oracle import CASES

pytestmark = pytest.mark.gpu

REL = 1e-3


def _rel_ok(name, g, g64):
    g64 = np.asarray(g64, np.float64)
    g = np.asarray(g, np.float64)
    n64 = np.linalg.norm(g64)
    if n64 == 0:
        assert not g.any(), f"{name}: expected no gradient"
        return 0.0
    rel = np.linalg.norm(g - g64) / n64
    assert rel <= REL, f"{name}: relative error {rel:.3g}"
    assert np.abs(g - g64).max() <= REL * np.abs(g64).max(), f"{name}: element error"
    return rel


def _run(scene):
    job = oracle_job(scene)
    scene.rerecord()            # the oracle consumed the recording; the GPU gets the same one
    return job


@pytest.fixture(autouse=True)
def _device():
    if pf.lib().pfdInit() != 0:
        pytest.fail("no GPU: " + pf.lib().pfdLastError().decode())


@pytest.mark.parametrize("case", list(CASES))
def test_forward_and_adjoint_match_fp64(case):
    s = Scene(**CASES[case])
    try:
        job = _run(s)
        p64 = params_of(s)
        p32 = [a.astype(np.float32) for a in p64]
        img = s.gpu_forward(*p32)
        # integer pass: bit-identical RGBA8 target
        assert np.array_equal(s.buf[: s.W * s.H], job.color()), "integer image differs from the oracle"
        img64 = job.forward(*p64)
        assert np.abs(img - img64).max() <= 2e-5
        target = target_of(s, job, color_sigma=0.0 if s.textured else 0.05)
        d = 2.0 * (img64 - target)
        g = s.gpu_backward(*p32, d.astype(np.float32))
        g64 = job.backward(*p64, d)
        for name, a, b in zip(("texels", "colors", "ldiff", "lspec"), g, g64):
            _rel_ok(f"{case}/{name}", a, b)
    finally:
        s.close()


def test_rerender_with_new_parameters_keeps_coverage():
    s = Scene(seed=11)
    try:
        job = _run(s)
        t, c, d, sp = params_of(s)
        s.gpu_forward(*[a.astype(np.float32) for a in (t, c, d, sp)])
        rng = np.random.default_rng(3)
        t2 = np.clip(t + rng.normal(0, 0.2, t.shape), 0, 1)
        c2 = np.clip(c + rng.normal(0, 0.1, c.shape), 0, 1)
        img2 = s.gpu_forward(*[a.astype(np.float32) for a in (t2, c2, d * 0.5, sp * 2.0)])
        assert np.abs(img2 - job.forward(t2, c2, d * 0.5, sp * 2.0)).max() <= 2e-5
    finally:
        s.close()


def test_c5_size_gradients():
    """512x512, 20k-triangle grid mesh, 256x256x4 float texels, per-vertex
    float colours, 2 lights; L2 loss against a target from a texture
    perturbed by N(0, 0.1)."""
    s = Scene(width=512, height=512, nx=100, nz=100, tex=(256, 256), lights=2, seed=5)
    try:
        job = _run(s)
        p64 = params_of(s)
        p32 = [a.astype(np.float32) for a in p64]
        img = s.gpu_forward(*p32)
        img64 = job.forward(*p64)
        assert np.abs(img - img64).max() <= 2e-5
        target = target_of(s, job)
        d = 2.0 * (img64 - target)
        g = s.gpu_backward(*p32, d.astype(np.float32))
        g64 = job.backward(*p64, d)
        rels = [_rel_ok(n, a, b) for n, a, b in zip(("texels", "colors", "ldiff", "lspec"), g, g64)]
        print("C5 relative errors", rels, flush=True)
    finally:
        s.close()


def test_autograd_function_wrapper():
    """torch.autograd.Function over the C entry points (ctypes), as a
    harness would use it: gradients equal the FP64 ones, and a few steps of
    gradient descent on the texels lower the loss."""
    torch = pytest.importorskip("torch")
    s = Scene(width=96, height=72, nx=12, nz=9, tex=(32, 32), seed=12)
    try:
        job = _run(s)

        class Render(torch.autograd.Function):
            @staticmethod
            def forward(ctx, texels, colors, ldiff, lspec):
                ctx.save_for_backward(texels, colors, ldiff, lspec)
                img = s.gpu_forward(*(a.detach().cpu().numpy() for a in (texels, colors, ldiff, lspec)))
                return torch.from_numpy(img)

            @staticmethod
            def backward(ctx, gimg):
                args = [a.detach().cpu().numpy() for a in ctx.saved_tensors]
                gt, gc, gd, gs = s.gpu_backward(*args, gimg.contiguous().cpu().numpy())
                return tuple(torch.from_numpy(x) for x in (gt, gc, gd, gs))

        p64 = params_of(s)
        target = torch.from_numpy(target_of(s, job).astype(np.float32))
        ps = [torch.tensor(a.astype(np.float32), requires_grad=True) for a in p64]
        loss = ((Render.apply(*ps) - target) ** 2).sum()
        loss.backward()
        img64 = job.forward(*p64)
        g64 = job.backward(*p64, 2.0 * (img64 - target.double().numpy()))
        for name, a, b in zip(("texels", "colors", "ldiff", "lspec"), ps, g64):
            _rel_ok(name, a.grad.numpy(), b)
        losses = [float(loss)]
        tex = ps[0]
        for _ in range(5):
            tex.grad = None
            loss = ((Render.apply(tex, *ps[1:]) - target) ** 2).sum()
            loss.backward()
            with torch.no_grad():
                tex -= 0.05 * tex.grad
            losses.append(float(loss))
        assert losses[-1] < 0.5 * losses[0], losses
    finally:
        s.close()
