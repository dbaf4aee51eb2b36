"""Host library: exported symbols, error semantics of the state machine,
and the recorded frames it hands to the executor (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np

import pixelforge_amd as pf
from conftest import ROOT, OracleExecutor


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"\b(pf[A-Za-z0-9_]+)\s*\(", text)) - {"pfdInit"} | ({"pfdInit"} if "pfdInit" in text else set())


def test_library_exports_every_declared_symbol():
    lib = pf.lib()
    names = set()
    for h in ("pixelforge.h", "pf_ext.h", "pf_device.h"):
        names |= {n for n in _declared(h) if not n.startswith("pfx") or h == "pf_ext.h"}
    names = {n for n in names if n not in ("pfRectfv",)}
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def _ctx(w=32, h=24):
    L = pf.lib()
    buf = np.zeros(w * h + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, w, h, 7, 0)
    assert ctx
    L.pfMakeCurrent(ctx)
    return L, ctx, buf


def test_error_codes_are_sticky_and_cleared():
    L, ctx, _ = _ctx()
    try:
        assert L.pfGetError() == 0
        L.pfBlendFunc(99)
        assert L.pfGetError() == 1                     # PF_INVALID_ENUM
        assert L.pfGetError() == 0
        L.pfDepthFunc(6)
        L.pfCullFace(2)
        assert L.pfGetError() == 1
        L.pfMatrixMode(1)                              # PF_PROJECTION
        L.pfPushMatrix(); L.pfPushMatrix(); L.pfPushMatrix()
        assert L.pfGetError() == 3                     # PF_STACK_OVERFLOW (stack of 2)
        for _ in range(3):
            L.pfPopMatrix()
        assert L.pfGetError() == 4                     # PF_STACK_UNDERFLOW
        L.pfEnableLight(9)
        assert L.pfGetError() == 2                     # PF_INVALID_VALUE
        L.pfEnableLight(1); L.pfEnableLight(1)
        assert L.pfGetError() == 5                     # already enabled: PF_INVALID_OPERATION
        L.pfViewport(-40, 0, 32, 24)
        assert L.pfGetError() == 5
    finally:
        L.pfDeleteContext(ctx)


def test_unsupported_formats_are_rejected():
    """Render targets: the nineteen pairs whose reference accessors touch
    only their own pixel (RGBA, RGB and BGR bytes; RGB/BGR 5_6_5; RGBA/BGRA
    5_5_5_1 and 4_4_4_4; luminance-alpha, RGB, BGR, RGBA and BGRA in half
    and float components); every other pair is PF_INVALID_ENUM for
    pfCreateContext and pfGenFramebuffer."""
    L = pf.lib()
    buf = np.zeros(8 * 8 * 4 + 64, np.uint32)
    targets = {(7, 0), (6, 0), (8, 0), (6, 2), (8, 2), (7, 3), (9, 3), (7, 4), (9, 4)}
    targets |= {(f, t) for f in (5, 6, 7, 8, 9) for t in (9, 10)}
    accepted = set()
    for f in range(10):
        for t in range(12):
            ctx = L.pfCreateContext(buf.ctypes.data, 8, 8, f, t)
            if ctx:
                accepted.add((f, t))
                L.pfDeleteContext(ctx)
    assert accepted == targets

    class FB(C.Structure):
        _fields_ = [("texture", C.c_void_p), ("zbuffer", C.c_void_p)]
    L2, ctx, _ = _ctx()
    L2.pfGenFramebuffer.restype = FB
    try:
        for f, t in [(9, 0), (4, 0), (4, 9), (0, 10), (5, 0), (0, 0)]:
            fb = L2.pfGenFramebuffer(8, 8, f, t)
            assert not fb.texture and L2.pfGetError() == 1
        for f, t in sorted(targets):
            fb = L2.pfGenFramebuffer(8, 8, f, t)
            assert fb.texture and L2.pfGetError() == 0
            L2.pfDeleteFramebuffer(C.byref(fb))
    finally:
        L2.pfDeleteContext(ctx)
    L2, ctx, _ = _ctx()
    try:
        L2.pfGenTexture.restype = C.c_void_p
        assert not L2.pfGenTexture(buf.ctypes.data, 8, 8, 7, 1)  # RGBA / UNSIGNED_SHORT
        assert L2.pfGetError() == 1
    finally:
        L2.pfDeleteContext(ctx)


def test_texture_formats_with_a_getter_are_accepted():
    """pfGenTexture takes exactly the (format, type) pairs of the reference's
    texel getter table (pixel.h:3050-3095): 10 formats; UNSIGNED_BYTE,
    HALF_FLOAT and FLOAT for all, 5_6_5 for RGB/BGR, 5_5_5_1 and 4_4_4_4
    for RGBA/BGRA."""
    L, ctx, _ = _ctx()
    buf = np.zeros(64 * 16, np.uint32)
    L.pfGenTexture.restype = C.c_void_p
    L.pfDeleteTexture.argtypes = [C.POINTER(C.c_void_p), C.c_bool]
    try:
        accepted = set()
        for f in range(10):
            for t in range(12):
                tex = C.c_void_p(L.pfGenTexture(buf.ctypes.data, 4, 4, f, t))
                err = L.pfGetError()
                if tex.value:
                    accepted.add((f, t))
                    assert err == 0
                    L.pfDeleteTexture(C.byref(tex), False)
                else:
                    assert err == 1                    # PF_INVALID_ENUM
        expect = {(f, t) for f in range(10) for t in (0, 9, 10)}
        expect |= {(f, 2) for f in (6, 8)} | {(f, t) for f in (7, 9) for t in (3, 4)}
        assert accepted == expect and len(expect) == 36
    finally:
        L.pfDeleteContext(ctx)


EXEC = pf.EXECUTOR


class FrameLog:
    """Executor that records what each frame contains."""

    def __init__(self):
        self.frames = []
        self.cb = EXEC(self._run)

    def _run(self, fp, user):
        f = C.cast(fp, C.POINTER(Frame)).contents
        cmds = [(f.cmds[i].kind, f.cmds[i].first, f.cmds[i].last) for i in range(f.numCmds)]
        info = [f.tris.info[i] for i in range(f.tris.count)]
        vid = [f.tris.vertexId[i] for i in range(3 * f.tris.count)] if f.tris.vertexId else None
        self.frames.append(dict(w=f.width, h=f.height, ntri=f.tris.count, nstates=f.numStates,
                                cmds=cmds, info=info, tex=bool(f.tris.texcoord), nrm=bool(f.tris.normal),
                                vid=vid))
        return 0


class Tris(C.Structure):
    _fields_ = [("count", C.c_uint32), ("position", C.c_void_p), ("color", C.c_void_p),
                ("texcoord", C.c_void_p), ("normal", C.c_void_p), ("info", C.POINTER(C.c_uint32)),
                ("vertexId", C.POINTER(C.c_uint32))]


class Cmd(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("first", C.c_uint32), ("last", C.c_uint32),
                ("clearColor", C.c_uint32), ("clearDepth", C.c_float), ("pad", C.c_uint32 * 3)]


class Frame(C.Structure):
    """PFFframe (include/pf_frame.h)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("color", C.c_void_p), ("depth", C.c_void_p),
                ("tris", Tris), ("states", C.c_void_p), ("numStates", C.c_uint32),
                ("textures", C.c_void_p), ("numTextures", C.c_uint32),
                ("cmds", C.POINTER(Cmd)), ("numCmds", C.c_uint32),
                ("pixelOps", C.c_void_p), ("numPixelOps", C.c_uint32),
                ("points", C.c_void_p), ("numPoints", C.c_uint32),
                ("lines", C.c_void_p), ("numLines", C.c_uint32),
                ("arrays", C.c_void_p), ("numArrays", C.c_uint32), ("target", C.c_uint32)]


def test_recorded_frame_structure():
    L, ctx, _ = _ctx()
    log = FrameLog()
    L.pfxSetExecutor(C.cast(log.cb, C.c_void_p), None)
    try:
        L.pfClear(3)
        L.pfDisable(0x10)                              # culling off: both passes
        L.pfBegin(2)
        for v in [(-1, -1), (1, -1), (0, 1), (0, 0), (1, 1), (-1, 1)]:
            L.pfVertex2f(C.c_float(v[0]), C.c_float(v[1]))
        L.pfEnd()
        L.pfEnable(0x10)                               # cull back faces: FRONT pass only
        L.pfBegin(5)                                   # PF_QUADS -> fan of two
        for v in [(-1, -1), (1, -1), (1, 1), (-1, 1)]:
            L.pfVertex2f(C.c_float(v[0]), C.c_float(v[1]))
        L.pfEnd()
        L.pfFinish()
    finally:
        L.pfxSetExecutor(None, None)
        L.pfDeleteContext(ctx)
    assert len(log.frames) == 1
    f = log.frames[0]
    assert f["ntri"] == 4 and f["cmds"] == [(1, 0, 0), (2, 0, 4)]
    both, front = 0x03000000, 0x01000000
    assert [i & 0xFF000000 for i in f["info"]] == [both, both, front, front]
    assert not f["tex"] and not f["nrm"]
    assert f["vid"] is None                            # immediate mode: no colour-array indices


def test_target_change_splits_frames():
    L, ctx, _ = _ctx()
    log = FrameLog()
    L.pfxSetExecutor(C.cast(log.cb, C.c_void_p), None)
    L.pfGenFramebuffer.restype = None
    try:
        class FB(C.Structure):
            _fields_ = [("texture", C.c_void_p), ("zbuffer", C.c_void_p)]
        L.pfGenFramebuffer.restype = FB
        fb = L.pfGenFramebuffer(16, 8, 7, 0)
        L.pfClear(3)
        L.pfBindFramebuffer(C.byref(fb))
        L.pfEnable(0x2)                                # PF_FRAMEBUFFER
        L.pfClear(3)
        L.pfFinish()
        L.pfDisable(0x2)
        L.pfDeleteFramebuffer(C.byref(fb))
    finally:
        L.pfxSetExecutor(None, None)
        L.pfDeleteContext(ctx)
    assert [(f["w"], f["h"]) for f in log.frames] == [(32, 24), (16, 8)]


def test_gpu_path_fails_loudly_without_device():
    """With no GPU the default executor reports an error; nothing renders
    on the CPU behind the caller's back."""
    L = pf.lib()
    if L.pfdInit() == 0:
        return                                         # a GPU is present: covered by the gpu tests
    _, ctx, buf = _ctx()
    try:
        L.pfClear(3)
        L.pfFinish()
        assert L.pfGetError() == 5                     # PF_INVALID_OPERATION
        assert not buf.any()
    finally:
        L.pfDeleteContext(ctx)
