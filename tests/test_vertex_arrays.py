"""Vertex arrays, getters, framebuffer pixel setters and the out-of-path
entry points of the host library (no GPU needed).

The images of pfDrawElements / pfDrawArrays are pinned by the reference
goldens of tests/scenes.py ARRAY_SCENES (test_oracle_golden.py); here the
recording itself is checked: the same mesh drawn in immediate mode and
through every array path gives the same recorded triangles, array
texcoords skip the texture matrix (context.c:1335-1352), corners carry the
colour-array index, and the error codes follow context.c:1161-1435.
"""
import ctypes as C
import os

import numpy as np
import pytest

import pixelforge_amd as pf
import scenes
from conftest import GOLDEN, _oracle
from test_host_api import Frame, EXEC

PF_FLOAT, PF_DOUBLE, PF_UBYTE, PF_USHORT, PF_UINT, PF_SHORT, PF_INT = 10, 11, 0, 1, 5, 7, 8
PF_VERTEX_ARRAY, PF_NORMAL_ARRAY, PF_COLOR_ARRAY, PF_TEXTURE_COORD_ARRAY = 0x100, 0x200, 0x400, 0x800
PF_TRIANGLES, PF_QUADS = 2, 5
INVALID_ENUM, INVALID_VALUE, INVALID_OPERATION = 1, 2, 5


def _ctx(w=32, h=24):
    L = pf.lib()
    buf = np.zeros(w * h + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, w, h, 7, 0)
    assert ctx
    L.pfMakeCurrent(ctx)
    return L, ctx, buf


@pytest.fixture(params=[True, False], ids=["array-draws", "vertex-loop"])
def array_draws(request):
    """Runs a test with pfDrawElements / pfDrawArrays recorded whole (the
    default: the triangles are formed later, here by the oracle's
    restatement of the reference's vertex loop) and vertex by vertex on the
    host; both must record the same triangles."""
    pf.lib().pfxSetArrayDraws(1 if request.param else 0)
    yield request.param
    pf.lib().pfxSetArrayDraws(1)


class Capture:
    """Executor keeping copies of each frame's triangle streams, with array
    draws expanded by the oracle."""

    def __init__(self):
        self.frames = []
        self.arrays = []
        self.cb = EXEC(self._run)

    def _run(self, fp, user):
        raw = C.cast(fp, C.POINTER(Frame)).contents
        self.arrays.append(raw.numArrays)
        f = Frame()
        assert _oracle.pforacle_expand(C.c_void_p(fp), C.byref(f)) == 0
        n = f.tris.count
        get = lambda p, k, t: np.ctypeslib.as_array(C.cast(p, C.POINTER(t)), (n * k,)).copy() if p else None
        self.frames.append(dict(pos=get(f.tris.position, 12, C.c_float), col=get(f.tris.color, 3, C.c_uint32),
                                uv=get(f.tris.texcoord, 6, C.c_float), nrm=get(f.tris.normal, 9, C.c_float),
                                vid=get(f.tris.vertexId, 3, C.c_uint32)))
        _oracle.pforacle_expand_free()
        return 0

    def __enter__(self):
        pf.lib().pfxSetExecutor(C.cast(self.cb, C.c_void_p), None)
        return self

    def __exit__(self, *exc):
        pf.lib().pfxSetExecutor(None, None)


def _arr(a, t):
    a = np.ascontiguousarray(a, t)
    return a, a.ctypes.data


def _fl(*v):
    return [C.c_float(x) for x in v]


def test_array_goldens_equal_immediate_mode():
    """The reference draws the same pixels for the mesh in immediate mode
    and through every array path (its own goldens say so)."""
    def load(n):
        g = np.load(os.path.join(GOLDEN, f"{n}.npz"), allow_pickle=False)
        return g["color"], g["depth"]
    c0, d0 = load("mesh_immediate")
    assert np.count_nonzero(c0 != c0[0, 0]) > 1000
    for n in ("mesh_elements_uint", "mesh_elements_ubyte", "mesh_elements_ushort", "mesh_draw_arrays",
              "mesh_double_positions"):
        c, d = load(n)
        assert np.array_equal(c, c0) and np.array_equal(d.view(np.uint32), d0.view(np.uint32)), n
    c1, _ = load("mesh_colors_ubyte")
    for n in ("mesh_colors_ushort", "mesh_colors_uint", "mesh_colors_double"):
        assert np.array_equal(load(n)[0], c1), n


def test_draw_elements_records_the_immediate_mode_triangles(array_draws):
    L, ctx, _ = _ctx()
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (6, 3)).astype(np.float32)
    col = (rng.integers(0, 256, (6, 4)) / 256.0).astype(np.float32)
    uv = rng.uniform(0, 1, (6, 2)).astype(np.float32)
    nrm = rng.uniform(-1, 1, (6, 3)).astype(np.float32)
    idx = np.array([0, 1, 2, 3, 4, 5, 5, 2, 0], np.uint32)
    try:
        L.pfDisable(0x10)
        L.pfMatrixMode(2)                              # PF_TEXTURE: arrays must ignore it
        L.pfTranslatef(*_fl(0.5, 0.25, 0.0))
        L.pfMatrixMode(0)
        with Capture() as cap:
            L.pfBegin(PF_TRIANGLES)
            L.pfMatrixMode(2); L.pfLoadIdentity(); L.pfMatrixMode(0)
            for i in idx:
                L.pfColor4f(*_fl(*col[i]))
                L.pfTexCoord2f(*_fl(*uv[i]))
                L.pfNormal3f(*_fl(*nrm[i]))
                L.pfVertex3f(*_fl(*pos[i]))
            L.pfEnd()
            L.pfFinish()
            L.pfMatrixMode(2); L.pfTranslatef(*_fl(0.5, 0.25, 0.0)); L.pfMatrixMode(0)
            keep = [_arr(pos, np.float32), _arr(col, np.float32), _arr(uv, np.float32), _arr(nrm, np.float32),
                    _arr(idx, np.uint32)]
            L.pfEnable(PF_VERTEX_ARRAY | PF_NORMAL_ARRAY | PF_COLOR_ARRAY | PF_TEXTURE_COORD_ARRAY)
            L.pfVertexPointer(3, PF_FLOAT, 12, C.c_void_p(keep[0][1]))
            L.pfColorPointer(4, PF_FLOAT, 999, C.c_void_p(keep[1][1]))   # stride is ignored
            L.pfTexCoordPointer(PF_FLOAT, 0, C.c_void_p(keep[2][1]))
            L.pfNormalPointer(PF_FLOAT, 0, C.c_void_p(keep[3][1]))
            L.pfDrawElements(PF_TRIANGLES, len(idx), PF_UINT, C.c_void_p(keep[4][1]))
            L.pfFinish()
        assert L.pfGetError() == 0
    finally:
        L.pfDeleteContext(ctx)
    imm, arr = cap.frames
    assert cap.arrays == [0, 1 if array_draws else 0]
    for k in ("pos", "col", "uv", "nrm"):
        assert np.array_equal(imm[k], arr[k]), k
    assert imm["vid"] is None
    assert arr["vid"].tolist() == idx.tolist()


def test_draw_arrays_types_and_missing_attributes(array_draws):
    """Positions of every type, w = 1 and zero z for size 2; attributes with
    no enabled array: zero normal and texcoord, current colour, no index."""
    L, ctx, _ = _ctx()
    try:
        L.pfDisable(0x10)
        L.pfColor4ub(10, 20, 30, 40)
        with Capture() as cap:
            L.pfEnable(PF_VERTEX_ARRAY)
            for t, np_t in ((PF_SHORT, np.int16), (PF_INT, np.int32), (PF_DOUBLE, np.float64)):
                v, p = _arr([[9, 9], [9, 9], [1, 0], [0, 1], [-1, -1]], np_t)
                L.pfVertexPointer(2, t, 0, C.c_void_p(p))
                L.pfDrawArrays(PF_TRIANGLES, 2, 3)
            L.pfFinish()
    finally:
        L.pfDeleteContext(ctx)
    f = cap.frames[0]
    pos = f["pos"].reshape(-1, 3, 4)
    assert pos.shape[0] == 3
    for t in range(3):
        assert pos[t].tolist() == [[1, 0, 0, 1], [0, 1, 0, 1], [-1, -1, 0, 1]]
    assert f["nrm"] is None or not f["nrm"].any()
    assert set(f["col"].tolist()) == {10 | 20 << 8 | 30 << 16 | 40 << 24}
    assert f["vid"] is None


def test_color_array_conversions(array_draws):
    """ubyte as is, ushort >> 8, uint >> 24, float/double (ubyte)(x * 255);
    missing alpha is 255 (context.c:1356-1392)."""
    L, ctx, _ = _ctx()
    tri = np.array([[1, 0], [0, 1], [-1, -1]], np.float32)
    cases = [
        (PF_UBYTE, np.uint8, [[1, 2, 3, 4], [5, 6, 7, 8], [250, 251, 252, 253]], 4),
        (PF_USHORT, np.uint16, [[0x1234, 0xFFFF, 0x00FF, 0x8000]] * 3, 4),
        (PF_UINT, np.uint32, [[0x12345678, 0xFFFFFFFF, 0x00FFFFFF, 0x80000000]] * 3, 4),
        (PF_FLOAT, np.float32, [[0.5, 1.0, 0.999, 2.0]] * 3, 4),
        (PF_DOUBLE, np.float64, [[0.5, 1.0, 0.999, -0.5]] * 3, 4),
        (PF_UBYTE, np.uint8, [[9, 8, 7]] * 3, 3),
    ]
    expect = [[1, 2, 3, 4], [0x12, 0xFF, 0x00, 0x80], [0x12, 0xFF, 0x00, 0x80], [127, 255, 254, 254],
              [127, 255, 254, 129], [9, 8, 7, 255]]   # -127.5 -> (int)-127 -> byte 0x81
    try:
        L.pfDisable(0x10)
        L.pfEnable(PF_VERTEX_ARRAY | PF_COLOR_ARRAY)
        tv, tp = _arr(tri, np.float32)
        L.pfVertexPointer(2, PF_FLOAT, 0, C.c_void_p(tp))
        keep = []
        with Capture() as cap:
            for t, np_t, vals, size in cases:
                a, p = _arr(vals, np_t)
                keep.append(a)
                L.pfColorPointer(size, t, 0, C.c_void_p(p))
                L.pfDrawArrays(PF_TRIANGLES, 0, 3)
            L.pfFinish()
        assert L.pfGetError() == 0
    finally:
        L.pfDeleteContext(ctx)
    col = cap.frames[0]["col"].reshape(-1, 3)
    for i, e in enumerate(expect):
        got = [(int(col[i][0]) >> (8 * k)) & 0xFF for k in range(4)]
        assert got == e, (i, got, e)
    assert cap.frames[0]["vid"].tolist() == [0, 1, 2] * len(cases)


def test_array_errors():
    L, ctx, _ = _ctx()
    try:
        v, p = _arr(np.zeros((3, 3)), np.float32)
        L.pfDrawArrays(PF_TRIANGLES, 0, 3)
        assert L.pfGetError() == INVALID_OPERATION       # PF_VERTEX_ARRAY disabled
        L.pfVertexPointer(1, PF_FLOAT, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_VALUE
        L.pfVertexPointer(3, PF_UBYTE, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_ENUM
        L.pfColorPointer(2, PF_FLOAT, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_VALUE
        L.pfColorPointer(4, PF_SHORT, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_ENUM
        L.pfNormalPointer(PF_INT, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_ENUM
        L.pfTexCoordPointer(PF_UBYTE, 0, C.c_void_p(p))
        assert L.pfGetError() == INVALID_ENUM
        L.pfEnable(PF_VERTEX_ARRAY)
        L.pfDrawElements(PF_TRIANGLES, 3, PF_FLOAT, C.c_void_p(p))
        assert L.pfGetError() == INVALID_ENUM            # index type checked first
        L.pfDrawArrays(PF_TRIANGLES, 0, 3)
        assert L.pfGetError() == INVALID_OPERATION       # no position array bound
    finally:
        L.pfDeleteContext(ctx)


def test_getters():
    L, ctx, _ = _ctx(40, 30)
    try:
        i4 = (C.c_int32 * 16)()
        f16 = (C.c_float * 16)()
        d16 = (C.c_double * 16)()
        b = C.c_bool()
        L.pfGetIntegerv(10000, i4)                     # PF_VIEWPORT
        assert list(i4[:4]) == [0, 0, 40, 30]
        L.pfClearColor(10, 20, 30, 40)
        L.pfGetIntegerv(10001, i4)
        assert list(i4[:4]) == [10, 20, 30, 40]
        L.pfGetFloatv(10001, f16)
        assert f16[0] == np.float32(10 * np.float32(1 / 255.0))
        L.pfGetDoublev(10001, d16)
        assert d16[1] == 20 / 255.0
        L.pfGetBooleanv(0x10, C.byref(b))              # culling is on by default
        assert b.value
        L.pfGetBooleanv(0x4, C.byref(b))
        assert not b.value
        L.pfGetIntegerv(10021, i4)                     # PF_MAX_LIGHTS
        assert i4[0] == 8
        L.pfTranslatef(*_fl(1, 2, 3))
        L.pfGetFloatv(10015, f16)                      # PF_MODELVIEW_MATRIX
        assert list(f16[12:15]) == [1, 2, 3]
        L.pfPointSize(C.c_float(3.0)); L.pfLineWidth(C.c_float(2.5)); L.pfPixelZoom(*_fl(2, 4))
        L.pfGetFloatv(10011, f16); assert f16[0] == 3.0
        L.pfGetFloatv(10012, f16); assert f16[0] == 2.5
        L.pfGetFloatv(10033, f16); assert f16[0] == 4.0     # PF_ZOOM_Y
        L.pfPointSize(C.c_float(0.0))
        assert L.pfGetError() == INVALID_VALUE
        L.pfVertexPointer(3, PF_DOUBLE, 24, None)
        L.pfGetIntegerv(10022, i4); assert i4[0] == 3
        L.pfGetIntegerv(10023, i4); assert i4[0] == 24
        L.pfGetIntegerv(10024, i4); assert i4[0] == PF_DOUBLE
        L.pfGetIntegerv(12345, i4)
        assert L.pfGetError() == INVALID_ENUM
        ptr = C.c_void_p()
        L.pfGetPointerv(10009, C.byref(ptr))           # PF_DEPTH_FUNC
        assert C.cast(ptr, C.POINTER(C.c_int32))[0] == 2   # PF_LESS
    finally:
        L.pfDeleteContext(ctx)


def test_framebuffer_pixel_setters():
    L, ctx, buf = _ctx(8, 4)

    class FB(C.Structure):
        _fields_ = [("texture", C.c_void_p), ("zbuffer", C.POINTER(C.c_float))]

    class Color(C.Structure):
        _fields_ = [("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("a", C.c_uint8)]
    try:
        L.pfGenFramebuffer.restype = FB
        fb = L.pfGenFramebuffer(8, 4, 7, 0)
        L.pfGetFramebufferPixel.restype = Color
        L.pfGetFramebufferDepth.restype = C.c_float
        L.pfSetFramebufferPixel(C.byref(fb), 1, 2, Color(1, 2, 3, 4))
        c = L.pfGetFramebufferPixel(C.byref(fb), 1, 2)
        assert (c.r, c.g, c.b, c.a) == (1, 2, 3, 4)
        L.pfSetFramebufferPixelDepth(C.byref(fb), 3, 1, C.c_float(0.5), Color(9, 9, 9, 9))
        assert L.pfGetFramebufferDepth(C.byref(fb), 3, 1) == 0.5
        L.pfSetFramebufferPixelDepthTest(C.byref(fb), 3, 1, C.c_float(0.7), Color(7, 7, 7, 7), 2)   # LESS: fails
        assert L.pfGetFramebufferPixel(C.byref(fb), 3, 1).r == 9
        L.pfSetFramebufferPixelDepthTest(C.byref(fb), 3, 1, C.c_float(0.7), Color(7, 7, 7, 7), 1)   # NOTEQUAL
        assert L.pfGetFramebufferPixel(C.byref(fb), 3, 1).r == 7
        assert L.pfGetFramebufferDepth(C.byref(fb), 3, 1) == np.float32(0.7)
        L.pfSetFramebufferPixelDepthTest(C.byref(fb), 0, 0, C.c_float(0.1), Color(7, 7, 7, 7), 6)
        assert L.pfGetError() == INVALID_ENUM
        L.pfDeleteFramebuffer(C.byref(fb))
    finally:
        L.pfDeleteContext(ctx)


@pytest.mark.parametrize("name,kw", scenes.ARRAY_SCENES, ids=[n for n, _ in scenes.ARRAY_SCENES])
def test_vertex_loop_recording_matches_golden(name, kw):
    """The array scenes recorded vertex by vertex on the host (the path the
    library keeps for lists and point / line polygon modes) give the
    reference's image too; test_oracle_golden covers them recorded whole."""
    from conftest import render_oracle
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    pf.lib().pfxSetArrayDraws(0)
    try:
        c, d, e, _ = render_oracle(scenes.params(kw))
    finally:
        pf.lib().pfxSetArrayDraws(1)
    assert e == 0
    assert np.count_nonzero(c != c[0, 0]) > 1000
    assert np.array_equal(c, g["color"]) and np.array_equal(d.view(np.uint32), g["depth"].view(np.uint32))
