"""Render lists (renderlist.c:6-97, context.c:1580-1683): error codes, the
context state that pfEndList and pfCallList restore, and the triangles a
called list records.  Where the reference's behaviour is defined the same
call sequence also runs through the compiled reference and the two traces
must agree.  Images of list scenes are covered by the reference goldens
list_* (test_oracle_golden.py, test_gpu_parity.py)."""
import ctypes as C

import numpy as np
import pytest

import pixelforge_amd as pf
from conftest import REF_SO, OracleExecutor, have_ref
from test_host_api import FrameLog

PF_RGBA, PF_UNSIGNED_BYTE = 7, 0
PF_TRIANGLES, PF_QUADS = 2, 5
PF_BLEND = 0x0004
PF_CURRENT_COLOR, PF_CURRENT_NORMAL, PF_CURRENT_TEXTURE_COORDS = 10004, 10005, 10006
PF_INVALID_VALUE, PF_INVALID_OPERATION = 2, 5


def _bind(path):
    L = C.CDLL(path)
    L.pfCreateContext.restype = C.c_void_p
    L.pfCreateContext.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int, C.c_int]
    L.pfMakeCurrent.argtypes = [C.c_void_p]
    L.pfDeleteContext.argtypes = [C.c_void_p]
    L.pfGenList.restype = C.c_void_p
    L.pfNewList.argtypes = [C.c_void_p]
    L.pfCallList.argtypes = [C.c_void_p]
    L.pfDeleteList.argtypes = [C.POINTER(C.c_void_p)]
    L.pfGetError.restype = C.c_int
    L.pfIsEnabled.restype = C.c_bool
    L.pfIsEnabled.argtypes = [C.c_int]
    L.pfEnable.argtypes = [C.c_int]
    L.pfBegin.argtypes = [C.c_int]
    L.pfColor4ub.argtypes = [C.c_ubyte] * 4
    L.pfNormal3f.argtypes = [C.c_float] * 3
    L.pfTexCoord2f.argtypes = [C.c_float] * 2
    L.pfVertex3f.argtypes = [C.c_float] * 3
    L.pfGetIntegerv.argtypes = [C.c_int, C.c_void_p]
    L.pfGetFloatv.argtypes = [C.c_int, C.c_void_p]
    if hasattr(L, "pfxSetExecutor"):
        L.pfxSetExecutor.argtypes = [C.c_void_p, C.c_void_p]
        L.pfxSetExecutor.restype = None
    return L


def _snap(L):
    col = (C.c_int * 4)()
    nrm = (C.c_float * 3)()
    uv = (C.c_float * 2)()
    L.pfGetIntegerv(PF_CURRENT_COLOR, col)
    L.pfGetFloatv(PF_CURRENT_NORMAL, nrm)
    L.pfGetFloatv(PF_CURRENT_TEXTURE_COORDS, uv)
    return (tuple(col), tuple(nrm), tuple(uv), bool(L.pfIsEnabled(PF_BLEND)))


def _tri(L, z=-0.5):
    L.pfBegin(PF_TRIANGLES)
    L.pfVertex3f(-0.5, -0.5, z)
    L.pfVertex3f(0.5, -0.5, z)
    L.pfVertex3f(0.0, 0.5, z)
    L.pfEnd()


def _state_trace(L):
    """A call sequence with defined behaviour; returns every snapshot and
    error code it produces.  Nothing is drawn (lists are only called while
    another is being recorded)."""
    buf = np.zeros(16 * 16 + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, 16, 16, PF_RGBA, PF_UNSIGNED_BYTE)
    assert ctx
    L.pfMakeCurrent(ctx)
    out = []
    try:
        L.pfColor4ub(10, 20, 30, 40)
        L.pfNormal3f(0.0, 1.0, 0.0)
        L.pfTexCoord2f(0.25, 0.5)
        out.append(("before", _snap(L)))
        A = C.c_void_p(L.pfGenList())
        B = C.c_void_p(L.pfGenList())
        L.pfNewList(A)
        L.pfColor4ub(1, 2, 3, 4)
        L.pfNormal3f(1.0, 0.0, 0.0)
        L.pfTexCoord2f(0.75, 0.125)
        L.pfEnable(PF_BLEND)
        _tri(L)
        out.append(("recording", _snap(L)))
        L.pfEndList()
        out.append(("after EndList", _snap(L)))
        out.append(("err", L.pfGetError()))
        L.pfNewList(B)
        L.pfColor4ub(7, 8, 9, 10)
        L.pfCallList(A)
        out.append(("after nested CallList", _snap(L)))
        L.pfEndList()
        out.append(("after EndList B", _snap(L)))
        out.append(("err", L.pfGetError()))
        L.pfEndList()
        out.append(("EndList without NewList", L.pfGetError()))
        out.append(("after stray EndList", _snap(L)))
        L.pfNewList(None)
        out.append(("NewList(NULL)", L.pfGetError()))
        L.pfDeleteList(C.byref(A))
        L.pfDeleteList(C.byref(B))
        out.append(("deleted", (A.value, B.value)))
    finally:
        L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)
    return out


def test_list_state_restore_and_errors():
    t = dict((k, v) for k, v in _state_trace(_bind(pf.LIB_PATH)) if k != "err")
    before = ((10, 20, 30, 40), (0.0, 1.0, 0.0), (0.25, 0.5), False)
    assert t["before"] == before
    assert t["recording"] == ((1, 2, 3, 4), (1.0, 0.0, 0.0), (0.75, 0.125), True)
    assert t["after EndList"] == before
    # pfCallList saves the context in the one backup slot that pfEndList
    # restores from, so list B ends with the colour set inside it.
    assert t["after nested CallList"] == ((7, 8, 9, 10), (0.0, 1.0, 0.0), (0.25, 0.5), False)
    assert t["after EndList B"] == t["after nested CallList"]
    assert t["EndList without NewList"] == PF_INVALID_OPERATION
    assert t["after stray EndList"] == t["after EndList B"]
    assert t["NewList(NULL)"] == PF_INVALID_VALUE
    assert t["deleted"] == (None, None)


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built")
def test_list_state_trace_matches_reference():
    ours = _state_trace(_bind(pf.LIB_PATH))
    ref = _state_trace(_bind(REF_SO))
    assert ours == ref


def test_called_list_records_its_triangles():
    """A list of a triangle and a quad, called twice, records the same
    triangles as drawing them twice (the quad as a fan of two)."""
    L = _bind(pf.LIB_PATH)
    log = FrameLog()
    buf = np.zeros(32 * 24 + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, 32, 24, PF_RGBA, PF_UNSIGNED_BYTE)
    L.pfMakeCurrent(ctx)
    L.pfxSetExecutor(C.cast(log.cb, C.c_void_p), None)
    try:
        A = C.c_void_p(L.pfGenList())
        L.pfNewList(A)
        _tri(L)
        L.pfBegin(PF_QUADS)
        for x, y in [(-1, -1), (1, -1), (1, 1), (-1, 1)]:
            L.pfVertex3f(x, y, -0.25)
        L.pfEnd()
        L.pfEndList()
        L.pfFinish()                      # nothing recorded while the list was built
        L.pfCallList(A)
        L.pfCallList(A)
        L.pfFinish()
        _tri(L)
        L.pfBegin(PF_QUADS)
        for x, y in [(-1, -1), (1, -1), (1, 1), (-1, 1)]:
            L.pfVertex3f(x, y, -0.25)
        L.pfEnd()
        L.pfFinish()
        L.pfDeleteList(C.byref(A))
    finally:
        L.pfxSetExecutor(None, None)
        L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)
    drawn = [f["ntri"] for f in log.frames if f["ntri"]]
    assert drawn == [6, 3]


def test_list_undefined_reference_cases_are_errors():
    """Cases where the reference dereferences NULL or indexes before a
    list's first call: here they set an error and change nothing."""
    L = _bind(pf.LIB_PATH)
    buf = np.zeros(8 * 8 + 16, np.uint32)
    ctx = L.pfCreateContext(buf.ctypes.data, 8, 8, PF_RGBA, PF_UNSIGNED_BYTE)
    L.pfMakeCurrent(ctx)
    try:
        with OracleExecutor():
            L.pfCallList(None)
            assert L.pfGetError() == PF_INVALID_VALUE
            A = C.c_void_p(L.pfGenList())
            L.pfNewList(A)
            L.pfVertex3f(0.0, 0.0, 0.0)             # no pfBegin since pfNewList
            assert L.pfGetError() == PF_INVALID_OPERATION
            L.pfEndList()
            assert L.pfGetError() == 0
            # Calling a list while recording into it appends copies of its
            # calls; the replay stops at the count it started with.
            L.pfNewList(A)
            _tri(L)
            L.pfCallList(A)
            L.pfCallList(A)
            L.pfEndList()
            assert L.pfGetError() == 0
            L.pfDeleteList(C.byref(A))
            assert A.value is None
            L.pfDeleteList(None)
    finally:
        L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)


@pytest.mark.parametrize("name", [n for n, _ in __import__("scenes").LIST_SCENES])
def test_list_calls_recorded_as_array_draws(name):
    """A sealed list called with filled triangle modes and no colour
    material is recorded as array draws of its vertices (the GPU forms the
    triangles); replayed through immediate mode instead, the image is the
    same, and both equal the reference's golden."""
    import os
    import scenes
    from conftest import GOLDEN, render_oracle
    kw = dict(scenes.LIST_SCENES)[name]
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    images = []
    for on in (1, 0):
        pf.lib().pfxSetArrayDraws(on)
        try:
            c, d, e, _ = render_oracle(scenes.params(kw))
        finally:
            pf.lib().pfxSetArrayDraws(1)
        assert e == 0
        images.append((c, d))
    for c, d in images:
        assert np.array_equal(c, g["color"]) and np.array_equal(d.view(np.uint32), g["depth"].view(np.uint32))
    # The calls that can be array draws are: the recorded frame says so.
    from test_host_api import Frame, EXEC
    counts = []
    cb = EXEC(lambda fp, user: (counts.append(C.cast(fp, C.POINTER(Frame)).contents.numArrays), 0)[1])
    pf.lib().pfxSetExecutor(C.cast(cb, C.c_void_p), None)
    try:
        pf.render(scenes.params(kw))
    finally:
        pf.lib().pfxSetExecutor(None, None)
    if kw.get("colorMaterial"):
        assert counts == [0]                      # colour material: immediate-mode replay
    else:
        assert counts and counts[0] > 0
