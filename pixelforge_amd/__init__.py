"""pixelforge_amd -- MI355X-native PixelForge triangle fragment pipeline.

The product is the C library libpixelforge_amd.so (include/pixelforge.h is
its drop-in API, include/pf_device.h the C ABI of its HIP kernels).  This
module only loads the built libraries with ctypes and mirrors the structs a
Python caller needs; it has no rendering code of its own and no fallback:
`lib()` raises if the native library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PIXELFORGE_AMD_LIB selects another build of the library (the sanitized
# host layer of `python -m pixelforge_amd.build --sanitize`).
LIB_PATH = os.path.abspath(os.environ.get("PIXELFORGE_AMD_LIB", os.path.join(HERE, "libpixelforge_amd.so")))
SCENES_PATH = os.path.join(HERE, "libpf_scenes.so")

# Scene kinds (csrc/pf_scenes.c)
RANDOM_NDC, HEIGHTFIELD, SPHERES, RANDOM_3D, MESH, LIST = 1, 2, 3, 4, 5, 6

# Enum values of include/pixelforge.h used by callers
PF_FRONT, PF_BACK = 0, 1
PF_TRIANGLES, PF_TRIANGLE_FAN, PF_TRIANGLE_STRIP, PF_QUADS, PF_QUAD_FAN, PF_QUAD_STRIP = range(2, 8)
PF_EQUAL, PF_NOTEQUAL, PF_LESS, PF_LEQUAL, PF_GREATER, PF_GEQUAL = range(6)
(PF_BLEND_AVERAGE, PF_BLEND_ALPHA, PF_BLEND_ADD, PF_BLEND_SUB, PF_BLEND_MUL,
 PF_BLEND_SCREEN, PF_BLEND_LIGHTEN, PF_BLEND_DARKEN) = range(8)
PF_REPEAT, PF_MIRRORED_REPEAT, PF_CLAMP_TO_EDGE = range(3)
PF_NEAREST, PF_BILINEAR = range(2)
PF_FLAT, PF_SMOOTH = range(2)


class SceneParams(C.Structure):
    """PFSparams of csrc/pf_scenes.c."""
    _fields_ = [
        ("kind", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
        ("ntris", C.c_uint32), ("seed", C.c_uint64),
        ("cull", C.c_int32), ("depthTest", C.c_int32), ("depthFunc", C.c_int32),
        ("blend", C.c_int32), ("blendMode", C.c_int32),
        ("texture", C.c_int32), ("texW", C.c_int32), ("texH", C.c_int32),
        ("wrap", C.c_int32), ("filter", C.c_int32), ("shade", C.c_int32),
        ("lighting", C.c_int32), ("numLights", C.c_int32), ("spot", C.c_int32),
        ("attenuation", C.c_int32), ("blendFraction", C.c_float),
        ("gridX", C.c_int32), ("gridZ", C.c_int32), ("spheres", C.c_int32),
        ("sphereLat", C.c_int32), ("sphereLon", C.c_int32),
        ("clearFirst", C.c_int32), ("colorMaterial", C.c_int32),
        ("triScale", C.c_float), ("flushEvery", C.c_int32), ("useQuads", C.c_int32),
        ("arrays", C.c_int32), ("indexType", C.c_int32), ("posType", C.c_int32),
        ("colType", C.c_int32), ("colSize", C.c_int32),
        ("texFormat", C.c_int32), ("texType", C.c_int32), ("texSwizzle", C.c_int32),
        ("pixelOps", C.c_int32), ("pixFormat", C.c_int32), ("pixType", C.c_int32),
        ("fog", C.c_int32), ("fogStart", C.c_float), ("fogEnd", C.c_float), ("fogAlpha", C.c_int32),
        ("points", C.c_int32), ("pointSize", C.c_float), ("polyPoint", C.c_int32),
        ("lines", C.c_int32), ("lineWidth", C.c_float), ("polyLine", C.c_int32),
        ("rects", C.c_int32), ("uvMin", C.c_float), ("uvMax", C.c_float), ("listOptions", C.c_int32),
        ("arrayMode", C.c_int32), ("fbFormat", C.c_int32), ("fbType", C.c_int32), ("fbMain", C.c_int32),
    ]


def scene(**kw):
    """SceneParams with the defaults of a smooth-shaded, depth-tested scene."""
    p = SceneParams()
    p.cull = -1
    p.depthTest = 1
    p.depthFunc = PF_LESS
    p.shade = PF_SMOOTH
    p.blendMode = PF_BLEND_ALPHA
    p.clearFirst = 1
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


class PFDstats(C.Structure):
    _fields_ = [("rasterTriangles", C.c_uint64), ("tilePairs", C.c_uint64),
                ("fragments", C.c_uint64), ("inputTriangles", C.c_uint64),
                ("numTiles", C.c_uint32), ("tileShift", C.c_uint32),
                ("deviceBytes", C.c_uint64), ("maxTileList", C.c_uint32), ("busyTiles", C.c_uint32),
                ("clippedTriangles", C.c_uint32), ("reserved", C.c_uint32)]


class PFDtimes(C.Structure):
    _fields_ = [("setupMs", C.c_float), ("binMs", C.c_float),
                ("rasterMs", C.c_float), ("totalMs", C.c_float)]


class PFDconfig(C.Structure):
    _fields_ = [("tileShift", C.c_uint32), ("rasterMode", C.c_uint32),
                ("reserved", C.c_uint32 * 6)]


EXECUTOR = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
BATCH_EXECUTOR = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_void_p)

_lib = None
_scenes = None


def lib():
    """The native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m pixelforge_amd.build")
        L = C.CDLL(LIB_PATH)
        L.pfdInit.restype = C.c_int
        L.pfdLastError.restype = C.c_char_p
        L.pfdJobCreate.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_void_p)]
        for name in ("pfdJobUploadTargets", "pfdJobLaunch", "pfdJobSync", "pfdJobDownload"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.pfdJobRunTimed.argtypes = [C.c_void_p, C.POINTER(PFDtimes)]
        L.pfdJobTimeReplays.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]
        L.pfdJobStats.argtypes = [C.c_void_p, C.POINTER(PFDstats)]
        L.pfdJobSetConfig.argtypes = [C.c_void_p, C.POINTER(PFDconfig)]
        L.pfdJobAutotune.argtypes = [C.c_void_p, C.POINTER(PFDconfig)]
        L.pfdSetDefaultConfig.argtypes = [C.POINTER(PFDconfig)]
        L.pfdJobDestroy.argtypes = [C.c_void_p]
        L.pfdJobDestroy.restype = None
        L.pfdTakeCapturedJob.restype = C.c_void_p
        L.pfxSetExecutor.argtypes = [C.c_void_p, C.c_void_p]
        L.pfxSetExecutor.restype = None
        L.pfxSetBatchExecutor.argtypes = [C.c_void_p, C.c_void_p]
        L.pfxSetBatchExecutor.restype = None
        L.pfSubmitBatch.argtypes = [C.c_void_p, C.c_uint32]
        L.pfSubmitBatch.restype = None
        L.pfCreateContext.restype = C.c_void_p
        L.pfCreateContext.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int]
        L.pfMakeCurrent.argtypes = [C.c_void_p]
        L.pfDeleteContext.argtypes = [C.c_void_p]
        L.pfGetError.restype = C.c_int
        L.pfdJobTarget.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.pfdJobFrameCount.argtypes = [C.c_void_p]
        L.pfdExecuteFrame.argtypes = [C.c_void_p]
        L.pfdExecuteFrameNoAlloc.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def scenes():
    global _scenes
    if _scenes is None:
        if not os.path.exists(SCENES_PATH):
            raise RuntimeError(f"{SCENES_PATH} is not built: run python -m pixelforge_amd.build")
        S = C.CDLL(SCENES_PATH)
        S.pfs_bind.restype = C.c_void_p
        S.pfs_bind.argtypes = [C.c_char_p]
        S.pfs_dlerror.restype = C.c_char_p
        S.pfs_render.argtypes = [C.c_void_p, C.POINTER(SceneParams), C.c_void_p, C.c_void_p]
        S.pfs_render.restype = C.c_int
        S.pfs_render_batch.argtypes = [C.c_void_p, C.POINTER(SceneParams), C.c_uint32, C.POINTER(C.c_void_p)]
        S.pfs_render_batch.restype = C.c_int
        S.pfs_record.argtypes = [C.c_void_p, C.POINTER(SceneParams), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        S.pfs_record.restype = C.c_int
        S.pfs_free_texture.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        S.pfs_free_texture.restype = None
        S.pfs_tape_make.restype = C.c_void_p
        S.pfs_tape_make.argtypes = [C.POINTER(SceneParams)]
        S.pfs_tape_free.argtypes = [C.c_void_p]
        S.pfs_tape_free.restype = None
        S.pfs_tape_calls.argtypes = [C.c_void_p]
        S.pfs_tape_calls.restype = C.c_uint64
        S.pfs_tape_time.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        S.pfs_tape_time.restype = C.c_int
        S.pfs_tape_time_loop.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        S.pfs_tape_time_loop.restype = C.c_int
        _scenes = S
    return _scenes


_apis = {}


def api(path=LIB_PATH):
    """Scene-driver binding (PFSapi*) of a library exporting the API."""
    path = os.path.abspath(path)
    if path not in _apis:
        a = scenes().pfs_bind(path.encode())
        if not a:
            raise RuntimeError(f"cannot bind PixelForge API of {path}: {scenes().pfs_dlerror()}")
        _apis[path] = a
    return _apis[path]


# Bytes per pixel of the render-target (format, type) pairs.
TARGET_BYTES = {(7, 0): 4, (6, 0): 3, (8, 0): 3, (6, 2): 2, (8, 2): 2, (7, 3): 2, (9, 3): 2, (7, 4): 2, (9, 4): 2,
                (5, 9): 4, (5, 10): 8, (6, 9): 6, (8, 9): 6, (6, 10): 12, (8, 10): 12,
                (7, 9): 8, (9, 9): 8, (7, 10): 16, (9, 10): 16}


def render(params, path=LIB_PATH):
    """Renders one scene through the library at `path`; returns
    (color, depth float32[H, W], error code): color is uint32[H, W] for an
    RGBA8 target, else the target's bytes as uint8[H, W, bytes per pixel]."""
    h, w = params.height, params.width
    bpp = TARGET_BYTES[(params.fbFormat, params.fbType)] if params.fbFormat > 0 else 4
    color = np.zeros((h * w + 16,), np.uint32) if bpp <= 4 else np.zeros(((h * w + 16) * bpp,), np.uint8)
    depth = np.zeros((h * w + 16,), np.float32)
    err = scenes().pfs_render(api(path), C.byref(params), color.ctypes.data, depth.ctypes.data)
    if params.fbFormat > 0:
        return color.view(np.uint8)[: h * w * bpp].reshape(h, w, bpp).copy(), depth[: h * w].reshape(h, w), err
    return color[: h * w].reshape(h, w), depth[: h * w].reshape(h, w), err


def render_batch(params_list, path=LIB_PATH):
    """Renders several scenes, one context each, submitted as one batch;
    returns (list of color uint32[H, W], error code)."""
    n = len(params_list)
    arr = (SceneParams * n)(*params_list)
    bufs = [np.zeros((p.height * p.width + 16,), np.uint32) for p in params_list]
    ptrs = (C.c_void_p * n)(*[b.ctypes.data for b in bufs])
    err = scenes().pfs_render_batch(api(path), arr, n, ptrs)
    return [b[: p.height * p.width].reshape(p.height, p.width) for b, p in zip(bufs, params_list)], err


class Tape:
    """A scene's API calls recorded once (csrc/pf_scenes.c, call tapes), to
    be replayed through a library: timing a replay measures the library, not
    the scene generator."""

    def __init__(self, params):
        self.params = params
        self.ptr = scenes().pfs_tape_make(C.byref(params))
        if not self.ptr:
            raise ValueError("scene kind or option not covered by call tapes")
        self.calls = int(scenes().pfs_tape_calls(self.ptr))

    def time(self, frames, path=LIB_PATH):
        """Renders the scene `frames` times through the library at `path`
        into an application colour buffer (pfClear + draws + pfFinish per
        frame); returns (color uint32[H, W] of the last frame, per-frame
        milliseconds, error code)."""
        h, w = self.params.height, self.params.width
        color = np.zeros((h * w + 16,), np.uint32)
        ms = np.zeros(frames, np.float64)
        err = scenes().pfs_tape_time(api(path), self.ptr, frames, color.ctypes.data, ms.ctypes.data)
        return color[: h * w].reshape(h, w), ms, err

    def time_loop(self, frames):
        """Per-frame milliseconds of replaying the frame's calls through
        no-op functions: the application loop's own share of time()."""
        ms = np.zeros(frames, np.float64)
        scenes().pfs_tape_time_loop(self.ptr, frames, ms.ctypes.data)
        return ms

    def close(self):
        if self.ptr:
            scenes().pfs_tape_free(self.ptr)
            self.ptr = None


def set_default_config(tile_shift=0, raster_mode=0):
    """Tile size (2**tile_shift) and raster mode of frames run by the
    default executor; tile_shift 0 (the initial setting) chooses both per
    frame."""
    c = PFDconfig()
    c.tileShift = tile_shift
    c.rasterMode = raster_mode
    rc = lib().pfdSetDefaultConfig(C.byref(c))
    if rc:
        raise RuntimeError(f"pfdSetDefaultConfig failed ({rc})")


def unpack(color):
    """uint32 RGBA8 image -> uint8 [H, W, 4] (r, g, b, a)."""
    return color.view(np.uint8).reshape(color.shape + (4,))


PF_RGBA, PF_UNSIGNED_BYTE = 7, 0


class Captured:
    """A recorded frame (or batch of frames) kept resident on the GPU as a
    PFDjob, with the host objects it refers to (contexts, target buffers,
    textures) kept alive until close()."""

    def __init__(self, job, ctxs, buffers, textures):
        self.job = C.c_void_p(job)
        self.ctxs, self.buffers, self.textures = ctxs, buffers, textures

    def targets(self):
        """The job's host render targets as (colour uint32[n], depth
        float32[n]) views of the buffers the job uploads and downloads."""
        L = lib()
        out = []
        for f in range(L.pfdJobFrameCount(self.job)):
            c, d, w, h = C.c_void_p(), C.c_void_p(), C.c_uint32(), C.c_uint32()
            if L.pfdJobTarget(self.job, f, C.byref(c), C.byref(d), C.byref(w), C.byref(h)):
                raise RuntimeError("pfdJobTarget failed")
            n = w.value * h.value
            out.append((np.ctypeslib.as_array(C.cast(c, C.POINTER(C.c_uint32)), (n,)),
                        np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_float)), (n,))))
        return out

    def snapshot(self):
        """Copies of the host render targets (to restore between runs)."""
        return [(c.copy(), d.copy()) for c, d in self.targets()]

    def restore(self, snap):
        """Writes a snapshot back into the host render targets and uploads
        them, so that the next run starts from the same pixels."""
        for (c, d), (sc, sd) in zip(self.targets(), snap):
            c[:] = sc
            d[:] = sd
        if lib().pfdJobUploadTargets(self.job):
            raise RuntimeError("pfdJobUploadTargets failed")

    def images(self):
        """Copies of the host colour targets after a download."""
        return [c.copy() for c, _ in self.targets()]

    def close(self):
        L = lib()
        if self.job.value:
            L.pfdJobDestroy(self.job)
            self.job = C.c_void_p(0)
        for ctx, (tex, px) in zip(self.ctxs, self.textures):
            L.pfMakeCurrent(ctx)
            scenes().pfs_free_texture(api(), tex, px)
            L.pfDeleteContext(ctx)
        L.pfMakeCurrent(None)
        self.ctxs, self.textures = [], []


def _record(params):
    L = lib()
    color = np.zeros(params.width * params.height + 16, np.uint32)
    ctx = L.pfCreateContext(color.ctypes.data, params.width, params.height, PF_RGBA, PF_UNSIGNED_BYTE)
    if not ctx:
        raise RuntimeError("pfCreateContext failed")
    L.pfMakeCurrent(ctx)
    tex, texpx = C.c_void_p(), C.c_void_p()
    err = scenes().pfs_record(api(), C.byref(params), C.byref(tex), C.byref(texpx))
    if err:
        raise RuntimeError(f"scene recording failed: {err}")
    return ctx, color, (tex, texpx)


def capture(params):
    """Records one scene through the API and returns it as a resident job."""
    L = lib()
    ctx, color, tex = _record(params)
    L.pfxSetExecutor(C.cast(L.pfdCaptureExecutor, C.c_void_p), None)
    try:
        L.pfFinish()
    finally:
        L.pfxSetExecutor(None, None)
    job = L.pfdTakeCapturedJob()
    if not job:
        raise RuntimeError(f"frame capture failed: {L.pfdLastError().decode()}")
    return Captured(job, [ctx], [color], [tex])


def capture_batch(params_list):
    """Records one context per scene and captures them, as pfSubmitBatch
    hands them over, into one resident job."""
    L = lib()
    recs = [_record(p) for p in params_list]
    ctxs = (C.c_void_p * len(recs))(*[r[0] for r in recs])
    L.pfxSetBatchExecutor(C.cast(L.pfdCaptureBatchExecutor, C.c_void_p), None)
    try:
        L.pfSubmitBatch(ctxs, len(recs))
    finally:
        L.pfxSetBatchExecutor(None, None)
    job = L.pfdTakeCapturedJob()
    if not job:
        raise RuntimeError(f"batch capture failed: {L.pfdLastError().decode()}")
    return Captured(job, [r[0] for r in recs], [r[1] for r in recs], [r[2] for r in recs])


def batch_workload(n=256, seed=4, max_side=1024, max_tris=50000, max_tex=512):
    """The batched workload: n independent contexts, framebuffer sides
    uniform in [64, max_side], triangle counts log-uniform in
    [100, max_tris] (jittered height-field grids about 10 cells wide per
    cell deep), RGBA8 noise textures with sides uniform in [64, max_tex]
    (bilinear, REPEAT), depth LEQUAL, back faces culled; every second
    context draws all its triangles with additive blending."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = (int(v) for v in rng.integers(64, max_side + 1, 2))
        tris = float(np.exp(rng.uniform(np.log(100.0), np.log(max_tris))))
        gz = max(1, int(round((tris / 20.0) ** 0.5)))
        gx = max(1, int(round(tris / (2.0 * gz))))
        tw, th = (int(v) for v in rng.integers(64, max_tex + 1, 2))
        add = i % 2 == 1
        out.append(scene(kind=HEIGHTFIELD, width=w, height=h, seed=1000 + i, gridX=gx, gridZ=gz,
                         texture=1, texW=tw, texH=th, wrap=PF_REPEAT, filter=PF_BILINEAR,
                         depthFunc=PF_LEQUAL, cull=PF_BACK, blend=1 if add else 0,
                         blendMode=PF_BLEND_ADD, blendFraction=1.0 if add else 0.0))
    return out
