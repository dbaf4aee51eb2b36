"""Differentiable scene of the float-domain path: an indexed grid mesh with
float vertex colours, a float RGBA texture and Phong lights, issued through
the PixelForge API of libpixelforge_amd.so (pfDrawElements with float
colour, normal and texture-coordinate arrays), then rendered and
differentiated with pfDiffRender / pfDiffBackward.

Parameters are numpy arrays: texels [h, w, 4], vertex colours [V, 4], light
diffuse and specular [8, 3] by light number.  Used by the benchmark and the
tests; randomness from numpy's default_rng(seed).
"""
import ctypes as C

import numpy as np

import pixelforge_amd as pf

PF_FLOAT, PF_UINT = 10, 5
PF_TRIANGLES = 2


class DiffParams(C.Structure):          # PFdiffparams (include/pixelforge.h)
    _fields_ = [("texture", C.c_void_p), ("texels", C.c_void_p), ("numVertices", C.c_uint32),
                ("vertexColors", C.c_void_p), ("lightDiffuse", C.c_void_p), ("lightSpecular", C.c_void_p)]


class DiffGrads(C.Structure):           # PFdiffgrads
    _fields_ = [("texels", C.c_void_p), ("vertexColors", C.c_void_p), ("lightDiffuse", C.c_void_p),
                ("lightSpecular", C.c_void_p)]


def _f(*v):
    return [C.c_float(float(x)) for x in v]


class Scene:
    """Grid mesh of nx * nz cells (2 triangles each) facing a perspective
    camera, drawn with pfDrawElements; Phong with `lights` point lights."""

    def __init__(self, width=64, height=48, nx=8, nz=6, tex=(16, 16), lights=2, seed=1, filt=pf.PF_BILINEAR,
                 wrap=pf.PF_REPEAT, attenuation=False, spot=False, lit=True, textured=True, uv_scale=1.0):
        rng = np.random.default_rng(seed)
        self.W, self.H = width, height
        self.nv = (nx + 1) * (nz + 1)
        aspect = width / height
        i, j = np.meshgrid(np.arange(nx + 1), np.arange(nz + 1))
        x = (2.0 * i / nx - 1.0) * 1.25 * aspect
        y = (2.0 * j / nz - 1.0) * 1.25
        x = x + rng.uniform(-0.2, 0.2, x.shape) * (2.5 * aspect / nx) * ((i > 0) & (i < nx))
        y = y + rng.uniform(-0.2, 0.2, y.shape) * (2.5 / nz) * ((j > 0) & (j < nz))
        z = -2.4 + 0.15 * np.sin(2.1 * x) * np.cos(1.7 * y)
        self.pos = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
        nrm = np.stack([-0.3 * np.cos(2.1 * x), 0.25 * np.sin(1.7 * y), np.ones_like(x)], -1)
        self.nrm = (nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)).reshape(-1, 3).astype(np.float32)
        self.uv = (np.stack([i / nx, j / nz], -1).reshape(-1, 2) * uv_scale).astype(np.float32)
        self.colors = rng.uniform(0.2, 1.0, (self.nv, 4)).astype(np.float32)
        self.colors[:, 3] = 1.0
        idx = []
        for jj in range(nz):
            for ii in range(nx):
                v00 = jj * (nx + 1) + ii
                v10, v01 = v00 + 1, v00 + nx + 1
                idx += [v00, v10, v01 + 1, v00, v01 + 1, v01]
        self.idx = np.array(idx, np.uint32)
        self.texels = rng.uniform(0.0, 1.0, (tex[1], tex[0], 4)).astype(np.float32)
        self.light_pos = rng.uniform(-1.0, 1.0, (lights, 3)).astype(np.float32) + np.array([0, 0, -1.0], np.float32)
        self.ldiff = np.zeros((8, 3), np.float32)
        self.lspec = np.zeros((8, 3), np.float32)
        self.light_ids = [1 + 2 * k for k in range(lights)]          # non-contiguous light numbers
        for k, lid in enumerate(self.light_ids):
            self.ldiff[lid] = rng.uniform(0.2, 0.6, 3)
            self.lspec[lid] = rng.uniform(0.1, 0.4, 3)
        self.filt, self.wrap, self.att, self.spot = filt, wrap, attenuation, spot
        self.lit, self.textured = lit, textured
        self.ctx = None

    def record(self):
        """Creates the context and records the frame (not executed yet)."""
        L = pf.lib()
        self.buf = np.zeros(self.W * self.H + 16, np.uint32)
        self.ctx = L.pfCreateContext(self.buf.ctypes.data, self.W, self.H, 7, 0)
        assert self.ctx
        L.pfMakeCurrent(self.ctx)
        L.pfGenTexture.restype = C.c_void_p
        self.tex = L.pfGenTexture(C.c_void_p(self.texels.ctypes.data), self.texels.shape[1], self.texels.shape[0], 7,
                                  PF_FLOAT)
        assert self.tex
        L.pfTextureParameter(C.c_void_p(self.tex), self.wrap, self.filt)
        L.pfDisable(0x10)                                      # both culling passes
        L.pfEnable(0x8)                                        # depth test, LESS
        L.pfMatrixMode(1)
        L.pfLoadIdentity()
        a, n = self.W / self.H, 0.5
        L.pfFrustum(*_f(-a * 0.5 * n, a * 0.5 * n, -0.5 * n, 0.5 * n, n, 10.0))
        L.pfMatrixMode(0)
        L.pfLoadIdentity()
        L.pfRotatef(*_f(-10.0, 1.0, 0.0, 0.0))
        if self.textured:
            L.pfBindTexture(C.c_void_p(self.tex))
            L.pfEnable(0x1)
        if self.lit:
            L.pfEnable(0x40)
            L.pfLightModel(1)                                  # PF_PHONG
            for k, lid in enumerate(self.light_ids):
                L.pfLightfv(lid, 7, (C.c_float * 3)(*self.light_pos[k]))
                L.pfLightfv(lid, 3, (C.c_float * 3)(*self.ldiff[lid]))
                L.pfLightfv(lid, 4, (C.c_float * 3)(*self.lspec[lid]))
                if self.att:
                    L.pfLightf(lid, 13, C.c_float(0.3))
                    L.pfLightf(lid, 14, C.c_float(0.05))
                if self.spot:
                    d = -self.light_pos[k] + np.array([0, 0, -2.4], np.float32)
                    L.pfLightfv(lid, 8, (C.c_float * 3)(*d))
                    L.pfLightf(lid, 10, C.c_float(25.0))
                    L.pfLightf(lid, 11, C.c_float(40.0))
                L.pfEnableLight(lid)
            L.pfMaterialfv(2, 4, (C.c_float * 3)(0.8, 0.7, 0.9))    # PF_FRONT_AND_BACK, PF_SPECULAR
            L.pfMaterialf(2, 6, C.c_float(24.0))                   # PF_SHININESS
        L.pfClearColor(10, 20, 30, 255)
        L.pfClear(3)
        L.pfEnable(0x100 | 0x200 | 0x400 | 0x800)
        L.pfVertexPointer(3, PF_FLOAT, 0, C.c_void_p(self.pos.ctypes.data))
        L.pfNormalPointer(PF_FLOAT, 0, C.c_void_p(self.nrm.ctypes.data))
        L.pfTexCoordPointer(PF_FLOAT, 0, C.c_void_p(self.uv.ctypes.data))
        L.pfColorPointer(4, PF_FLOAT, 0, C.c_void_p(self.colors.ctypes.data))
        L.pfDrawElements(PF_TRIANGLES, len(self.idx), PF_UINT, C.c_void_p(self.idx.ctypes.data))
        assert L.pfGetError() == 0

    def rerecord(self):
        """Drops the current context and records the same frame again."""
        L = pf.lib()
        L.pfDeleteContext(C.c_void_p(self.ctx))
        L.pfDeleteTexture(C.byref(C.c_void_p(self.tex)), False)
        self.ctx = None
        self.record()

    def close(self):
        if self.ctx:
            L = pf.lib()
            L.pfDeleteContext(C.c_void_p(self.ctx))
            L.pfDeleteTexture(C.byref(C.c_void_p(self.tex)), False)
            L.pfMakeCurrent(None)
            self.ctx = None

    # -- GPU --------------------------------------------------------------
    def gpu_params(self, texels, colors, ldiff, lspec):
        keep = [np.ascontiguousarray(a, np.float32) for a in (texels, colors, ldiff, lspec)]
        p = DiffParams(self.tex if self.textured else None, keep[0].ctypes.data, self.nv, keep[1].ctypes.data,
                       keep[2].ctypes.data, keep[3].ctypes.data)
        return p, keep

    def gpu_forward(self, texels, colors, ldiff, lspec):
        L = pf.lib()
        L.pfMakeCurrent(C.c_void_p(self.ctx))
        p, keep = self.gpu_params(texels, colors, ldiff, lspec)
        img = np.zeros((self.H, self.W, 4), np.float32)
        L.pfDiffRender(C.byref(p), img.ctypes.data_as(C.c_void_p))
        err = L.pfGetError()
        assert err == 0, f"pfDiffRender error {err}"
        return img

    def gpu_backward(self, texels, colors, ldiff, lspec, dimg):
        L = pf.lib()
        L.pfMakeCurrent(C.c_void_p(self.ctx))
        p, keep = self.gpu_params(texels, colors, ldiff, lspec)
        dimg = np.ascontiguousarray(dimg, np.float32)
        gt = np.zeros_like(self.texels)
        gc = np.zeros_like(self.colors)
        gd = np.zeros((8, 3), np.float32)
        gs = np.zeros((8, 3), np.float32)
        g = DiffGrads(gt.ctypes.data, gc.ctypes.data, gd.ctypes.data, gs.ctypes.data)
        L.pfDiffBackward(C.byref(p), dimg.ctypes.data_as(C.c_void_p), C.byref(g))
        err = L.pfGetError()
        assert err == 0, f"pfDiffBackward error {err}"
        return gt, gc, gd, gs

    def gpu_time(self, texels, colors, ldiff, lspec, iters=20):
        """Mean device ms of the integer pass, the float forward and the
        adjoint on resident data (pfxDiffTime)."""
        L = pf.lib()
        L.pfMakeCurrent(C.c_void_p(self.ctx))
        p, keep = self.gpu_params(texels, colors, ldiff, lspec)
        r, f, b = C.c_float(), C.c_float(), C.c_float()
        rc = L.pfxDiffTime(C.byref(p), iters, C.byref(r), C.byref(f), C.byref(b))
        if rc:
            raise RuntimeError(f"pfxDiffTime failed ({rc}): {L.pfdLastError().decode()}")
        return r.value, f.value, b.value


def params_of(scene):
    """The scene's own parameter values in float64."""
    return scene.texels.astype(np.float64), scene.colors.astype(np.float64), \
        scene.ldiff.astype(np.float64), scene.lspec.astype(np.float64)
