"""FP64 oracle side of the differentiable test scenes
(pixelforge_amd/diffmesh.py): `oracle_job(scene)` records the scene's frame
and captures it into oracle/pf_oracle.c (pforacle_diff_*).
"""
import ctypes as C

import numpy as np

import pixelforge_amd as pf
from pixelforge_amd.diffmesh import Scene, params_of  # noqa: F401
from conftest import _oracle

LIGHT_VALS = 48


class OracleParams(C.Structure):        # PFOdiffParams (oracle/pf_oracle.c)
    _fields_ = [("texture", C.c_int32), ("texels", C.c_void_p), ("numVertices", C.c_uint32),
                ("vertexColors", C.c_void_p), ("lights", C.c_void_p)]


_oracle.pforacle_diff_prepare.restype = C.c_void_p
_oracle.pforacle_diff_prepare.argtypes = [C.c_void_p]
_oracle.pforacle_diff_free.argtypes = [C.c_void_p]
_oracle.pforacle_diff_eval.argtypes = [C.c_void_p, C.POINTER(OracleParams)] + [C.c_void_p] * 5
_oracle.pforacle_diff_winners.restype = C.c_void_p
_oracle.pforacle_diff_winners.argtypes = [C.c_void_p]
_oracle.pforacle_diff_color.restype = C.c_void_p
_oracle.pforacle_diff_color.argtypes = [C.c_void_p]



def oracle_job(scene):
    """Records the frame and captures it into the FP64 oracle."""
    scene.record()
    handles = []

    def run(fp, user):
        h = _oracle.pforacle_diff_prepare(fp)
        handles.append(h)
        return 0 if h else -1
    cb = pf.EXECUTOR(run)
    L = pf.lib()
    L.pfxSetExecutor(C.cast(cb, C.c_void_p), None)
    try:
        L.pfFinish()
    finally:
        L.pfxSetExecutor(None, None)
    assert len(handles) == 1 and handles[0]
    return OracleJob(handles[0], scene)


class OracleJob:
    def __init__(self, h, scene):
        self.h, self.s = h, scene

    def __del__(self):
        if self.h:
            _oracle.pforacle_diff_free(self.h)
            self.h = None

    def _params(self, texels, colors, ldiff, lspec):
        keep = [np.ascontiguousarray(texels, np.float64), np.ascontiguousarray(colors, np.float64),
                np.ascontiguousarray(np.concatenate([np.ravel(ldiff), np.ravel(lspec)]), np.float64)]
        return OracleParams(0 if self.s.textured else -1, keep[0].ctypes.data, self.s.nv, keep[1].ctypes.data,
                            keep[2].ctypes.data), keep

    def forward(self, texels, colors, ldiff, lspec):
        p, keep = self._params(texels, colors, ldiff, lspec)
        img = np.zeros((self.s.H, self.s.W, 4), np.float64)
        _oracle.pforacle_diff_eval(self.h, C.byref(p), img.ctypes.data, None, None, None, None)
        return img

    def backward(self, texels, colors, ldiff, lspec, dimg):
        p, keep = self._params(texels, colors, ldiff, lspec)
        dimg = np.ascontiguousarray(dimg, np.float64)
        gt = np.zeros(np.shape(texels), np.float64)
        gc = np.zeros(np.shape(colors), np.float64)
        gl = np.zeros(LIGHT_VALS, np.float64)
        _oracle.pforacle_diff_eval(self.h, C.byref(p), None, dimg.ctypes.data, gt.ctypes.data, gc.ctypes.data,
                                   gl.ctypes.data)
        return gt, gc, gl[:24].reshape(8, 3), gl[24:].reshape(8, 3)

    def winners(self):
        n = self.s.W * self.s.H
        return np.ctypeslib.as_array(C.cast(_oracle.pforacle_diff_winners(self.h), C.POINTER(C.c_int32)), (n,)).copy()

    def color(self):
        n = self.s.W * self.s.H
        return np.ctypeslib.as_array(C.cast(_oracle.pforacle_diff_color(self.h), C.POINTER(C.c_uint32)), (n,)).copy()


def target_of(scene, job, sigma=0.1, seed=99, color_sigma=0.0):
    """Image rendered from a perturbed texture (N(0, sigma) per texel) and,
    with color_sigma, perturbed vertex colours."""
    rng = np.random.default_rng(seed)
    t, c, d, s = params_of(scene)
    return job.forward(t + rng.normal(0.0, sigma, t.shape), c + rng.normal(0.0, color_sigma, c.shape), d, s)
