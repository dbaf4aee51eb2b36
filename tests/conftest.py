"""Shared fixtures: the native library, the CPU oracle and the compiled
reference, plus helpers that render one scene through each of them.

The oracle (oracle/_build/libpforacle.so) is the checker: it renders the
frames the host library records, through the library's executor hook, so a
GPU frame and its oracle frame come from the same recorded commands.  The
compiled reference (oracle/_ref/libpfref.so) renders the same scene through
its own API; it exists only where it was built from the reference sources.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import pixelforge_amd as pf  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libpforacle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    from pixelforge_amd import build
    if not build.built():
        build.build()
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)


_ensure_built()
_oracle = C.CDLL(ORACLE_SO)
_oracle.pforacle_pow.restype = C.c_float
_oracle.pforacle_pow.argtypes = [C.c_float, C.c_float]
_oracle.pforacle_blend.restype = C.c_uint32
_oracle.pforacle_blend.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
_oracle.pforacle_bary_smooth.restype = C.c_uint32
_oracle.pforacle_bary_smooth.argtypes = [C.c_uint32] * 3 + [C.c_float] * 3
_oracle.pforacle_sample.restype = C.c_uint32
_oracle.pforacle_sample.argtypes = [C.c_void_p, C.c_float, C.c_float]
_frag_counter = C.c_uint64(0)


@pytest.fixture(scope="session")
def oracle():
    return _oracle


class OracleExecutor:
    """Context manager routing the host library's frames to the oracle."""

    def __enter__(self):
        _frag_counter.value = 0
        pf.lib().pfxSetExecutor(C.cast(_oracle.pforacle_executor, C.c_void_p),
                                C.cast(C.pointer(_frag_counter), C.c_void_p))
        return self

    def __exit__(self, *exc):
        pf.lib().pfxSetExecutor(None, None)

    @property
    def fragments(self):
        return _frag_counter.value


def render_oracle(params):
    """Scene -> (color, depth, err, fragments) through host library + oracle."""
    with OracleExecutor() as ex:
        c, d, e = pf.render(params, pf.LIB_PATH)
        return c, d, e, ex.fragments


def render_gpu(params):
    """Scene -> (color, depth, err) through host library + HIP kernels."""
    pf.lib().pfxSetExecutor(None, None)
    return pf.render(params, pf.LIB_PATH)


def render_ref(params):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    return pf.render(params, REF_SO)


def have_ref():
    return os.path.exists(REF_SO)


def assert_same(a_color, a_depth, b_color, b_depth, what=""):
    """Bit-exact colour and depth (NaN patterns compared as bits)."""
    dc = np.count_nonzero(a_color != b_color)
    dd = np.count_nonzero(a_depth.view(np.uint32) != b_depth.view(np.uint32))
    assert dc == 0 and dd == 0, f"{what}: {dc} colour and {dd} depth words differ"


def parity_report(ref_color, ref_depth, color, depth):
    """North-star parity measures between a reference image and ours:
    fraction of pixels within +-1 LSB on every channel, max channel diff,
    fraction of pixels whose depth differs."""
    a = pf.unpack(ref_color).astype(np.int16)
    b = pf.unpack(color).astype(np.int16)
    d = np.abs(a - b).max(axis=2)
    return {
        "within1": float((d <= 1).mean()),
        "maxdiff": int(d.max()),
        "depth_mismatch": float((ref_depth.view(np.uint32) != depth.view(np.uint32)).mean()),
    }
