"""The 1080p height-field workload's triangle screen areas.

The benchmark scene is specified as 200k triangles of a jittered height
field whose screen areas are log-normal with a median of 8 pixels.  The
oracle counts, per recorded triangle, the samples it covers (both culling
passes, clipped fans summed; pforacle_set_coverage), over the frame at its
full size: the median is 8 px and log(area) has the spread of a
log-normal with sigma near 1.
"""
import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402
from conftest import _oracle, render_oracle  # noqa: E402


def test_heightfield_screen_areas_are_log_normal_with_median_8px():
    p = bench.workload()
    n = 200000
    counts = np.zeros(n, np.uint32)
    _oracle.pforacle_set_coverage.argtypes = [C.c_void_p, C.c_uint32]
    _oracle.pforacle_set_coverage(counts.ctypes.data, n)
    try:
        _, _, err, frags = render_oracle(p)
    finally:
        _oracle.pforacle_set_coverage(None, 0)
    assert err == 0
    assert int(counts.sum()) == frags            # every covered sample belongs to one triangle
    a = counts[counts > 0].astype(np.float64)
    assert a.size > 0.7 * n                      # most triangles are on screen and front-facing
    assert np.median(a) == 8.0
    lg = np.log(a)
    assert 7.5 < np.exp(lg.mean()) < 9.0         # geometric mean: the log-normal's median
    assert 0.8 < lg.std() < 1.2                  # sigma of log(area)
    # quartiles of a log-normal sit at median * exp(-+0.674 sigma)
    q1, q3 = np.percentile(a, [25, 75])
    assert 3.0 <= q1 <= 5.0 and 13.0 <= q3 <= 18.0
