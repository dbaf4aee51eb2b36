"""Benchmark of the fragment pipeline on one MI355X.

Workload (one "step" = one full frame through the hot path):
  1920x1080 RGBA8 + float32 depth, 200,000 triangles of a seeded jittered
  height-field grid (1000 x 100 cells, two triangles each) seen through a
  perspective camera; 1024x1024 RGBA8 uniform-noise texture, bilinear,
  REPEAT; depth test LEQUAL; 30% of the triangles (seeded choice) drawn
  after the opaque ones with alpha 128 and PF_BLEND_ALPHA; back faces
  culled; the frame starts with pfClear.

The scene is recorded once through the PixelForge C API into a frame that
stays resident on the GPU (inputs in HBM); each step replays the whole
kernel chain (setup, binning, tile raster) as one HIP graph.  The metric is
covered fragments (before the depth test, counted by the raster kernel)
per second of wall time over the timed steps.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline build of the reference uses OpenMP; fix its thread count
# before its runtime loads.
_CPU_THREADS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
os.environ.setdefault("OMP_NUM_THREADS", str(min(_CPU_THREADS, 16)))

import numpy as np  # noqa: E402

import pixelforge_amd as pf  # noqa: E402

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")
PEAK_HBM_GBS = 8000.0
PF_RGBA, PF_UNSIGNED_BYTE = 7, 0


def workload(width=1920, height=1080, seed=2):
    return pf.scene(kind=pf.HEIGHTFIELD, width=width, height=height, seed=seed, gridX=1000, gridZ=100,
                    texture=1, texW=1024, texH=1024, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR,
                    depthFunc=pf.PF_LEQUAL, blendFraction=0.3, cull=pf.PF_BACK)


def record_job(params):
    """Records the scene through the API and returns a resident GPU job."""
    lib = pf.lib()
    n = params.width * params.height
    color = np.zeros(n + 16, np.uint32)
    ctx = lib.pfCreateContext(color.ctypes.data, params.width, params.height, PF_RGBA, PF_UNSIGNED_BYTE)
    if not ctx:
        raise RuntimeError("pfCreateContext failed")
    lib.pfMakeCurrent(ctx)
    tex, texpx = C.c_void_p(), C.c_void_p()
    err = pf.scenes().pfs_record(pf.api(), C.byref(params), C.byref(tex), C.byref(texpx))
    if err:
        raise RuntimeError(f"scene recording failed: {err}")
    lib.pfxSetExecutor(C.cast(lib.pfdCaptureExecutor, C.c_void_p), None)
    lib.pfFinish()
    lib.pfxSetExecutor(None, None)
    job = C.c_void_p(lib.pfdTakeCapturedJob())
    if not job.value:
        raise RuntimeError(f"frame capture failed: {lib.pfdLastError().decode()}")
    return job, (ctx, color, tex, texpx)


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {pf.lib().pfdLastError().decode()}")


def cpu_baseline(params, frags, frames=5):
    """The reference's own CPU path (oracle/_ref) on the same scene."""
    if not os.path.exists(REF_SO):
        return None
    pf.render(params, REF_SO)                       # warm-up
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        pf.render(params, REF_SO)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": frags / med, "unit": "fragments/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "reference",
            "sample": f"1 warm-up + median of {frames} full frames of the same scene (API calls incl. scene "
                      f"generation), reference built -O3 -mavx2 -fopenmp; median {med * 1e3:.1f} ms/frame"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-shift", type=int, default=0, help="3..6 fixes the tile size; 0 autotunes")
    ap.add_argument("--raster-mode", type=int, default=0, help="with --tile-shift: 0 packed, 1 in-place")
    ap.add_argument("--profile-runs", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    lib = pf.lib()
    check(lib.pfdInit(), "pfdInit")
    params = workload()
    t0 = time.perf_counter()
    job, keep = record_job(params)
    rec_s = time.perf_counter() - t0
    check(lib.pfdJobUploadTargets(job), "upload")

    cfg = pf.PFDconfig()
    if args.tile_shift:
        cfg.tileShift = args.tile_shift
        cfg.rasterMode = args.raster_mode
        check(lib.pfdJobSetConfig(job, C.byref(cfg)), "config")
    else:
        check(lib.pfdJobAutotune(job, C.byref(cfg)), "autotune")

    for _ in range(max(args.warmup, 2)):
        check(lib.pfdJobLaunch(job), "launch")
    check(lib.pfdJobSync(job), "sync")

    t0 = time.perf_counter()
    for _ in range(args.steps):
        check(lib.pfdJobLaunch(job), "launch")
    check(lib.pfdJobSync(job), "sync")
    elapsed = time.perf_counter() - t0

    st = pf.PFDstats()
    check(lib.pfdJobStats(job, C.byref(st)), "stats")
    frags = int(st.fragments)
    value = frags * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps

    # Per-kernel durations (HIP events on the job's stream).
    raster, setup, binning = [], [], []
    for _ in range(args.profile_runs):
        t = pf.PFDtimes()
        check(lib.pfdJobRunTimed(job, C.byref(t)), "timed run")
        raster.append(t.rasterMs)
        setup.append(t.setupMs)
        binning.append(t.binMs)
    raster_ms = statistics.mean(raster)
    W, H = params.width, params.height
    tex_bytes = 1024 * 1024 * 4
    raster_bytes = W * H * 8 + tex_bytes            # colour + depth written once, texels read once
    achieved = raster_bytes / (raster_ms * 1e-3) / 1e9

    out = {
        "metric": "shaded fragments/s at 1080p (1 GPU)",
        "value": value,
        "unit": "fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "1920x1080 RGBA8+f32 depth, 200k-triangle jittered height-field grid, "
                               "1024^2 RGBA8 noise texture bilinear REPEAT, LEQUAL, 30% alpha-blended",
                   "fragments_per_frame": frags, "raster_triangles": int(st.rasterTriangles),
                   "tile_pairs": int(st.tilePairs), "tile": 1 << int(st.tileShift),
                   "max_tile_list": int(st.maxTileList), "busy_tiles": int(st.busyTiles),
                   "raster_mode": ["packed", "in-place"][cfg.rasterMode],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "k_raster_packed" if cfg.rasterMode == 0 else "k_raster",
                     "kernel_ms": raster_ms, "algorithmic_bytes": raster_bytes},
        "phases_ms": {"setup": statistics.mean(setup), "binning": statistics.mean(binning), "raster": raster_ms},
        "record_s": rec_s,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(params, frags)
    else:
        out["cpu_baseline"] = None
    lib.pfdJobDestroy(job)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
