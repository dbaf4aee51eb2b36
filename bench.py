"""Benchmark of the fragment pipeline on one MI355X.

Three measurements, one JSON line (the headline is the first):

1. Fragments/s at 1080p (`value`).  One step = one full frame:
   1920x1080 RGBA8 + float32 depth, 200,000 triangles of a seeded jittered
   height-field grid (1000 x 100 cells, two triangles each) seen through a
   perspective camera; 1024x1024 RGBA8 uniform-noise texture, bilinear,
   REPEAT; depth test LEQUAL; 30% of the triangles (seeded choice) drawn
   after the opaque ones with alpha 128 and PF_BLEND_ALPHA; back faces
   culled; the frame starts with pfClear.  Fragments are covered samples
   (before the depth test) counted by the raster kernel.
2. Batched frames/s (`batched`): 256 independent contexts of uneven size
   (sides uniform in [64, 1024]), triangle counts log-uniform in
   [100, 50000], RGBA8 textures with sides in [64, 512], every second
   context additively blended; all contexts render in the same launches
   (pfSubmitBatch), tile size autotuned over 8x8 .. 64x64.
3. Differentiable (`differentiable`): 512x512, 20,000-triangle grid mesh,
   256x256x4 float texels, per-vertex float colours and two lights'
   diffuse/specular as parameters; L2 loss against a target rendered from
   the texture perturbed by N(0, 0.1); device time of forward and adjoint,
   and the gradients' relative error against the FP64 oracle.

Scenes are recorded once through the PixelForge C API into jobs that stay
resident on the GPU (inputs in HBM); each step replays the whole kernel
chain (setup, binning, tile raster) as one HIP graph.

`end_to_end` times the 1080p frame as an application of the reference sees
it: pfClear, the scene's API calls and pfFinish (upload of the recorded
frame, the kernels, colour and depth copied back), replayed from a call
tape so that the scene generator is not timed; `cpu_baseline` times the
reference on the same tape.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu]
"""
import argparse
import ctypes as C
import hashlib
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
from pixelforge_amd.diffmesh import Scene, params_of  # noqa: E402

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfref.so")
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libpforacle.so")
PEAK_HBM_GBS = 8000.0
SETUP_BYTES_PER_TRI = 88          # position 3x16 + colour 3x4 + texcoord 3x8 + info 4
RASTER_MODES = ["packed", "in-place", "packed, 4 px/lane", "in-place, 4 px/lane"]


def workload(width=1920, height=1080, seed=2):
    return pf.scene(kind=pf.HEIGHTFIELD, width=width, height=height, seed=seed, gridX=1000, gridZ=100,
                    texture=1, texW=1024, texH=1024, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR,
                    depthFunc=pf.PF_LEQUAL, blendFraction=0.3, cull=pf.PF_BACK)


def measured_traffic(key):
    """HBM bytes per dispatch of a kernel configuration, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by a copy of known
    size (profiles/traffic.json, written from profiles/*.md); None if that
    configuration was not profiled."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f).get(key)
        return None if t is None else t["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {pf.lib().pfdLastError().decode()}")


def configure(job, tile_shift, raster_mode):
    lib = pf.lib()
    cfg = pf.PFDconfig()
    if tile_shift:
        cfg.tileShift = tile_shift
        cfg.rasterMode = raster_mode
        check(lib.pfdJobSetConfig(job, C.byref(cfg)), "config")
    else:
        check(lib.pfdJobAutotune(job, C.byref(cfg)), "autotune")
    return cfg


def time_replays(job, steps, warmup):
    lib = pf.lib()
    for _ in range(max(warmup, 2)):
        check(lib.pfdJobLaunch(job), "launch")
    check(lib.pfdJobSync(job), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        check(lib.pfdJobLaunch(job), "launch")
    check(lib.pfdJobSync(job), "sync")
    return time.perf_counter() - t0


def verified_replays(cap, steps, warmup):
    """Runs the job once from its captured targets (sizing run, then the
    first graph replay from the same targets), then times `steps` graph
    replays after `warmup` more, and downloads the frames again: the timed
    replays must leave exactly the first replay's images, or the line
    fails.  Returns (seconds, first images, sha1 of the first images)."""
    lib = pf.lib()
    job = cap.job
    snap = cap.snapshot()
    for _ in range(2):
        cap.restore(snap)
        check(lib.pfdJobLaunch(job), "launch")
        check(lib.pfdJobSync(job), "sync")
    check(lib.pfdJobDownload(job), "download")
    first = cap.images()
    elapsed = time_replays(job, steps, warmup)
    check(lib.pfdJobDownload(job), "download")
    for i, (a, b) in enumerate(zip(first, cap.images())):
        if not np.array_equal(a, b):
            raise RuntimeError(f"frame {i}: the timed replays left {np.count_nonzero(a != b)} pixels different "
                               "from the first replay")
    cap.restore(snap)
    digest = hashlib.sha1(b"".join(a.tobytes() for a in first)).hexdigest()[:16]
    return elapsed, first, digest


def oracle_check(params, image):
    """cpu_baseline leg: the oracle renders the same recorded frame; True
    when the GPU's first replay equals it bit for bit."""
    if not os.path.exists(ORACLE_SO):
        return None
    o = C.CDLL(ORACLE_SO)
    lib = pf.lib()
    lib.pfxSetExecutor(C.cast(o.pforacle_executor, C.c_void_p), None)
    try:
        ref, _, err = pf.render(params)
    finally:
        lib.pfxSetExecutor(None, None)
    return err == 0 and bool(np.array_equal(ref.ravel(), image))


def phase_times(job, runs):
    lib = pf.lib()
    raster, setup, binning = [], [], []
    for _ in range(runs):
        t = pf.PFDtimes()
        check(lib.pfdJobRunTimed(job, C.byref(t)), "timed run")
        raster.append(t.rasterMs)
        setup.append(t.setupMs)
        binning.append(t.binMs)
    return statistics.mean(setup), statistics.mean(binning), statistics.mean(raster)


def stats(job):
    st = pf.PFDstats()
    check(pf.lib().pfdJobStats(job, C.byref(st)), "stats")
    return st


def cpu_frame_rate(params_list, frames=3):
    """The reference's own CPU path (oracle/_ref) on the given scenes,
    replayed from call tapes (the scene generator is not timed): median
    seconds per pass over all of them, each frame being pfClear and the
    scene's draws, after one warm-up pass."""
    tapes = [pf.Tape(p) for p in params_list]
    try:
        per = np.zeros((len(tapes), frames))
        for i, t in enumerate(tapes):
            _, ms, err = t.time(frames + 1, REF_SO)
            if err:
                raise RuntimeError(f"reference frame failed ({err})")
            per[i] = ms[1:]
        return float(np.median(per.sum(axis=0))) * 1e-3
    finally:
        for t in tapes:
            t.close()


def end_to_end(params, frames, expect):
    """The drop-in path as an application sees it: per frame, the scene's
    API calls (pfClear, state, pfBegin / pfColor / pfTexCoord / pfVertex /
    pfEnd) then pfFinish, which uploads the recorded frame, runs it and
    copies colour and depth back into the application's buffers.  Replayed
    from a call tape, so the scene generator is not timed.  The first frame
    is a warm-up.  The host part alone is timed with an executor that
    discards the frame, and the application's own loop (decoding the tape,
    calling through function pointers) by replaying the tape through no-op
    functions: library_ms is the frame time less that loop.  The last frame
    must equal `expect` (the resident job's image)."""
    tape = pf.Tape(params)
    try:
        img, ms, err = tape.time(frames + 1)
        check(err, "end-to-end frame")
        if not np.array_equal(img.ravel(), expect):
            raise RuntimeError(f"end-to-end frame differs from the resident job's on "
                               f"{np.count_nonzero(img.ravel() != expect)} pixels")
        lib = pf.lib()
        cb = pf.EXECUTOR(lambda frame, user: 0)
        lib.pfxSetExecutor(C.cast(cb, C.c_void_p), None)
        try:
            _, rec_ms, err = tape.time(frames + 1)
        finally:
            lib.pfxSetExecutor(None, None)
        check(err, "recording")
        loop_ms = tape.time_loop(frames + 1)
        e2e_ms = float(np.median(ms[1:]))
        app_ms = float(np.median(loop_ms[1:]))
        return {"ms_per_frame": e2e_ms, "frames": frames,
                "host_record_ms": float(np.median(rec_ms[1:])), "api_calls_per_frame": tape.calls,
                "app_loop_ms": app_ms, "library_ms": e2e_ms - app_ms,
                "equals_resident_job": True}
    finally:
        tape.close()


# ---------------------------------------------------------------------------
# 1. 1080p frame
# ---------------------------------------------------------------------------

def bench_frame(args):
    params = workload()
    t0 = time.perf_counter()
    cap = pf.capture(params)
    rec_s = time.perf_counter() - t0
    job = cap.job
    lib = pf.lib()
    check(lib.pfdJobUploadTargets(job), "upload")
    cfg = configure(job, args.tile_shift, args.raster_mode)
    elapsed, first, digest = verified_replays(cap, args.steps, args.warmup)
    st = stats(job)
    frags = int(st.fragments)
    setup_ms, bin_ms, raster_ms = phase_times(job, args.profile_runs)
    W, H = params.width, params.height
    raster_bytes = W * H * 8 + 1024 * 1024 * 4      # colour + depth written once, texels read once
    achieved = raster_bytes / (raster_ms * 1e-3) / 1e9
    out = {
        "value": frags * args.steps / elapsed,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "config": {"workload": "1920x1080 RGBA8+f32 depth, 200k-triangle jittered height-field grid, "
                               "1024^2 RGBA8 noise texture bilinear REPEAT, LEQUAL, 30% alpha-blended",
                   "fragments_per_frame": frags, "raster_triangles": int(st.rasterTriangles),
                   "tile_pairs": int(st.tilePairs), "tile": 1 << int(st.tileShift),
                   "max_tile_list": int(st.maxTileList), "busy_tiles": int(st.busyTiles),
                   "raster_mode": RASTER_MODES[cfg.rasterMode], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(f"k_raster 1080p tile {1 << int(st.tileShift)} mode {cfg.rasterMode}"),
                     "kernel": "k_raster", "kernel_ms": raster_ms, "algorithmic_bytes": raster_bytes},
        "phases_ms": {"setup": setup_ms, "binning": bin_ms, "raster": raster_ms},
        "record_s": rec_s,
        "replays": {"timed_equal_first": True, "frame_sha1": digest},
    }
    if not args.no_cpu:
        out["replays"]["first_equals_oracle"] = oracle_check(params, first[0])
    e2e = end_to_end(params, args.e2e_frames, first[0])
    e2e["fragments_per_s"] = frags / (e2e["ms_per_frame"] * 1e-3)
    out["end_to_end"] = e2e
    if not args.no_cpu and os.path.exists(REF_SO):
        med = cpu_frame_rate([params], frames=5)
        out["cpu_baseline"] = {
            "value": frags / med, "unit": "fragments/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "reference",
            "sample": f"1 warm-up + median of 5 frames of the same scene replayed from a call tape (pfClear + "
                      f"draws; scene generation not timed), reference built -O3 -mavx2 -fopenmp; "
                      f"median {med * 1e3:.1f} ms/frame"}
        e2e["cpu_reference_ms_per_frame"] = med * 1e3
        e2e["speedup_vs_reference"] = med * 1e3 / e2e["ms_per_frame"]
    else:
        out["cpu_baseline"] = None
    cap.close()
    return out


# ---------------------------------------------------------------------------
# 1b. the other single-frame configurations
# ---------------------------------------------------------------------------

def config1():
    """800x600, 10,000 triangles with vertices uniform in NDC [-1, 1]^3 and
    the default matrices (the reference's 2D path), random RGBA8 vertex
    colours interpolated (smooth), depth test LESS, culling disabled, no
    texture, no blending; the frame starts with pfClear."""
    return pf.scene(kind=pf.RANDOM_NDC, width=800, height=600, ntris=10000, seed=1, cull=-1,
                    depthFunc=pf.PF_LESS)


def config3():
    """3840x2160, 256 spheres of 31 x 63 x 2 = 3,906 triangles (999,936 in
    all) with per-vertex normals, lit per fragment (Phong) by 4 point
    lights, shininess 32; 512x512 RGBA8 texture, NEAREST, CLAMP_TO_EDGE;
    back faces culled, depth test LESS; the frame starts with pfClear."""
    return pf.scene(kind=pf.SPHERES, width=3840, height=2160, spheres=256, sphereLat=31, sphereLon=63, seed=3,
                    lighting=2, numLights=4, texture=1, texW=512, texH=512, wrap=pf.PF_CLAMP_TO_EDGE,
                    filter=pf.PF_NEAREST, cull=pf.PF_BACK)


def bench_config(params, workload, tex_bytes, args, cpu_frames, e2e_frames):
    """Resident replays of one frame (autotuned), phase times, the raster
    kernel's roofline, the end-to-end frame and the reference's CPU time on
    the same scene (call tape)."""
    cap = pf.capture(params)
    job = cap.job
    lib = pf.lib()
    check(lib.pfdJobUploadTargets(job), "upload")
    cfg = configure(job, 0, 0)
    elapsed, first, digest = verified_replays(cap, args.config_steps, 3)
    st = stats(job)
    frags = int(st.fragments)
    setup_ms, bin_ms, raster_ms = phase_times(job, max(3, args.profile_runs // 4))
    W, H = params.width, params.height
    raster_bytes = W * H * 8 + tex_bytes
    achieved = raster_bytes / (raster_ms * 1e-3) / 1e9
    ms = elapsed * 1e3 / args.config_steps
    out = {
        "value": frags / (ms * 1e-3), "unit": "fragments/s", "ms_per_frame": ms, "steps": args.config_steps,
        "config": {"workload": workload, "fragments_per_frame": frags, "raster_triangles": int(st.rasterTriangles),
                   "tile_pairs": int(st.tilePairs), "tile": 1 << int(st.tileShift),
                   "raster_mode": RASTER_MODES[cfg.rasterMode]},
        "phases_ms": {"setup": setup_ms, "binning": bin_ms, "raster": raster_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "kernel": "k_raster", "kernel_ms": raster_ms,
                     "algorithmic_bytes": raster_bytes},
        "replays": {"timed_equal_first": True, "frame_sha1": digest},
    }
    if e2e_frames:
        e2e = end_to_end(params, e2e_frames, first[0])
        e2e["fragments_per_s"] = frags / (e2e["ms_per_frame"] * 1e-3)
        out["end_to_end"] = e2e
    if not args.no_cpu and os.path.exists(REF_SO):
        med = cpu_frame_rate([params], frames=cpu_frames)
        out["cpu_baseline"] = {
            "value": frags / med, "unit": "fragments/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "reference",
            "sample": f"1 warm-up + median of {cpu_frames} frames of the same scene replayed from a call tape "
                      f"(scene generation not timed), reference built -O3 -mavx2 -fopenmp; "
                      f"median {med * 1e3:.1f} ms/frame"}
        if e2e_frames:
            out["end_to_end"]["cpu_reference_ms_per_frame"] = med * 1e3
            out["end_to_end"]["speedup_vs_reference"] = med * 1e3 / out["end_to_end"]["ms_per_frame"]
    else:
        out["cpu_baseline"] = None
    cap.close()
    return out


# ---------------------------------------------------------------------------
# 2. batched contexts
# ---------------------------------------------------------------------------

def bench_batch(args):
    params = pf.batch_workload(args.contexts)
    cap = pf.capture_batch(params)
    job = cap.job
    lib = pf.lib()
    check(lib.pfdJobUploadTargets(job), "upload")
    cfg = configure(job, args.batch_tile_shift, args.batch_raster_mode)
    elapsed, first, digest = verified_replays(cap, args.batch_steps, 3)
    st = stats(job)
    setup_ms, bin_ms, raster_ms = phase_times(job, max(3, args.profile_runs // 4))
    ms = elapsed * 1e3 / args.batch_steps
    px = sum(p.width * p.height for p in params)
    tex = sum(p.texW * p.texH * 4 for p in params)
    tris = int(st.inputTriangles)
    algo = px * 8 + tex + tris * SETUP_BYTES_PER_TRI
    out = {
        "value": len(params) * args.batch_steps / elapsed, "unit": "frames/s", "contexts": len(params),
        "replays": {"timed_equal_first": True, "frames_sha1": digest},
        "ms_per_batch": ms, "fragments": int(st.fragments), "input_triangles": tris,
        "pixels": px, "tile": 1 << int(st.tileShift), "raster_mode": RASTER_MODES[cfg.rasterMode],
        "phases_ms": {"setup": setup_ms, "binning": bin_ms, "raster": raster_ms},
        "roofline": {"bound": "hbm", "achieved": algo / (ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": algo / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, "scope": "whole batch step",
                     "algorithmic_bytes": algo,
                     "raster_kernel": {"achieved": (px * 8 + tex) / (raster_ms * 1e-3) / 1e9, "kernel_ms": raster_ms,
                                       "algorithmic_bytes": px * 8 + tex,
                                       "traffic": measured_traffic(f"k_raster batch{len(params)} tile "
                                                                   f"{1 << int(st.tileShift)} mode {cfg.rasterMode}")}},
    }
    if not args.no_cpu:
        out["replays"]["first_equals_oracle"] = all(oracle_check(p, first[i]) for i, p in
                                                    enumerate(params[: args.cpu_contexts]))
    if not args.no_cpu and os.path.exists(REF_SO):
        sample = params[: args.cpu_contexts]
        med = cpu_frame_rate(sample, frames=3)
        out["cpu_baseline"] = {
            "value": len(sample) / med, "unit": "frames/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "reference",
            "sample": f"first {len(sample)} of the {len(params)} contexts, one after another (the reference has no "
                      f"batched path), replayed from call tapes (scene generation not timed), 1 warm-up + median "
                      f"of 3; {med * 1e3:.0f} ms per pass"}
    else:
        out["cpu_baseline"] = None
    cap.close()
    return out


# ---------------------------------------------------------------------------
# 3. differentiable path
# ---------------------------------------------------------------------------

class _OracleParams(C.Structure):        # PFOdiffParams of the FP64 oracle
    _fields_ = [("texture", C.c_int32), ("texels", C.c_void_p), ("numVertices", C.c_uint32),
                ("vertexColors", C.c_void_p), ("lights", C.c_void_p)]


def _fp64_gradients(scene, p64, target):
    """cpu_baseline leg: the FP64 oracle's forward + adjoint of the same
    frame, timed, and its gradients (to measure the GPU's error)."""
    o = C.CDLL(ORACLE_SO)
    o.pforacle_diff_prepare.restype = C.c_void_p
    o.pforacle_diff_prepare.argtypes = [C.c_void_p]
    o.pforacle_diff_eval.argtypes = [C.c_void_p, C.POINTER(_OracleParams)] + [C.c_void_p] * 5
    o.pforacle_diff_free.argtypes = [C.c_void_p]
    handles = []
    cb = pf.EXECUTOR(lambda fp, user: (handles.append(o.pforacle_diff_prepare(fp)), 0)[1])
    scene.record()
    L = pf.lib()
    L.pfxSetExecutor(C.cast(cb, C.c_void_p), None)
    try:
        L.pfFinish()
    finally:
        L.pfxSetExecutor(None, None)
    h = handles[0]
    t, c, d, s = (np.ascontiguousarray(a, np.float64) for a in p64)
    lights = np.concatenate([d.ravel(), s.ravel()])
    prm = _OracleParams(0, t.ctypes.data, scene.nv, c.ctypes.data, lights.ctypes.data)
    t0 = time.perf_counter()
    img = np.zeros((scene.H, scene.W, 4))
    o.pforacle_diff_eval(h, C.byref(prm), img.ctypes.data, None, None, None, None)
    dimg = 2.0 * (img - target)
    g = [np.zeros_like(t), np.zeros_like(c), np.zeros(48)]
    o.pforacle_diff_eval(h, C.byref(prm), None, dimg.ctypes.data, g[0].ctypes.data, g[1].ctypes.data, g[2].ctypes.data)
    dt = time.perf_counter() - t0
    o.pforacle_diff_free(h)
    return img, (g[0], g[1], g[2][:24].reshape(8, 3), g[2][24:].reshape(8, 3)), dt


def bench_diff(args):
    s = Scene(width=512, height=512, nx=100, nz=100, tex=(256, 256), lights=2, seed=5)
    p64 = params_of(s)
    p32 = [a.astype(np.float32) for a in p64]
    rng = np.random.default_rng(99)
    target_tex = (p64[0] + rng.normal(0.0, 0.1, p64[0].shape)).astype(np.float32)
    s.record()
    target = s.gpu_forward(target_tex, *p32[1:]).astype(np.float64)
    img = s.gpu_forward(*p32)
    grads = s.gpu_backward(*p32, (2.0 * (img - target)).astype(np.float32))
    raster_ms, fwd_ms, bwd_ms = s.gpu_time(*p32, iters=args.diff_iters)
    out = {
        "config": {"workload": "512x512, 20k-triangle grid mesh (pfDrawElements), 256x256x4 f32 texels, "
                               "10201 RGBA f32 vertex colours, 2 Phong lights' diffuse+specular; L2 loss vs a "
                               "target from the texture perturbed by N(0, 0.1)"},
        "integer_pass_ms": raster_ms, "forward_ms": fwd_ms, "adjoint_ms": bwd_ms,
        "fwd_bwd_ms": raster_ms + fwd_ms + bwd_ms,
        "fwd_bwd_over_fwd": (raster_ms + fwd_ms + bwd_ms) / (raster_ms + fwd_ms),
        "loss": float(((img - target) ** 2).sum()),
    }
    s.close()
    if not args.no_cpu and os.path.exists(ORACLE_SO):
        s2 = Scene(width=512, height=512, nx=100, nz=100, tex=(256, 256), lights=2, seed=5)
        img64, g64, dt = _fp64_gradients(s2, p64, target)
        s2.close()
        rel = {}
        for name, a, b in zip(("texels", "vertex_colors", "light_diffuse", "light_specular"), grads, g64):
            n = np.linalg.norm(b)
            rel[name] = float(np.linalg.norm(np.asarray(a, np.float64) - b) / n) if n else 0.0
        out["grad_rel_err_vs_fp64"] = rel
        out["image_max_abs_err_vs_fp64"] = float(np.abs(img - img64).max())
        out["cpu_baseline"] = {"value": dt * 1e3, "unit": "ms (forward + adjoint, FP64)", "cores": 1, "kind": "port",
                               "sample": "the FP64 oracle's float-domain forward and adjoint of the same frame "
                                         "(coverage from its integer pass, not timed)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-shift", type=int, default=0, help="3..6 fixes the tile size; 0 autotunes")
    ap.add_argument("--raster-mode", type=int, default=0, help="with --tile-shift: raster mode (pf_device.h)")
    ap.add_argument("--batch-tile-shift", type=int, default=0, help="the same for the batched job")
    ap.add_argument("--batch-raster-mode", type=int, default=0)
    ap.add_argument("--profile-runs", type=int, default=20)
    ap.add_argument("--contexts", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--cpu-contexts", type=int, default=16)
    ap.add_argument("--diff-iters", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--e2e-frames", type=int, default=10, help="timed end-to-end frames (API calls + pfFinish)")
    ap.add_argument("--no-configs", action="store_true", help="skip the 800x600 and 4K single-frame workloads")
    ap.add_argument("--config-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    check(pf.lib().pfdInit(), "pfdInit")
    frame = bench_frame(args)
    print(f"1080p: {frame['value']:.4g} fragments/s, {frame['ms_per_step']:.3f} ms/frame; end to end "
          f"{frame['end_to_end']['ms_per_frame']:.2f} ms/frame", file=sys.stderr, flush=True)
    out = {
        "metric": "shaded fragments/s at 1080p (1 GPU) + batched frames/s; fwd+bwd time at 1e-3 grad err",
        "value": frame["value"],
        "unit": "fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": frame["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": frame["config"],
        "roofline": frame["roofline"],
        "cpu_baseline": frame["cpu_baseline"],
        "phases_ms": frame["phases_ms"],
        "record_s": frame["record_s"],
        "end_to_end": frame["end_to_end"],
    }
    if not args.no_configs:
        out["config1"] = bench_config(config1(), "800x600, 10k random NDC triangles (2D path), smooth colours, LESS, "
                                      "culling off", 0, args, cpu_frames=3, e2e_frames=args.e2e_frames)
        print(f"config 1: {out['config1']['value']:.4g} fragments/s, {out['config1']['ms_per_frame']:.3f} ms/frame",
              file=sys.stderr, flush=True)
        out["config3"] = bench_config(config3(), "3840x2160, 256 spheres (1M triangles), Phong with 4 point lights, "
                                      "512^2 RGBA8 NEAREST CLAMP, back faces culled", 512 * 512 * 4, args,
                                      cpu_frames=2, e2e_frames=0)
        print(f"config 3: {out['config3']['value']:.4g} fragments/s, {out['config3']['ms_per_frame']:.3f} ms/frame",
              file=sys.stderr, flush=True)
    if not args.no_batch:
        out["batched"] = bench_batch(args)
        print(f"batched: {out['batched']['value']:.4g} frames/s", file=sys.stderr, flush=True)
    if not args.no_diff:
        out["differentiable"] = bench_diff(args)
        print(f"differentiable: {out['differentiable']['fwd_bwd_ms']:.3f} ms fwd+bwd", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
