/*
 * pf_device.h -- the thin C ABI between the host library and the HIP
 * kernels of libpixelforge_amd.so.
 *
 * No reference counterpart: the reference rasterises on the CPU inside
 * pfVertex* (context.c:1671-1674).  Here a recorded frame (pf_frame.h) is
 * uploaded into a "job" resident in GPU memory and run by the kernel chain
 *   setup -> pair count/scan -> pair emit -> tile sort -> tile raster
 * captured once as a HIP graph.  A job holds one or more frames (one per
 * render target); several frames in one job render in the same launches
 * (the batched path of pfSubmitBatch).
 *
 * Plain C types only.  Every function returns 0 on success and a negative
 * PFD_E* code on failure; pfdLastError() gives a message.  Nothing here
 * falls back to the CPU: without a usable GPU every entry point fails.
 */
#ifndef PF_DEVICE_H
#define PF_DEVICE_H

#include <stdint.h>
#include "pf_frame.h"

#if defined(__cplusplus)
extern "C" {
#endif

#define PFD_OK          0
#define PFD_E_NODEVICE (-1)   /* no GPU, or HIP runtime failure at init   */
#define PFD_E_NOMEM    (-2)   /* device allocation failed                 */
#define PFD_E_INVALID  (-3)   /* malformed frame or argument              */
#define PFD_E_LAUNCH   (-4)   /* a kernel launch or HIP call failed       */

typedef struct PFDjob PFDjob;

/* Tile configuration of a job.  tileShift: log2 of the square screen tile
 * one workgroup rasterises (3..6 -> 8x8 .. 64x64) -- also the binning
 * granularity.  rasterMode: 0 packs covered fragments of several triangles
 * into full wavefronts before shading; 1 shades each triangle in place on
 * the lanes it covers (better when triangles cover whole 8x8 blocks).
 * Zero fields take defaults. */
typedef struct {
    uint32_t tileShift;
    uint32_t rasterMode;
    uint32_t reserved[6];
} PFDconfig;

/* Per-job counters, filled by pfdJobStats after a run. */
typedef struct {
    uint64_t rasterTriangles;  /* triangles emitted by setup (after cull/clip) */
    uint64_t tilePairs;        /* (triangle, tile) pairs binned               */
    uint64_t fragments;        /* covered samples before the depth test       */
    uint64_t inputTriangles;
    uint32_t numTiles;
    uint32_t tileShift;
    uint64_t deviceBytes;      /* device memory held by the job               */
    uint32_t maxTileList;      /* longest per-tile pair list (last pass)      */
    uint32_t busyTiles;        /* tiles with a non-empty list (last pass)     */
} PFDstats;

/* Per-kernel timing of one run, in milliseconds (HIP events on the job's
 * stream). */
typedef struct {
    float setupMs;             /* setup count + scan + setup write            */
    float binMs;               /* pair count/scan/emit + sort                 */
    float rasterMs;            /* tile raster (clear fused)                   */
    float totalMs;
} PFDtimes;

/* Device initialisation; idempotent.  Returns PFD_E_NODEVICE without a GPU. */
int         pfdInit(void);
const char* pfdLastError(void);

/* Configuration of jobs created without one (pfdExecuteFrame(s), the host
 * library's default executor). */
int pfdSetDefaultConfig(const PFDconfig* cfg);

/* One-shot execution used by the host library's default executor: uploads
 * the frame and the render target, runs it and downloads the target. */
int pfdExecuteFrame(const PFFframe* frame);

/* Same for several frames (distinct render targets) in one batched job. */
int pfdExecuteFrames(const PFFframe* frames, uint32_t count);

/* Resident jobs, for replay and measurement. */
int  pfdJobCreate(const PFFframe* frames, uint32_t count, const PFDconfig* cfg, PFDjob** out);
int  pfdJobSetConfig(PFDjob* job, const PFDconfig* cfg);
int  pfdJobUploadTargets(PFDjob* job);           /* host target buffers -> device */
int  pfdJobLaunch(PFDjob* job);                  /* async, on the job's stream    */
int  pfdJobSync(PFDjob* job);
int  pfdJobDownload(PFDjob* job);                /* device -> host target buffers */
int  pfdJobRunTimed(PFDjob* job, PFDtimes* t);   /* one run with per-phase events */
int  pfdJobTimeReplays(PFDjob* job, uint32_t iters, float* msTotal); /* graph replays */
int  pfdJobStats(PFDjob* job, PFDstats* st);
int  pfdJobAutotune(PFDjob* job, PFDconfig* chosen);
void pfdJobDestroy(PFDjob* job);

/* Frame executor that, instead of running the frame, keeps it as a
 * resident job; pfdTakeCapturedJob() hands the last one over. */
int     pfdCaptureExecutor(const PFFframe* frame, void* user);
PFDjob* pfdTakeCapturedJob(void);

#if defined(__cplusplus)
}
#endif

#endif /* PF_DEVICE_H */
