/*
 * pf_device.h -- the thin C ABI between the host library and the HIP
 * kernels of libpixelforge_amd.so.
 *
 * No reference counterpart: the reference rasterises on the CPU inside
 * pfVertex* (context.c:1671-1674).  Here a recorded frame (pf_frame.h) is
 * uploaded into a "job" resident in GPU memory and run by the kernel chain
 *   setup count -> block scan -> setup write (records + tile pairs) ->
 *   stable radix sort by tile -> tile ranges -> tile raster
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

#include <stddef.h>
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
 * that is binned and rasterised by one workgroup (3..6 -> 8x8 .. 64x64).
 * rasterMode: 0 packs the covered fragments of several triangles into full
 * wavefronts before shading; 1 shades each triangle in place on the lanes
 * it covers; 2 and 3 are 0 and 1 with four pixels per lane (a wave renders
 * 16x16 pixels per walk of the tile list; tiles of 8x8 fall back to 0/1).
 * Other values act as 0.
 * Zero fields take defaults.  A zero tileShift in the default configuration
 * (pfdSetDefaultConfig; the initial one) makes one-shot frames choose their
 * tiles from the first pass's counts: larger tiles and in-place shading
 * when triangles span many 8x8 tiles. */
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
    uint32_t clippedTriangles; /* recorded triangles that needed clipping     */
    uint32_t reserved;
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
 * library's default executor); initially automatic (tileShift 0). */
int pfdSetDefaultConfig(const PFDconfig* cfg);

/* One-shot execution used by the host library's default executor: uploads
 * the frame and the render target, runs it and downloads the target.  The
 * library keeps one job's streams and device buffers between calls
 * (buffers only grow), so a frame no larger than an earlier one creates no
 * stream and calls no hipMalloc or hipFree. */
int pfdExecuteFrame(const PFFframe* frame);

/* Same for several frames (distinct render targets) in one batched job. */
int pfdExecuteFrames(const PFFframe* frames, uint32_t count);

/* pfdExecuteFrame without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * around HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the frame -- its uploaded inputs and
 * render target, raster records, tile pairs, sort and scan buffers and
 * counters -- is carved from a static device array of PFD_SCRATCH_BYTES
 * (128 MiB) per GPU, which the library holds from its set-up (pfdInit)
 * until the process ends; it calls no hipMalloc or hipFree.  A frame that
 * needs more fails with PFD_E_NOMEM and a message naming the limit, before
 * its raster pass (a 1920x1080 frame of 200k textured triangles needs about
 * 75 MB).  Calls on one GPU share that array, which keeps the last call's
 * contents; overlapping calls would race on it, so the library serialises
 * the calls of a process with a lock.  Results equal pfdExecuteFrame's. */
#define PFD_SCRATCH_BYTES_MIB 128
int pfdExecuteFrameNoAlloc(const PFFframe* frame);

/* Host memory for the host library's own buffers: recorded triangle
 * streams, depth buffers and offscreen colour buffers (no reference
 * counterpart).  Page-locked when a GPU is present, so a frame's uploads
 * and downloads of them are direct DMA transfers; plain heap memory
 * otherwise.  pfdHostFree takes only pointers from pfdHostAlloc (NULL is
 * ignored). */
void* pfdHostAlloc(size_t bytes);
void  pfdHostFree(void* p);

/* Device copies of host bytes that outlive frames (no reference
 * counterpart): render lists keep their recorded vertices resident, and
 * frames that call them reference the copy (PFFarrays.device) instead of
 * uploading the vertices again. */
typedef struct PFDresident PFDresident;
int         pfdResidentCreate(const void* host, size_t bytes, PFDresident** out);
const void* pfdResidentPointer(const PFDresident* r);
void        pfdResidentDestroy(PFDresident* r);

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
/* Host render target of frame `frame` (the buffers pfdJobUploadTargets reads
 * and pfdJobDownload writes; colour pixels in the frame's PFF_TARGET_*
 * format): a caller replaying a job restores them between runs.  Any output
 * pointer may be NULL. */
int  pfdJobTarget(PFDjob* job, uint32_t frame, void** color, float** depth, uint32_t* width, uint32_t* height);
int  pfdJobFrameCount(PFDjob* job);               /* frames in the job, or < 0 */
int  pfdJobAutotune(PFDjob* job, PFDconfig* chosen);
void pfdJobDestroy(PFDjob* job);

/* ------------------------------------------------------------------------
 * Differentiable rendering (no reference counterpart).
 *
 * A differentiable job renders one frame through the integer pipeline
 * once, keeping for every pixel the raster triangle whose fragment was
 * written last (coverage and depth winners are then fixed: edges stay
 * hard), and shades those fragments again in floating point without
 * quantisation:
 *   colour  = sum_k b_k * corner colour_k (corner colours are parameters
 *             where the corner came from a colour array; times material
 *             diffuse / 255 when lit)
 *   texture = texel taps of the reference's index maps, bilinear weights
 *             (1-fx)(1-fy) .. fx fy, times the colour per channel
 *   Phong   = clamp(sum over lights of ambient + kd * diffuse * D +
 *             ks * specular * S, 0, 1), alpha 0, with exact sqrt and pow
 * and returns gradients of a scalar loss with respect to the texels of one
 * texture, the vertex colours and the lights' diffuse and specular colours
 * given dL/dimage.  Frames with blending, flat shading, Gouraud lighting
 * or a clear after the first command are rejected (PFD_E_INVALID).
 * ------------------------------------------------------------------------ */

typedef struct {
    int32_t      texture;        /* index into PFFframe.textures whose texels are parameters, -1: none */
    const float* texels;         /* [h][w][4] RGBA; NULL: that texture's own RGBA-float pixels */
    uint32_t     numVertices;    /* colour parameters, addressed by PFFtriangles.vertexId */
    const float* vertexColors;   /* [numVertices][4]; NULL: recorded RGBA8 / 255, no gradient */
    const float* lightDiffuse;   /* [8][3] by light number; NULL: recorded colours / 255 */
    const float* lightSpecular;  /* [8][3] */
} PFDdiffParams;

typedef struct {                 /* outputs; NULL members are not computed */
    float* texels;               /* [h][w][4] */
    float* vertexColors;         /* [numVertices][4] */
    float* lightDiffuse;         /* [8][3] */
    float* lightSpecular;        /* [8][3] */
} PFDdiffGrads;

typedef struct {
    float rasterMs;              /* integer pass (setup, binning, raster with winners) */
    float forwardMs;             /* float-domain forward */
    float backwardMs;            /* adjoint (gradient clears, scatter, light sums)     */
} PFDdiffTimes;

int pfdDiffCreate(const PFFframe* frame, PFDjob** out);
/* Float image [H][W][4] of the current parameters (host memory). */
int pfdDiffForward(PFDjob* job, const PFDdiffParams* params, float* image);
/* Gradients for dL/dimage [H][W][4] (host memory); outputs are overwritten. */
int pfdDiffBackward(PFDjob* job, const PFDdiffParams* params, const float* dImage, PFDdiffGrads* grads);
/* Mean device times of `iters` integer passes, forwards and adjoints on
 * resident data (HIP events on the job's stream). */
int pfdDiffTime(PFDjob* job, const PFDdiffParams* params, uint32_t iters, PFDdiffTimes* times);

/* Frame executor that, instead of running the frame, keeps it as a
 * resident job; pfdTakeCapturedJob() hands the last one over. */
int     pfdCaptureExecutor(const PFFframe* frame, void* user);
/* Same for the frames of one pfSubmitBatch call (one job, all frames). */
int     pfdCaptureBatchExecutor(const PFFframe* frames, uint32_t count, void* user);
PFDjob* pfdTakeCapturedJob(void);

#if defined(__cplusplus)
}
#endif

#endif /* PF_DEVICE_H */
