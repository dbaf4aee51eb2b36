/*
 * pf_ext.h -- entry points of libpixelforge_amd.so with no reference
 * counterpart, beyond pfFinish/pfFlush/pfSubmitBatch in pixelforge.h.
 */
#ifndef PF_EXT_H
#define PF_EXT_H

#include "pixelforge.h"
#include "pf_frame.h"

#if defined(__cplusplus)
extern "C" {
#endif

/* Runs a recorded frame; returns 0 on success. */
typedef int (*PFXexecutor)(const PFFframe* frame, void* user);

/* Replaces the frame executor.  The default (NULL) is the GPU executor,
 * pfdExecuteFrame.  A test harness installs its own to inspect the
 * recorded frames; the library never installs one by itself. */
PF_API void pfxSetExecutor(PFXexecutor fn, void* user);

/* Number of frames the current context has executed so far. */
PF_API PFuint pfxFrameCount(void);

/* Whether pfDrawElements / pfDrawArrays calls whose primitives all become
 * filled triangles are recorded whole, their triangles formed on the GPU
 * (the default), or vertex by vertex on the host as immediate-mode calls
 * are.  Both record the same triangles; a test harness compares them. */
PF_API void pfxSetArrayDraws(PFboolean enabled);

/* Runs the frames of one pfSubmitBatch call together; returns 0 on success. */
typedef int (*PFXbatchExecutor)(const PFFframe* frames, uint32_t count, void* user);

/* Replaces the executor of pfSubmitBatch.  The default (NULL) runs the
 * frames in one GPU job (pfdExecuteFrames), or one by one through the frame
 * executor when one is installed. */
PF_API void pfxSetBatchExecutor(PFXbatchExecutor fn, void* user);

/* Mean device times (ms) over `iters` runs of the current context's
 * differentiable frame (the last pfDiffRender): integer pass, float
 * forward, adjoint.  Returns 0 on success. */
PF_API int pfxDiffTime(const PFdiffparams* params, PFuint iters, float* rasterMs, float* forwardMs,
                       float* backwardMs);

#if defined(__cplusplus)
}
#endif

#endif /* PF_EXT_H */
