/*
 * pf_host.c -- host side of the drop-in: the PixelForge C API state machine.
 *
 * Mirrors the reference context (context.c, texture.c, framebuffer.c,
 * internal/context/context.c) call for call, except that a completed
 * triangle is appended to the context's recorded frame (pf_frame.h) instead
 * of being rasterised on the spot.  Frames are executed by the frame
 * executor -- the GPU (pfdExecuteFrame) unless a harness installed another
 * -- at the synchronisation points listed in pixelforge.h.
 *
 * Floating-point evaluation order follows the reference operation by
 * operation (no contraction: built with -ffp-contract=off) so that matrices,
 * texture coordinates and normals are bit-identical to the reference's.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pixelforge.h"
#include "pf_frame.h"
#include "pf_ext.h"
#include "pf_device.h"

#define PF_PI_D        3.14159265358979323846      /* PFM_PI, pfm.h:59 */
#define PF_DEG2RAD_D   (PF_PI_D / 180.0)           /* PFM_DEG2RAD, pfm.h:62 */
#define MAX_PROJ_STACK 2                           /* config.h:25-27 */
#define MAX_MV_STACK   8                           /* config.h:29-31 */
#define MAX_TEX_STACK  4                           /* config.h:33-35 */
#define MAX_LIGHTS     8                           /* config.h:37-39 */

/* ------------------------------------------------------------------------ */
/* Objects                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {                       /* struct PFItex, context.h:160-176 */
    void*           pixels;
    uint32_t        w, h;
    PFpixelformat   format;
    PFdatatype      type;
    PFtexturewrap   wrap;
    PFtexturefilter filter;
    int             hostPixels;        /* pixels from pfdHostAlloc (pfGenFramebuffer) */
} Tex;

typedef struct {                       /* PFIlight, context.h:216-229 */
    float   position[3], direction[3];
    float   innerCutOff, outerCutOff;
    float   attConstant, attLinear, attQuadratic;
    PFcolor ambient, diffuse, specular;
    int     next;                      /* index of the next active light, -1 */
} Light;

typedef struct {                       /* PFImaterial, context.h:234-240 */
    PFcolor ambient, diffuse, specular, emission;
    float   shininess;
} Material;

typedef struct {                       /* PFIvertex, context.h:203-210 */
    float   position[4];
    float   normal[3];
    float   texcoord[2];
    PFcolor color;
    uint32_t vid;                      /* colour-array index, PFF_NO_VERTEX */
} Vert;

typedef struct {                       /* PFIvertexattribbuffer, context.h:242-247 */
    const void* buffer;
    PFsizei     stride;                /* stored, never used (context.c:1285) */
    PFint       size;
    PFenum      type;
} Attrib;

typedef struct {                       /* PFIrendercall, context.h:270-279 */
    Material    faceMaterial[2];
    Tex*        texture;
    PFdrawmode  drawMode;
    float*      pos;                   /* 4 floats per vertex */
    float*      uv;                    /* 2 floats per vertex */
    float*      nrm;                   /* 3 floats per vertex */
    PFcolor*    col;
    size_t      n, cap;
    void*       packed;                /* sealed: pos, nrm, uv, col in one block */
    PFDresident* resident;             /* sealed: its device copy, or NULL */
} ListCall;

/* The calls of a list once pfEndList has sealed them: immutable, shared by
 * the list and the pending frames that call it (reference counted, so that
 * pfNewList / pfDeleteList while a frame is pending leave it intact). */
typedef struct {
    ListCall*   calls;
    size_t      n;
    int         refs;
} ListStore;

typedef struct {                       /* PFIrenderlist, context.h:298 */
    ListCall*   calls;
    size_t      n, cap;
    ListStore*  store;                 /* sealed calls (owns `calls`), or NULL */
} List;

typedef struct {                       /* PFIctxbackup, context.h:300-316 */
    Material    faceMaterial[2];
    float       uv[2], nrm[3];
    PFcolor     color;
    Tex*        texture;
    uint32_t    state;
} Backup;

typedef struct {                       /* the frame being recorded */
    void*      targetColor;
    float*     targetDepth;
    uint32_t   targetW, targetH, targetKind;
    float*     pos;  uint32_t* col;  float* uv;  float* nrm;  uint32_t* info;  uint32_t* vid;
    size_t     ntri, captri;
    PFFstate*  states;   size_t nstates, capstates;
    PFFtexture* texs;    const Tex** texKeys; size_t ntex, captex;
    PFFcmd*    cmds;     size_t ncmd, capcmd;
    PFFpixels* pix;      size_t npix, cappix;   /* pfDrawPixels calls; images owned */
    PFFpoint*  pts;      size_t npts, cappts;
    PFFline*   lns;      size_t nlns, caplns;
    PFFarrays* arr;      size_t narr, caparr;   /* array draws; element copies owned */
    uint8_t*   arrOwned;  size_t capowned;      /* per array draw: 1 if its block is owned */
    ListStore** stores;  size_t nstores, capstores;   /* lists called, one reference each */
    int        anyTex, anyLight, anyVid;
} Rec;

typedef struct {
    PFframebuffer  mainFB;
    PFframebuffer* boundFB;
    PFframebuffer* curFB;
    void*          auxBuffer;

    int32_t  vpPos[2];  uint32_t vpDim[2];  int32_t vpMin[2], vpMax[2];

    uint32_t       state;
    PFdepthmode    depthFunc;
    PFblendmode    blendMode;
    PFshademode    shadeModel;
    PFlightmode    lightMode;
    PFface         cullFace;
    PFpolygonmode  polyMode[2];
    PFcolor        clearColor;
    float          clearDepth;

    float matProj[16], matTex[16], matModel[16], matView[16], matMVP[16], matNormal[16];
    float stackProj[MAX_PROJ_STACK][16], stackMV[MAX_MV_STACK][16], stackTex[MAX_TEX_STACK][16];
    uint32_t nProj, nMV, nTex;
    PFmatrixmode matMode;
    float*   curMat;
    int      modelUsed;

    Light    lights[MAX_LIGHTS];
    int      activeHead;               /* -1: no active light */
    Material mat[2];
    PFface   followFace;
    PFenum   followMode;

    float    curNormal[3], curUV[2];
    PFcolor  curColor;
    PFdrawmode drawMode;
    Vert     vb[6];
    uint32_t vcount;
    Tex*     curTex;
    Attrib   aPos, aNrm, aUV, aCol;    /* vertex arrays (context.c:1161-1229) */
    List*    curList;                  /* list being recorded (pfNewList), NULL */
    Backup   backup;                   /* one slot, shared by pfNewList and pfCallList */

    float    pointSize, lineWidth;     /* out-of-path state, kept for the getters */
    float    pixelZoom[2], rasterPos[4];
    int32_t  fogMode;                  /* PFIfog, context.h:250-256 */
    float    fogDensity, fogStart, fogEnd;
    PFcolor  fogColor;

    PFerrcode err;
    Rec      rec;
    int      stateDirty;
    uint32_t texGen;
    uint32_t frames;

    PFDjob*    diffJob;                /* last frame rendered by pfDiffRender */
    const Tex** diffTex;               /* its textures, by frame texture index */
    uint32_t   diffNumTex;
} Ctx;

/* Current context, one per thread.  Initial-exec TLS: every API call reads
 * it, and the default model for a shared library calls __tls_get_addr. */
static __thread Ctx* G __attribute__((tls_model("initial-exec"))) = NULL;
static PFXexecutor g_exec = NULL;
static void* g_execUser = NULL;
static PFXbatchExecutor g_batchExec = NULL;
static void* g_batchUser = NULL;
static uint32_t g_texGen = 1;          /* bumped by pfTextureParameter */
static int g_arrayDraws = 1;           /* pfxSetArrayDraws: record array draws whole */

/* ------------------------------------------------------------------------ */
/* Matrix helpers (pfm.h); same evaluation order as the reference.           */
/* ------------------------------------------------------------------------ */

static void m4_identity(float* d)
{
    memset(d, 0, 16 * sizeof(float));
    d[0] = d[5] = d[10] = d[15] = 1.0f;
}

/* pfmMat4Mul (pfm.h:2176-2198): dst[i*4+j] = sum_k L[i*4+k] * R[k*4+j]. */
static void m4_mul(float* dst, const float* L, const float* R)
{
    float r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += L[i * 4 + k] * R[k * 4 + j];
            r[i * 4 + j] = s;
        }
    memcpy(dst, r, sizeof r);
}

static void m4_transpose(float* d, const float* s)     /* pfm.h:2053-2066 */
{
    float r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r[i * 4 + j] = s[j * 4 + i];
    memcpy(d, r, sizeof r);
}

/* pfmMat4Invert (pfm.h:2082-2122): cofactor expansion through the twelve
 * 2x2 minors of the upper and lower row pairs. */
static void m4_invert(float* d, const float* s)
{
    float a00 = s[0],  a01 = s[1],  a02 = s[2],  a03 = s[3];
    float a10 = s[4],  a11 = s[5],  a12 = s[6],  a13 = s[7];
    float a20 = s[8],  a21 = s[9],  a22 = s[10], a23 = s[11];
    float a30 = s[12], a31 = s[13], a32 = s[14], a33 = s[15];
    float b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10;
    float b02 = a00 * a13 - a03 * a10, b03 = a01 * a12 - a02 * a11;
    float b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    float b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30;
    float b08 = a20 * a33 - a23 * a30, b09 = a21 * a32 - a22 * a31;
    float b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    float inv = 1.0f / (b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06);
    d[0]  = (a11 * b11 - a12 * b10 + a13 * b09) * inv;
    d[1]  = (-a01 * b11 + a02 * b10 - a03 * b09) * inv;
    d[2]  = (a31 * b05 - a32 * b04 + a33 * b03) * inv;
    d[3]  = (-a21 * b05 + a22 * b04 - a23 * b03) * inv;
    d[4]  = (-a10 * b11 + a12 * b08 - a13 * b07) * inv;
    d[5]  = (a00 * b11 - a02 * b08 + a03 * b07) * inv;
    d[6]  = (-a30 * b05 + a32 * b02 - a33 * b01) * inv;
    d[7]  = (a20 * b05 - a22 * b02 + a23 * b01) * inv;
    d[8]  = (a10 * b10 - a11 * b08 + a13 * b06) * inv;
    d[9]  = (-a00 * b10 + a01 * b08 - a03 * b06) * inv;
    d[10] = (a30 * b04 - a31 * b02 + a33 * b00) * inv;
    d[11] = (-a20 * b04 + a21 * b02 - a23 * b00) * inv;
    d[12] = (-a10 * b09 + a11 * b07 - a12 * b06) * inv;
    d[13] = (a00 * b09 - a01 * b07 + a02 * b06) * inv;
    d[14] = (-a30 * b03 + a31 * b01 - a32 * b00) * inv;
    d[15] = (a20 * b03 - a21 * b01 + a22 * b00) * inv;
}

static void m4_ortho(float* d, float l, float r, float b, float t, float n, float f) /* pfm.h:2419-2440 */
{
    memset(d, 0, 16 * sizeof(float));
    float rl = r - l, tb = t - b, fn = f - n;
    d[0] = 2.0f / rl;
    d[5] = 2.0f / tb;
    d[10] = -2.0f / fn;
    d[11] = 0.0f;
    d[12] = -(l + r) / rl;
    d[13] = -(t + b) / tb;
    d[14] = -(f + n) / fn;
    d[15] = 1.0f;
}

static void m4_frustum(float* d, float l, float r, float b, float t, float n, float f) /* pfm.h:2375-2392 */
{
    memset(d, 0, 16 * sizeof(float));
    float rl = r - l, tb = t - b, fn = f - n;
    d[0] = (n * 2.0f) / rl;
    d[5] = (n * 2.0f) / tb;
    d[8] = (r + l) / rl;
    d[9] = (t + b) / tb;
    d[10] = -(f + n) / fn;
    d[11] = -1.0f;
    d[14] = -(f * n * 2.0f) / fn;
}

/* ------------------------------------------------------------------------ */
/* Errors and small helpers                                                  */
/* ------------------------------------------------------------------------ */

static void set_err(PFerrcode e) { if (G) G->err = e; }

static uint32_t pack(PFcolor c)
{
    return (uint32_t)c.r | ((uint32_t)c.g << 8) | ((uint32_t)c.b << 16) | ((uint32_t)c.a << 24);
}

static PFubyte f2u8(float f)           /* (PFubyte)(float) as gcc emits it on x86-64 */
{
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return 0;
    return (PFubyte)(uint32_t)(int32_t)f;
}

static PFcolor rgb_from_f(const float* v)     /* context.c:951-975: alpha forced to 255 */
{
    PFcolor c = { f2u8(v[0] * 255.0f), f2u8(v[1] * 255.0f), f2u8(v[2] * 255.0f), 255 };
    return c;
}

static int grow(void** p, size_t* cap, size_t need, size_t elem)
{
    if (need <= *cap) return 1;
    size_t nc = *cap ? *cap : 256;
    while (nc < need) nc *= 2;
    void* q = realloc(*p, nc * elem);
    if (!q) return 0;
    *p = q;
    *cap = nc;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Recording                                                                 */
/* ------------------------------------------------------------------------ */

static void store_release(ListStore* st);

static void rec_reset(Rec* r)
{
    for (size_t i = 0; i < r->npix; i++) free((void*)r->pix[i].pixels);
    for (size_t i = 0; i < r->narr; i++)
        if (r->arrOwned[i]) free((void*)r->arr[i].position);             /* one block per draw */
    r->narr = 0;
    for (size_t i = 0; i < r->nstores; i++) store_release(r->stores[i]);
    r->nstores = 0;
    r->npix = 0;
    r->npts = 0;
    r->nlns = 0;
    r->ntri = r->nstates = r->ntex = r->ncmd = 0;
    r->anyTex = r->anyLight = r->anyVid = 0;
    r->targetColor = NULL;
    r->targetDepth = NULL;
}

static void rec_free(Rec* r)
{
    for (size_t i = 0; i < r->npix; i++) free((void*)r->pix[i].pixels);
    for (size_t i = 0; i < r->narr; i++)
        if (r->arrOwned[i]) free((void*)r->arr[i].position);
    for (size_t i = 0; i < r->nstores; i++) store_release(r->stores[i]);
    free(r->stores);
    free(r->arrOwned);
    free(r->arr);
    free(r->pix);
    free(r->pts);
    free(r->lns);
    pfdHostFree(r->pos); pfdHostFree(r->col); pfdHostFree(r->uv); pfdHostFree(r->nrm);
    pfdHostFree(r->info); pfdHostFree(r->vid);
    free(r->states); free(r->texs); free(r->texKeys); free(r->cmds);
    memset(r, 0, sizeof *r);
}

static Tex* fb_tex(const PFframebuffer* fb) { return (Tex*)fb->texture; }

static int exec_frame(const PFFframe* f)
{
    if (g_exec) return g_exec(f, g_execUser);
    return pfdExecuteFrame(f);
}

static void build_frame(Ctx* c, PFFframe* f)
{
    Rec* r = &c->rec;
    memset(f, 0, sizeof *f);
    f->width = r->targetW;
    f->height = r->targetH;
    f->color = r->targetColor;
    f->depth = r->targetDepth;
    f->target = r->targetKind;
    f->tris.count = (uint32_t)r->ntri;
    f->tris.position = r->pos;
    f->tris.color = r->col;
    f->tris.texcoord = r->anyTex ? r->uv : NULL;
    f->tris.normal = r->anyLight ? r->nrm : NULL;
    f->tris.info = r->info;
    f->tris.vertexId = r->anyVid ? r->vid : NULL;
    f->states = r->states;
    f->numStates = (uint32_t)r->nstates;
    f->textures = r->texs;
    f->numTextures = (uint32_t)r->ntex;
    f->cmds = r->cmds;
    f->numCmds = (uint32_t)r->ncmd;
    f->pixelOps = r->pix;
    f->numPixelOps = (uint32_t)r->npix;
    f->points = r->pts;
    f->numPoints = (uint32_t)r->npts;
    f->lines = r->lns;
    f->numLines = (uint32_t)r->nlns;
    f->arrays = r->arr;
    f->numArrays = (uint32_t)r->narr;
}

/* Executes the recorded frame, if any, and starts an empty one. */
static void submit(Ctx* c)
{
    if (!c || c->rec.ncmd == 0) { if (c) rec_reset(&c->rec); return; }
    PFFframe f;
    build_frame(c, &f);
    int rc = exec_frame(&f);
    if (rc != 0) {
        c->err = (rc == PFD_E_NOMEM) ? PF_ERROR_OUT_OF_MEMORY : PF_INVALID_OPERATION;
        if (!g_exec) fprintf(stderr, "pixelforge_amd: frame execution failed (%d): %s\n", rc, pfdLastError());
    }
    c->frames++;
    rec_reset(&c->rec);
    c->stateDirty = 1;
}

/* Makes sure the recorded frame targets the current framebuffer; a frame
 * for another target is executed first.  Returns 0 if there is no valid
 * render target. */
static int ensure_target(Ctx* c)
{
    Tex* t = fb_tex(c->curFB);
    if (!t || !t->pixels || !c->curFB->zbuffer) return 0;
    Rec* r = &c->rec;
    /* The frame being recorded already targets these buffers (a format
     * change of the same buffers goes through pfSetMainBuffer, which
     * submits the frame first). */
    if (r->ncmd && r->targetColor == t->pixels && r->targetDepth == c->curFB->zbuffer
        && r->targetW == t->w && r->targetH == t->h)
        return 1;
    const uint32_t kind = pff_target_of(PFF_SOURCE(t->format, t->type));
    if (kind == PFF_TARGET_NONE) return 0;
    if (r->ncmd) submit(c);
    if (r->ncmd == 0) {
        r->targetColor = t->pixels;
        r->targetDepth = c->curFB->zbuffer;
        r->targetW = t->w;
        r->targetH = t->h;
        r->targetKind = kind;
    }
    return 1;
}

static int rec_texture(Ctx* c, const Tex* t)
{
    Rec* r = &c->rec;
    for (size_t i = 0; i < r->ntex; i++)
        if (r->texKeys[i] == t && r->texs[i].wrap == (uint32_t)t->wrap
            && r->texs[i].filter == (uint32_t)t->filter && r->texs[i].pixels == t->pixels)
            return (int)i;
    size_t cap = r->captex;
    if (!grow((void**)&r->texs, &cap, r->ntex + 1, sizeof(PFFtexture))) return -1;
    cap = r->captex;
    if (!grow((void**)&r->texKeys, &cap, r->ntex + 1, sizeof(Tex*))) return -1;
    r->captex = cap;
    PFFtexture* e = &r->texs[r->ntex];
    memset(e, 0, sizeof *e);
    e->pixels = t->pixels;
    e->width = t->w;
    e->height = t->h;
    if (t->format == PF_RGBA && t->type == PF_UNSIGNED_BYTE) e->format = PFF_TEX_RGBA8;
    else if (t->format == PF_RGBA && t->type == PF_FLOAT) e->format = PFF_TEX_RGBAF32;
    else { e->format = PFF_TEX_SOURCE; e->source = PFF_SOURCE(t->format, t->type); }
    e->wrap = t->wrap;
    e->filter = t->filter;
    r->texKeys[r->ntex] = t;
    return (int)r->ntex++;
}

/* Snapshot of everything pfiProcessRasterize_TRIANGLE_IMPL and
 * Rasterize_Triangle read from the context (triangles.c:62-116, 287-396). */
static int current_state(Ctx* c)
{
    Rec* r = &c->rec;
    if (!c->stateDirty && r->nstates && c->texGen == g_texGen) return (int)r->nstates - 1;
    PFFstate s;
    memset(&s, 0, sizeof s);
    memcpy(s.mvp, c->matMVP, sizeof s.mvp);
    memcpy(s.matNormal, c->matNormal, sizeof s.matNormal);
    memcpy(s.viewRow2, c->matView + 8, 3 * sizeof(float));
    int lit = (c->state & PF_LIGHTING) && c->activeHead >= 0;
    if (lit) {                                           /* triangles.c:82-86 */
        float inv[16];
        m4_invert(inv, c->matView);
        memcpy(s.viewPos, inv + 12, 3 * sizeof(float));
    }
    memcpy(s.vpPos, c->vpPos, sizeof s.vpPos);
    memcpy(s.vpDim, c->vpDim, sizeof s.vpDim);
    memcpy(s.vpMin, c->vpMin, sizeof s.vpMin);
    memcpy(s.vpMax, c->vpMax, sizeof s.vpMax);
    s.flags = c->state & (PFF_TEXTURE_2D | PFF_BLEND | PFF_DEPTH_TEST);
    if (lit) s.flags |= PFF_LIGHTING;
    s.depthFunc = c->depthFunc;
    s.blendMode = c->blendMode;
    s.shadeModel = c->shadeModel;
    s.lightMode = c->lightMode;
    s.texture = -1;
    if ((c->state & PF_TEXTURE_2D) && c->curTex) {
        s.texture = rec_texture(c, c->curTex);
        if (s.texture < 0) return -1;
        r->anyTex = 1;
    }
    for (int f = 0; f < 2; f++) {
        s.material[f].ambient = pack(c->mat[f].ambient);
        s.material[f].diffuse = pack(c->mat[f].diffuse);
        s.material[f].specular = pack(c->mat[f].specular);
        s.material[f].emission = pack(c->mat[f].emission);
        s.material[f].shininess = c->mat[f].shininess;
    }
    if (lit) {
        r->anyLight = 1;
        for (int li = c->activeHead; li >= 0 && s.numLights < MAX_LIGHTS; li = c->lights[li].next) {
            const Light* L = &c->lights[li];
            PFFlight* d = &s.lights[s.numLights++];
            memcpy(d->position, L->position, sizeof d->position);
            memcpy(d->direction, L->direction, sizeof d->direction);
            d->innerCutOff = L->innerCutOff;
            d->outerCutOff = L->outerCutOff;
            d->attConstant = L->attConstant;
            d->attLinear = L->attLinear;
            d->attQuadratic = L->attQuadratic;
            d->ambient = pack(L->ambient);
            d->diffuse = pack(L->diffuse);
            d->specular = pack(L->specular);
            d->id = (uint32_t)li;
        }
    }
    if (r->nstates && memcmp(&r->states[r->nstates - 1], &s, sizeof s) == 0) {
        c->stateDirty = 0;
        c->texGen = g_texGen;
        return (int)r->nstates - 1;
    }
    if (!grow((void**)&r->states, &r->capstates, r->nstates + 1, sizeof(PFFstate))) return -1;
    r->states[r->nstates] = s;
    c->stateDirty = 0;
    c->texGen = g_texGen;
    return (int)r->nstates++;
}

static void push_cmd(Ctx* c, const PFFcmd* cmd)
{
    Rec* r = &c->rec;
    if (!grow((void**)&r->cmds, &r->capcmd, r->ncmd + 1, sizeof(PFFcmd))) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    r->cmds[r->ncmd++] = *cmd;
}

/* realloc for pfdHostAlloc memory: *p grows from `used` to `bytes`. */
static int host_grow(void** p, size_t used, size_t bytes)
{
    void* q = pfdHostAlloc(bytes);
    if (!q) return 0;
    if (*p && used) memcpy(q, *p, used);
    pfdHostFree(*p);
    *p = q;
    return 1;
}

/* Grows the triangle streams to hold at least `need` triangles.  They live
 * in pfdHostAlloc memory (page-locked with a GPU), from which a frame's
 * upload is a direct DMA transfer. */
static int rec_grow(Rec* r, size_t need)
{
    size_t nc = r->captri ? r->captri : 1024;
    while (nc < need) nc *= 2;
    const size_t old = r->captri;
    int ok = host_grow((void**)&r->pos, old * 48, nc * 48)
          && host_grow((void**)&r->col, old * 12, nc * 12)
          && host_grow((void**)&r->uv, old * 24, nc * 24)
          && host_grow((void**)&r->nrm, old * 36, nc * 36)
          && host_grow((void**)&r->info, old * 4, nc * 4)
          && host_grow((void**)&r->vid, old * 12, nc * 12);
    if (!ok) return 0;
    r->captri = nc;
    return 1;
}

/* Appends one triangle with the given culling passes (PFF_PASS_*).  Only
 * the streams its state reads are written: texture coordinates for a
 * textured state, normals for a lit one (the kernels and the oracle read
 * them for no other triangle), colour-array indices once any corner of the
 * frame has one (the triangles before it then get PFF_NO_VERTEX). */
static void emit_tri(Ctx* c, const Vert* a, const Vert* b, const Vert* v3, uint32_t passes)
{
    if (!ensure_target(c)) { set_err(PF_INVALID_OPERATION); return; }
    int si = current_state(c);
    if (si < 0) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    Rec* r = &c->rec;
    if (r->ntri + 1 > r->captri && !rec_grow(r, r->ntri + 1)) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    const size_t t = r->ntri;
    const Vert* v[3] = { a, b, v3 };
    float* pos = r->pos + 12 * t;
    uint32_t* col = r->col + 3 * t;
    for (int k = 0; k < 3; k++) {
        memcpy(pos + 4 * k, v[k]->position, 16);
        col[k] = pack(v[k]->color);
    }
    const PFFstate* st = &r->states[si];
    if (st->texture >= 0) {
        float* uv = r->uv + 6 * t;
        for (int k = 0; k < 3; k++) memcpy(uv + 2 * k, v[k]->texcoord, 8);
    }
    if (st->flags & PFF_LIGHTING) {
        float* nrm = r->nrm + 9 * t;
        for (int k = 0; k < 3; k++) memcpy(nrm + 3 * k, v[k]->normal, 12);
    }
    if (!r->anyVid && (a->vid != PFF_NO_VERTEX || b->vid != PFF_NO_VERTEX || v3->vid != PFF_NO_VERTEX)) {
        memset(r->vid, 0xFF, t * 3 * sizeof(uint32_t));
        r->anyVid = 1;
    }
    if (r->anyVid)
        for (int k = 0; k < 3; k++) r->vid[3 * t + k] = v[k]->vid;
    r->info[t] = (uint32_t)si | passes;
    r->ntri++;
    if (r->ncmd && r->cmds[r->ncmd - 1].kind == PFF_CMD_DRAW && r->cmds[r->ncmd - 1].last == t) {
        r->cmds[r->ncmd - 1].last = (uint32_t)r->ntri;
    } else {
        PFFcmd cmd;
        memset(&cmd, 0, sizeof cmd);
        cmd.kind = PFF_CMD_DRAW;
        cmd.first = (uint32_t)t;
        cmd.last = (uint32_t)r->ntri;
        push_cmd(c, &cmd);
    }
}

/* pfiProcessRasterize_TRIANGLE_FAN / _STRIP (triangles.c:127-152) for one
 * face: records the triangles with that face's pass only. */
static void emit_fan(Ctx* c, uint32_t passes, int n)
{
    for (int i = 0; i < n; i++) emit_tri(c, &c->vb[0], &c->vb[i + 1], &c->vb[i + 2], passes);
}

static void emit_strip(Ctx* c, uint32_t passes, int n)
{
    for (int i = 0; i < n; i++) {
        if (i % 2 == 0) emit_tri(c, &c->vb[i], &c->vb[i + 1], &c->vb[i + 2], passes);
        else            emit_tri(c, &c->vb[i + 2], &c->vb[i + 1], &c->vb[i], passes);
    }
}

static uint32_t face_pass(int face) { return face == PF_FRONT ? PFF_PASS_FRONT : PFF_PASS_BACK; }

/* pfiProcessAndRasterize (internal/context/context.c:94-244), triangle
 * modes.  Point and line polygon modes are not part of this library. */
/* Appends one point (points.c:35-57); consecutive points share one
 * PFF_CMD_POINTS command. */
static void emit_point(Ctx* c, const Vert* v)
{
    if (!ensure_target(c)) { set_err(PF_INVALID_OPERATION); return; }
    int si = current_state(c);
    if (si < 0) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    Rec* r = &c->rec;
    if (!grow((void**)&r->pts, &r->cappts, r->npts + 1, sizeof(PFFpoint))) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    PFFpoint* p = &r->pts[r->npts];
    memcpy(p->position, v->position, 16);
    p->color = pack(v->color);
    p->state = (uint32_t)si;
    p->size = c->pointSize;
    p->pad = 0;
    const uint32_t k = (uint32_t)r->npts++;
    if (r->ncmd && r->cmds[r->ncmd - 1].kind == PFF_CMD_POINTS && r->cmds[r->ncmd - 1].last == k) {
        r->cmds[r->ncmd - 1].last = k + 1;
        return;
    }
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_POINTS;
    cmd.first = k;
    cmd.last = k + 1;
    push_cmd(c, &cmd);
}

/* Appends one line (lines.c:56-105); consecutive lines share one
 * PFF_CMD_LINES command. */
static void emit_line(Ctx* c, const Vert* a, const Vert* b, int groupStart)
{
    if (!ensure_target(c)) { set_err(PF_INVALID_OPERATION); return; }
    int si = current_state(c);
    if (si < 0) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    Rec* r = &c->rec;
    if (!grow((void**)&r->lns, &r->caplns, r->nlns + 1, sizeof(PFFline))) { set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    PFFline* l = &r->lns[r->nlns];
    memcpy(l->position[0], a->position, 16);
    memcpy(l->position[1], b->position, 16);
    l->color[0] = pack(a->color);
    l->color[1] = pack(b->color);
    l->state = (uint32_t)si;
    l->width = c->lineWidth;
    l->groupStart = groupStart ? 1u : 0u;
    l->pad = 0;
    const uint32_t k = (uint32_t)r->nlns++;
    if (r->ncmd && r->cmds[r->ncmd - 1].kind == PFF_CMD_LINES && r->cmds[r->ncmd - 1].last == k) {
        r->cmds[r->ncmd - 1].last = k + 1;
        return;
    }
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_LINES;
    cmd.first = k;
    cmd.last = k + 1;
    push_cmd(c, &cmd);
}

/* pfiProcessRasterize_POLY_LINES (lines.c:79-105): the n edges, one group. */
static void emit_poly_lines(Ctx* c, int n)
{
    for (int i = 0; i < n; i++) emit_line(c, &c->vb[i], &c->vb[(i + 1) % n], i == 0);
}

/* pfiProcessRasterize_POLY_POINTS (points.c:45-57): the n corners. */
static void emit_poly_points(Ctx* c, int n)
{
    for (int i = 0; i < n; i++) emit_point(c, &c->vb[i]);
}

static void process_primitive(Ctx* c)
{
    int culled = (c->state & PF_CULL_FACE) != 0;
    int face = culled ? !c->cullFace : PF_FRONT_AND_BACK;
    switch (c->drawMode) {
    case PF_TRIANGLES:                       /* ictx.c:104-138: each face by its polygon mode */
        if (face == PF_FRONT_AND_BACK && c->polyMode[0] == PF_FILL && c->polyMode[1] == PF_FILL) {
            emit_tri(c, &c->vb[0], &c->vb[1], &c->vb[2], PFF_PASS_FRONT | PFF_PASS_BACK);
            break;
        }
        for (int f = 0; f < 2; f++) {
            if (face != PF_FRONT_AND_BACK && face != f) continue;
            if (c->polyMode[f] == PF_FILL) emit_tri(c, &c->vb[0], &c->vb[1], &c->vb[2], face_pass(f));
            else if (c->polyMode[f] == PF_POINT) emit_poly_points(c, 3);
            else emit_poly_lines(c, 3);
        }
        break;
    case PF_TRIANGLE_FAN:
    case PF_QUAD_FAN:
    case PF_TRIANGLE_STRIP:
    case PF_QUAD_STRIP: {
        int n = (c->drawMode == PF_TRIANGLE_FAN || c->drawMode == PF_TRIANGLE_STRIP) ? 2 : 4;
        int fan = (c->drawMode == PF_TRIANGLE_FAN || c->drawMode == PF_QUAD_FAN);
        for (int f = 0; f < 2; f++) {
            if (face != PF_FRONT_AND_BACK && face != f) continue;
            if (fan) emit_fan(c, face_pass(f), n); else emit_strip(c, face_pass(f), n);
        }
        break;
    }
    case PF_QUADS:
        for (int f = 0; f < 2; f++) {
            if (face != PF_FRONT_AND_BACK && face != f) continue;
            if (c->polyMode[f] == PF_FILL) emit_fan(c, face_pass(f), 2);
            else if (c->polyMode[f] == PF_POINT) emit_poly_points(c, 4);
            else emit_poly_lines(c, 4);
        }
        break;
    case PF_POINTS:                          /* ictx.c:98-100 */
        emit_point(c, &c->vb[0]);
        break;
    case PF_LINES:                           /* ictx.c:101-103 */
        emit_line(c, &c->vb[0], &c->vb[1], 1);
        break;
    default:
        break;
    }
}

static uint32_t mode_vertex_count(PFdrawmode m)   /* context.c:61-74 */
{
    switch (m) {
    case PF_POINTS: return 1;
    case PF_LINES: return 2;
    case PF_TRIANGLES: return 3;
    case PF_TRIANGLE_FAN: case PF_TRIANGLE_STRIP: return 4;
    case PF_QUADS: return 4;
    case PF_QUAD_FAN: case PF_QUAD_STRIP: return 6;
    }
    return 0;
}

static void reset_vertex_buffer(Ctx* c)            /* context.c:41-59 */
{
    switch (c->drawMode) {
    case PF_TRIANGLE_FAN:
    case PF_TRIANGLE_STRIP:
        c->vcount = 1;
        c->vb[0] = c->vb[3];
        break;
    case PF_QUAD_FAN:
    case PF_QUAD_STRIP:
        c->vcount = 2;
        c->vb[0] = c->vb[4];
        c->vb[1] = c->vb[5];
        break;
    default:
        c->vcount = 0;
        break;
    }
}

/* pfiUpdateMatrices (context.c:96-111). */
static void update_matrices(Ctx* c, int withNormal)
{
    if (c->modelUsed) {
        m4_mul(c->matMVP, c->matModel, c->matView);
        m4_mul(c->matMVP, c->matMVP, c->matProj);
        if (withNormal && (c->state & PF_LIGHTING)) {
            m4_invert(c->matNormal, c->matModel);
            m4_transpose(c->matNormal, c->matNormal);
        }
    } else {
        m4_mul(c->matMVP, c->matView, c->matProj);
        if (withNormal && (c->state & PF_LIGHTING)) m4_identity(c->matNormal);
    }
    c->stateDirty = 1;
}

/* ------------------------------------------------------------------------ */
/* Textures (texture.c:30-123)                                               */
/* ------------------------------------------------------------------------ */

/* Texture formats: every (format, type) pair with a texel getter
 * (pixel.h:3050-3095).  The reference accepts any pair in range
 * (pfiIsPixelFormatValid, pixel.h:3149-3153) and calls a NULL getter when a
 * pair without one is sampled; here such a pair is PF_INVALID_ENUM. */
static int format_ok(PFpixelformat f, PFdatatype t)
{
    return pff_source_texel_bytes(PFF_SOURCE(f, t)) != 0u;
}

PFtexture pfGenTexture(void* pixels, PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type)
{
    if (!format_ok(format, type)) { set_err(PF_INVALID_ENUM); return NULL; }
    Tex* t = (Tex*)calloc(1, sizeof(Tex));
    if (!t) { set_err(PF_ERROR_OUT_OF_MEMORY); return NULL; }
    t->pixels = pixels;
    t->w = width;
    t->h = height;
    t->format = format;
    t->type = type;
    t->wrap = PF_REPEAT;                 /* texture.c:59-60: NEAREST / REPEAT */
    t->filter = PF_NEAREST;
    return t;
}

void pfDeleteTexture(PFtexture* texture, PFboolean freeBuffer)
{
    Tex* t = texture ? (Tex*)*texture : NULL;
    if (!t) return;
    if (G) {
        if (G->curTex == t) G->curTex = NULL;
        if (G->rec.ncmd) submit(G);      /* recorded draws may read it */
    }
    if (freeBuffer && t->pixels) {
        if (t->hostPixels) pfdHostFree(t->pixels);
        else free(t->pixels);
    }
    free(t);
    *texture = NULL;
}

PFboolean pfIsValidTexture(const PFtexture texture)
{
    const Tex* t = (const Tex*)texture;
    return t && t->pixels && t->w > 0 && t->h > 0;
}

void pfTextureParameter(PFtexture texture, PFtexturewrap wrapMode, PFtexturefilter filterMode)
{
    Tex* t = (Tex*)texture;
    if (!(wrapMode <= PF_CLAMP_TO_EDGE && filterMode <= PF_BILINEAR)) { set_err(PF_INVALID_ENUM); return; }
    if (!t) return;
    t->wrap = wrapMode;
    t->filter = filterMode;
    g_texGen++;
}

void* pfGetTexturePixels(const PFtexture texture, PFsizei* width, PFsizei* height,
                         PFpixelformat* format, PFdatatype* type)
{
    const Tex* t = (const Tex*)texture;
    if (!t) return NULL;
    if (width) *width = t->w;
    if (height) *height = t->h;
    if (format) *format = t->format;
    if (type) *type = t->type;
    return t->pixels;
}

/* ------------------------------------------------------------------------ */
/* Context (context.c:115-385)                                               */
/* ------------------------------------------------------------------------ */

PFcontext pfCreateContext(void* targetBuffer, PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type)
{
    if (pff_target_of(PFF_SOURCE(format, type)) == PFF_TARGET_NONE) { set_err(PF_INVALID_ENUM); return NULL; }
    Ctx* c = (Ctx*)calloc(1, sizeof(Ctx));
    if (!c) return NULL;
    size_t n = (size_t)width * height;
    c->mainFB.texture = pfGenTexture(targetBuffer, width, height, format, type);
    c->mainFB.zbuffer = (float*)pfdHostAlloc((n ? n : 1) * sizeof(float));
    if (!c->mainFB.texture || !c->mainFB.zbuffer) {
        free(c->mainFB.texture); pfdHostFree(c->mainFB.zbuffer); free(c);
        return NULL;
    }
    for (size_t i = 0; i < n; i++) c->mainFB.zbuffer[i] = FLT_MAX;    /* context.c:136-138 */
    c->curFB = &c->mainFB;

    c->vpPos[0] = c->vpMin[0] = 0;                                     /* context.c:146-149 */
    c->vpPos[1] = c->vpMin[1] = 0;
    c->vpDim[0] = width - 1;  c->vpMax[0] = (int32_t)(width - 1);
    c->vpDim[1] = height - 1; c->vpMax[1] = (int32_t)(height - 1);

    c->blendMode = PF_BLEND_ALPHA;                                     /* context.c:155-161 */
    c->depthFunc = PF_LESS;
    c->clearColor = (PFcolor){ 0, 0, 0, 255 };
    c->clearDepth = FLT_MAX;
    c->polyMode[0] = c->polyMode[1] = PF_FILL;
    c->curColor = (PFcolor){ 255, 255, 255, 255 };

    for (int i = 0; i < MAX_LIGHTS; i++) {                             /* context.c:184-200 */
        Light* L = &c->lights[i];
        memset(L, 0, sizeof *L);
        L->innerCutOff = (float)PF_PI_D;
        L->outerCutOff = (float)PF_PI_D;
        L->attConstant = 1.0f;
        L->ambient = (PFcolor){ 51, 51, 51, 255 };
        L->diffuse = (PFcolor){ 255, 255, 255, 255 };
        L->specular = (PFcolor){ 255, 255, 255, 255 };
        L->next = -1;
    }
    c->activeHead = -1;
    for (int f = 0; f < 2; f++) {                                      /* context.c:213-223 */
        c->mat[f].ambient = (PFcolor){ 255, 255, 255, 255 };
        c->mat[f].diffuse = (PFcolor){ 255, 255, 255, 255 };
        c->mat[f].specular = (PFcolor){ 255, 255, 255, 255 };
        c->mat[f].emission = (PFcolor){ 0, 0, 0, 255 };
        c->mat[f].shininess = 64.0f;
    }
    c->followFace = PF_FRONT_AND_BACK;
    c->followMode = PF_AMBIENT_AND_DIFFUSE;
    c->pointSize = c->lineWidth = 1.0f;                                 /* context.c:166-180 */
    c->pixelZoom[0] = c->pixelZoom[1] = 1.0f;
    c->fogMode = PF_LINEAR;                                   /* context.c:203-209 */
    c->fogDensity = 1.0f;
    c->fogStart = 0.0f;
    c->fogEnd = 1.0f;
    c->fogColor = (PFcolor){ 0, 0, 0, 0 };
    c->rasterPos[3] = 1.0f;

    c->matMode = PF_MODELVIEW;                                         /* context.c:232-239 */
    c->curMat = c->matView;
    m4_ortho(c->matProj, -1.0f, 1.0f, -1.0f, 1.0f, -1.0f, 1.0f);
    m4_identity(c->matTex);
    m4_identity(c->matNormal);
    m4_identity(c->matModel);
    m4_identity(c->matView);

    c->state |= PF_CULL_FACE;                                          /* context.c:243-247 */
    c->shadeModel = PF_SMOOTH;
    c->lightMode = PF_GOURAUD;
    c->cullFace = PF_BACK;
    c->err = PF_NO_ERROR;
    c->stateDirty = 1;
    return c;
}

void pfDeleteContext(PFcontext ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return;
    submit(c);
    pfdJobDestroy(c->diffJob);
    free(c->diffTex);
    pfdHostFree(c->mainFB.zbuffer);
    free(c->mainFB.texture);
    rec_free(&c->rec);
    if (G == c) G = NULL;
    free(c);
}

void pfSetMainBuffer(void* targetBuffer, PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type)
{
    Ctx* c = G;
    if (!c) return;
    if (targetBuffer == NULL || width == 0 || height == 0) { c->err = PF_INVALID_VALUE; return; }
    if (pff_target_of(PFF_SOURCE(format, type)) == PFF_TARGET_NONE) { c->err = PF_INVALID_ENUM; return; }
    submit(c);
    Tex* old = fb_tex(&c->mainFB);
    uint32_t oldW = old->w, oldH = old->h;
    if (oldW != width || oldH != height) {
        /* context.c:279-312, with the depth buffer sized in floats (the
         * reference reallocates it to width*height bytes). */
        float* z = (float*)pfdHostAlloc((size_t)width * height * sizeof(float));
        if (!z) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
        for (uint32_t y = 0; y < height; y++)
            for (uint32_t x = 0; x < width; x++) {
                size_t i = (size_t)y * width + x;
                size_t oi = (size_t)y * width + x;     /* same linear index as the realloc kept */
                z[i] = (oi < (size_t)oldW * oldH) ? c->mainFB.zbuffer[oi] : c->clearDepth;
                if (x >= oldW || y >= oldH) z[i] = c->clearDepth;
            }
        pfdHostFree(c->mainFB.zbuffer);
        c->mainFB.zbuffer = z;
    }
    old->pixels = targetBuffer;
    old->w = width;
    old->h = height;
    old->format = format;
    old->type = type;
    c->auxBuffer = NULL;
}

void pfSetAuxBuffer(void* auxFramebuffer) { if (G) G->auxBuffer = auxFramebuffer; }

void pfSwapBuffers(void)
{
    Ctx* c = G;
    if (!c) return;
    if (!c->auxBuffer) { c->err = PF_INVALID_OPERATION; return; }
    submit(c);
    Tex* t = fb_tex(c->curFB);
    void* tmp = t->pixels;
    t->pixels = c->auxBuffer;
    c->auxBuffer = tmp;
}

PFcontext pfGetCurrentContext(void) { return G; }
void pfMakeCurrent(PFcontext ctx) { G = (Ctx*)ctx; }

PFboolean pfIsEnabled(PFstate state) { return G && (G->state & state) != 0; }

void pfEnable(PFstate state)
{
    Ctx* c = G;
    if (!c) return;
    c->state |= state;
    if (state & PF_FRAMEBUFFER) c->curFB = c->boundFB ? c->boundFB : &c->mainFB;
    c->stateDirty = 1;
}

void pfDisable(PFstate state)
{
    Ctx* c = G;
    if (!c) return;
    c->state &= ~(uint32_t)state;
    if (state & PF_FRAMEBUFFER) c->curFB = &c->mainFB;
    c->stateDirty = 1;
}

PFerrcode pfGetError(void)
{
    if (!G) return PF_NO_ERROR;
    PFerrcode e = G->err;
    G->err = PF_NO_ERROR;
    return e;
}

/* ------------------------------------------------------------------------ */
/* Matrices (context.c:395-548)                                              */
/* ------------------------------------------------------------------------ */

void pfMatrixMode(PFmatrixmode mode)
{
    Ctx* c = G;
    if (!c) return;
    switch (mode) {
    case PF_PROJECTION: c->curMat = c->matProj; break;
    case PF_MODELVIEW:  c->curMat = c->modelUsed ? c->matModel : c->matView; break;
    case PF_TEXTURE:    c->curMat = c->matTex; break;
    default: c->err = PF_INVALID_ENUM; return;
    }
    c->matMode = mode;
}

void pfPushMatrix(void)
{
    Ctx* c = G;
    if (!c) return;
    switch (c->matMode) {
    case PF_PROJECTION:
        if (c->nProj >= MAX_PROJ_STACK) { c->err = PF_STACK_OVERFLOW; return; }
        memcpy(c->stackProj[c->nProj++], c->matProj, 64);
        break;
    case PF_MODELVIEW:
        if (c->nMV >= MAX_MV_STACK) { c->err = PF_STACK_OVERFLOW; return; }
        if (c->modelUsed) memcpy(c->stackMV[c->nMV++], c->matModel, 64);
        else { c->curMat = c->matModel; c->modelUsed = 1; }
        break;
    case PF_TEXTURE:
        if (c->nTex >= MAX_TEX_STACK) { c->err = PF_STACK_OVERFLOW; return; }
        memcpy(c->stackTex[c->nTex++], c->matTex, 64);
        break;
    }
}

void pfPopMatrix(void)
{
    Ctx* c = G;
    if (!c) return;
    switch (c->matMode) {
    case PF_PROJECTION:
        if (c->nProj == 0) { c->err = PF_STACK_UNDERFLOW; return; }
        memcpy(c->matProj, c->stackProj[--c->nProj], 64);
        break;
    case PF_MODELVIEW:
        if (c->nMV == 0) {
            if (!c->modelUsed) { c->err = PF_STACK_UNDERFLOW; return; }
            m4_identity(c->matModel);
            c->modelUsed = 0;
            c->curMat = c->matView;
        } else {
            memcpy(c->matModel, c->stackMV[--c->nMV], 64);
        }
        break;
    case PF_TEXTURE:
        if (c->nTex == 0) { c->err = PF_STACK_UNDERFLOW; return; }
        memcpy(c->matTex, c->stackTex[--c->nTex], 64);
        break;
    }
}

void pfLoadIdentity(void) { if (G) m4_identity(G->curMat); }

void pfTranslatef(PFfloat x, PFfloat y, PFfloat z)         /* pfmMat4Translate, pfm.h:2218 */
{
    if (!G) return;
    float t[16];
    m4_identity(t);
    t[12] = x; t[13] = y; t[14] = z;
    m4_mul(G->curMat, t, G->curMat);
}

void pfRotatef(PFfloat angle, PFfloat x, PFfloat y, PFfloat z)  /* pfmMat4Rotate, pfm.h:2227-2263 */
{
    if (!G) return;
    float a = (float)(angle * PF_DEG2RAD_D);
    float d[16];
    m4_identity(d);
    float lsq = x * x + y * y + z * z;
    if (lsq != 1.0f && lsq != 0.0f) {
        float il = 1.0f / sqrtf(lsq);
        x *= il; y *= il; z *= il;
    }
    float sn = sinf(a), cs = cosf(a), t = 1.0f - cs;
    d[0] = x * x * t + cs;  d[1] = y * x * t + z * sn;  d[2] = z * x * t - y * sn;
    d[4] = x * y * t - z * sn;  d[5] = y * y * t + cs;  d[6] = z * y * t + x * sn;
    d[8] = x * z * t + y * sn;  d[9] = y * z * t - x * sn;  d[10] = z * z * t + cs;
    m4_mul(G->curMat, d, G->curMat);
}

void pfScalef(PFfloat x, PFfloat y, PFfloat z)             /* pfmMat4Scale, pfm.h:2368 */
{
    if (!G) return;
    float s[16];
    memset(s, 0, sizeof s);
    s[0] = x; s[5] = y; s[10] = z; s[15] = 1.0f;
    m4_mul(G->curMat, s, G->curMat);
}

void pfMultMatrixf(const PFfloat* mat) { if (G) m4_mul(G->curMat, G->curMat, mat); }

void pfFrustum(PFfloat l, PFfloat r, PFfloat b, PFfloat t, PFfloat n, PFfloat f)
{
    if (!G) return;
    float m[16];
    m4_frustum(m, l, r, b, t, n, f);
    m4_mul(G->curMat, G->curMat, m);
}

void pfOrtho(PFfloat l, PFfloat r, PFfloat b, PFfloat t, PFfloat n, PFfloat f)
{
    if (!G) return;
    float m[16];
    m4_ortho(m, l, r, b, t, n, f);
    m4_mul(G->curMat, G->curMat, m);
}

/* ------------------------------------------------------------------------ */
/* Render state (context.c:553-797)                                          */
/* ------------------------------------------------------------------------ */

void pfViewport(PFint x, PFint y, PFsizei width, PFsizei height)
{
    Ctx* c = G;
    if (!c) return;
    if (x <= -(PFint)width || y <= -(PFint)height) { c->err = PF_INVALID_OPERATION; return; }
    Tex* m = fb_tex(&c->mainFB);
    c->vpPos[0] = x;
    c->vpPos[1] = y;
    c->vpDim[0] = width - 1;
    c->vpDim[1] = height - 1;
    c->vpMin[0] = x > 0 ? x : 0;
    c->vpMin[1] = y > 0 ? y : 0;
    int64_t mx = (int64_t)x + width, my = (int64_t)y + height;
    c->vpMax[0] = (int32_t)(mx < (int64_t)m->w - 1 ? mx : (int64_t)m->w - 1);
    c->vpMax[1] = (int32_t)(my < (int64_t)m->h - 1 ? my : (int64_t)m->h - 1);
    c->stateDirty = 1;
}

void pfPolygonMode(PFface face, PFpolygonmode mode)
{
    Ctx* c = G;
    if (!c) return;
    if (!(mode == PF_POINT || mode == PF_LINE || mode == PF_FILL)) { c->err = PF_INVALID_ENUM; return; }
    switch (face) {
    case PF_FRONT: c->polyMode[0] = mode; break;
    case PF_BACK: c->polyMode[1] = mode; break;
    case PF_FRONT_AND_BACK: c->polyMode[0] = c->polyMode[1] = mode; break;
    default: c->err = PF_INVALID_ENUM; break;
    }
}

void pfShadeModel(PFshademode mode) { if (G) { G->shadeModel = mode; G->stateDirty = 1; } }
void pfLightModel(PFlightmode mode) { if (G) { G->lightMode = mode; G->stateDirty = 1; } }

void pfCullFace(PFface face)
{
    if (!G) return;
    if (face > PF_BACK) { G->err = PF_INVALID_ENUM; return; }
    G->cullFace = face;
}

void pfBlendFunc(PFblendmode mode)
{
    if (!G) return;
    if (mode > PF_BLEND_DARKEN) { G->err = PF_INVALID_ENUM; return; }
    G->blendMode = mode;
    G->stateDirty = 1;
}

void pfDepthFunc(PFdepthmode mode)
{
    if (!G) return;
    if (mode > PF_GEQUAL) { G->err = PF_INVALID_ENUM; return; }
    G->depthFunc = mode;
    G->stateDirty = 1;
}

void pfBindFramebuffer(PFframebuffer* framebuffer)
{
    Ctx* c = G;
    if (!c) return;
    c->boundFB = framebuffer;
    if (c->state & PF_FRAMEBUFFER) c->curFB = framebuffer ? framebuffer : &c->mainFB;
}

void pfBindTexture(PFtexture texture)
{
    if (!G) return;
    G->curTex = (Tex*)texture;
    G->stateDirty = 1;
}

void pfClear(PFclearflag flag)
{
    Ctx* c = G;
    if (!c || !flag) return;
    if (!ensure_target(c)) { c->err = PF_INVALID_OPERATION; return; }
    /* context.c:696: either bit clears colour and depth. */
    if (!(flag & (PF_COLOR_BUFFER_BIT | PF_DEPTH_BUFFER_BIT))) return;
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_CLEAR;
    cmd.clearColor = pack(c->clearColor);
    cmd.clearDepth = c->clearDepth;
    push_cmd(c, &cmd);
}

void pfClearDepth(PFfloat depth) { if (G) G->clearDepth = depth; }
void pfClearColor(PFubyte r, PFubyte g, PFubyte b, PFubyte a) { if (G) G->clearColor = (PFcolor){ r, g, b, a }; }

/* ------------------------------------------------------------------------ */
/* Lights and materials (context.c:802-1156)                                 */
/* ------------------------------------------------------------------------ */

void pfEnableLight(PFsizei light)
{
    Ctx* c = G;
    if (!c) return;
    if (light >= MAX_LIGHTS) { c->err = PF_INVALID_VALUE; return; }
    int* link = &c->activeHead;
    while (*link >= 0) {
        if (*link == (int)light) { c->err = PF_INVALID_OPERATION; return; }
        link = &c->lights[*link].next;
    }
    c->lights[light].next = -1;
    *link = (int)light;
    c->stateDirty = 1;
}

void pfDisableLight(PFsizei light)
{
    Ctx* c = G;
    if (!c) return;
    if (light >= MAX_LIGHTS) { c->err = PF_INVALID_VALUE; return; }
    int* link = &c->activeHead;
    while (*link >= 0) {
        if (*link == (int)light) {
            *link = c->lights[light].next;
            c->lights[light].next = -1;
            c->stateDirty = 1;
            return;
        }
        link = &c->lights[*link].next;
    }
    c->err = PF_INVALID_OPERATION;
}

PFboolean pfIsEnabledLight(PFsizei light)
{
    Ctx* c = G;
    if (!c) return PF_FALSE;
    if (light >= MAX_LIGHTS) { c->err = PF_INVALID_VALUE; return PF_FALSE; }
    for (int i = c->activeHead; i >= 0; i = c->lights[i].next)
        if (i == (int)light) return PF_TRUE;
    return PF_FALSE;
}

static int cutoff_ok(float v) { return (v >= 0 && v <= 90) || v == 180; }

void pfLightf(PFsizei light, PFenum param, PFfloat value)
{
    Ctx* c = G;
    if (!c) return;
    if (light >= MAX_LIGHTS) { c->err = PF_STACK_OVERFLOW; return; }
    Light* L = &c->lights[light];
    switch (param) {
    case PF_SPOT_INNER_CUTOFF:
        if (cutoff_ok(value)) L->innerCutOff = cosf((float)(value * PF_DEG2RAD_D)); else c->err = PF_INVALID_VALUE;
        break;
    case PF_SPOT_OUTER_CUTOFF:
        if (cutoff_ok(value)) L->outerCutOff = cosf((float)(value * PF_DEG2RAD_D)); else c->err = PF_INVALID_VALUE;
        break;
    case PF_CONSTANT_ATTENUATION: L->attConstant = value; break;
    case PF_LINEAR_ATTENUATION: L->attLinear = value; break;
    case PF_QUADRATIC_ATTENUATION: L->attQuadratic = value; break;
    default: c->err = PF_INVALID_ENUM; return;
    }
    c->stateDirty = 1;
}

void pfLightfv(PFsizei light, PFenum param, const void* value)
{
    Ctx* c = G;
    if (!c) return;
    if (light >= MAX_LIGHTS) { c->err = PF_STACK_OVERFLOW; return; }
    Light* L = &c->lights[light];
    const float* v = (const float*)value;
    switch (param) {
    case PF_POSITION: memcpy(L->position, v, 12); break;
    case PF_SPOT_DIRECTION: memcpy(L->direction, v, 12); break;
    case PF_SPOT_INNER_CUTOFF:
    case PF_SPOT_OUTER_CUTOFF:
        pfLightf(light, param, v[0]);
        return;
    case PF_CONSTANT_ATTENUATION: L->attConstant = v[0]; break;
    case PF_LINEAR_ATTENUATION: L->attLinear = v[0]; break;
    case PF_QUADRATIC_ATTENUATION: L->attQuadratic = v[0]; break;
    case PF_AMBIENT: L->ambient = rgb_from_f(v); break;
    case PF_DIFFUSE: L->diffuse = rgb_from_f(v); break;
    case PF_SPECULAR: L->specular = rgb_from_f(v); break;
    default: c->err = PF_INVALID_ENUM; return;
    }
    c->stateDirty = 1;
}

static int material_targets(Ctx* c, PFface face, Material** m0, Material** m1)
{
    switch (face) {
    case PF_FRONT: *m0 = *m1 = &c->mat[0]; return 1;
    case PF_BACK: *m0 = *m1 = &c->mat[1]; return 1;
    case PF_FRONT_AND_BACK: *m0 = &c->mat[0]; *m1 = &c->mat[1]; return 1;
    default: c->err = PF_INVALID_ENUM; return 0;
    }
}

static void material_set(Ctx* c, PFface face, PFenum param, const float* v3, float scalar, int vec)
{
    Material *m0, *m1;
    if (!material_targets(c, face, &m0, &m1)) return;
    float s3[3] = { scalar, scalar, scalar };
    const float* v = vec ? v3 : s3;
    switch (param) {
    case PF_AMBIENT: m0->ambient = m1->ambient = rgb_from_f(v); break;
    case PF_DIFFUSE: m0->diffuse = m1->diffuse = rgb_from_f(v); break;
    case PF_SPECULAR: m0->specular = m1->specular = rgb_from_f(v); break;
    case PF_EMISSION: m0->emission = m1->emission = rgb_from_f(v); break;
    case PF_SHININESS: m0->shininess = m1->shininess = vec ? v3[0] : scalar; break;
    case PF_AMBIENT_AND_DIFFUSE:
        m0->ambient = m1->ambient = rgb_from_f(v);
        m0->diffuse = m1->diffuse = rgb_from_f(v);
        break;
    default: c->err = PF_INVALID_ENUM; return;
    }
    c->stateDirty = 1;
}

void pfMaterialf(PFface face, PFenum param, PFfloat value) { if (G) material_set(G, face, param, NULL, value, 0); }
void pfMaterialfv(PFface face, PFenum param, const void* value) { if (G) material_set(G, face, param, (const float*)value, 0.0f, 1); }

void pfColorMaterial(PFface face, PFenum mode)
{
    Ctx* c = G;
    if (!c) return;
    if (face > PF_FRONT_AND_BACK) { c->err = PF_INVALID_ENUM; return; }
    if (mode < PF_AMBIENT_AND_DIFFUSE || mode > PF_EMISSION) { c->err = PF_INVALID_ENUM; return; }
    c->followFace = face;
    c->followMode = mode;
}

/* ------------------------------------------------------------------------ */
/* Immediate mode (context.c:1580-1913)                                      */
/* ------------------------------------------------------------------------ */

static void list_begin(Ctx* c, PFdrawmode mode);
static void list_vertex(Ctx* c, const float* v);

void pfBegin(PFdrawmode mode)
{
    Ctx* c = G;
    if (!c) return;
    if (c->curList) { list_begin(c, mode); return; }      /* context.c:1589-1602 */
    if (mode > PF_QUAD_STRIP) { c->err = PF_INVALID_ENUM; return; }
    update_matrices(c, !(mode == PF_POINTS || mode == PF_LINES));
    c->drawMode = mode;
    c->vcount = 0;
}

void pfEnd(void) { if (G) G->vcount = 0; }

void pfVertex4fv(const PFfloat* v)
{
    Ctx* c = G;
    if (!c) return;
    if (c->curList) { list_vertex(c, v); return; }         /* context.c:1675-1683 */
    Vert* d = &c->vb[c->vcount++];
    memcpy(d->position, v, 16);
    memcpy(d->normal, c->curNormal, 12);
    memcpy(d->texcoord, c->curUV, 8);
    d->color = c->curColor;
    d->vid = PFF_NO_VERTEX;
    if (c->vcount == mode_vertex_count(c->drawMode)) {
        process_primitive(c);
        reset_vertex_buffer(c);
    }
}

void pfVertex2i(PFint x, PFint y) { float v[4] = { (float)x, (float)y, 0.0f, 1.0f }; pfVertex4fv(v); }
void pfVertex2f(PFfloat x, PFfloat y) { float v[4] = { x, y, 0.0f, 1.0f }; pfVertex4fv(v); }
void pfVertex2fv(const PFfloat* p) { float v[4] = { p[0], p[1], 0.0f, 1.0f }; pfVertex4fv(v); }
void pfVertex3i(PFint x, PFint y, PFint z) { float v[4] = { (float)x, (float)y, (float)z, 1.0f }; pfVertex4fv(v); }
void pfVertex3f(PFfloat x, PFfloat y, PFfloat z) { float v[4] = { x, y, z, 1.0f }; pfVertex4fv(v); }
void pfVertex3fv(const PFfloat* p) { float v[4] = { p[0], p[1], p[2], 1.0f }; pfVertex4fv(v); }
void pfVertex4i(PFint x, PFint y, PFint z, PFint w) { float v[4] = { (float)x, (float)y, (float)z, (float)w }; pfVertex4fv(v); }
void pfVertex4f(PFfloat x, PFfloat y, PFfloat z, PFfloat w) { float v[4] = { x, y, z, w }; pfVertex4fv(v); }

void pfColor(PFcolor color)                                  /* context.c:1687-1713 */
{
    Ctx* c = G;
    if (!c) return;
    if (c->state & PF_COLOR_MATERIAL) {
        Material* m1 = &c->mat[0];
        Material* m2 = &c->mat[1];
        if (c->followFace == PF_FRONT) m2 = m1;
        else if (c->followFace == PF_BACK) m1 = m2;
        switch (c->followMode) {
        case PF_AMBIENT_AND_DIFFUSE: m1->ambient = m2->ambient = color; m1->diffuse = m2->diffuse = color; break;
        case PF_AMBIENT: m1->ambient = m2->ambient = color; break;
        case PF_DIFFUSE: m1->diffuse = m2->diffuse = color; break;
        case PF_SPECULAR: m1->specular = m2->specular = color; break;
        case PF_EMISSION: m1->emission = m2->emission = color; break;
        }
        c->stateDirty = 1;
    } else {
        c->curColor = color;
    }
}

void pfColor1ui(PFuint color)
{
    PFcolor c;
    memcpy(&c, &color, 4);
    pfColor(c);
}
void pfColor3ub(PFubyte r, PFubyte g, PFubyte b) { pfColor((PFcolor){ r, g, b, 255 }); }
void pfColor3ubv(const PFubyte* v) { pfColor((PFcolor){ v[0], v[1], v[2], 255 }); }
void pfColor3us(PFushort r, PFushort g, PFushort b) { pfColor((PFcolor){ (PFubyte)(r >> 8), (PFubyte)(g >> 8), (PFubyte)(b >> 8), 255 }); }
void pfColor3usv(const PFushort* v) { pfColor3us(v[0], v[1], v[2]); }
void pfColor3ui(PFuint r, PFuint g, PFuint b) { pfColor((PFcolor){ (PFubyte)(r >> 24), (PFubyte)(g >> 24), (PFubyte)(b >> 24), 255 }); }
void pfColor3uiv(const PFuint* v) { pfColor3ui(v[0], v[1], v[2]); }
void pfColor3f(PFfloat r, PFfloat g, PFfloat b) { pfColor((PFcolor){ f2u8(r * 255), f2u8(g * 255), f2u8(b * 255), 255 }); }
void pfColor3fv(const PFfloat* v) { pfColor3f(v[0], v[1], v[2]); }
void pfColor4ub(PFubyte r, PFubyte g, PFubyte b, PFubyte a) { pfColor((PFcolor){ r, g, b, a }); }
void pfColor4ubv(const PFubyte* v) { pfColor((PFcolor){ v[0], v[1], v[2], v[3] }); }
void pfColor4us(PFushort r, PFushort g, PFushort b, PFushort a) { pfColor((PFcolor){ (PFubyte)(r >> 8), (PFubyte)(g >> 8), (PFubyte)(b >> 8), (PFubyte)(a >> 8) }); }
void pfColor4usv(const PFushort* v) { pfColor4us(v[0], v[1], v[2], v[3]); }
void pfColor4ui(PFuint r, PFuint g, PFuint b, PFuint a) { pfColor((PFcolor){ (PFubyte)(r >> 24), (PFubyte)(g >> 24), (PFubyte)(b >> 24), (PFubyte)(a >> 24) }); }
void pfColor4uiv(const PFuint* v) { pfColor4ui(v[0], v[1], v[2], v[3]); }
void pfColor4f(PFfloat r, PFfloat g, PFfloat b, PFfloat a) { pfColor((PFcolor){ f2u8(r * 255), f2u8(g * 255), f2u8(b * 255), f2u8(a * 255) }); }
void pfColor4fv(const PFfloat* v) { pfColor4f(v[0], v[1], v[2], v[3]); }

void pfTexCoord2f(PFfloat u, PFfloat v)                      /* pfmVec2Transform, pfm.h:954-962 */
{
    Ctx* c = G;
    if (!c) return;
    const float* m = c->matTex;
    float tu = m[0] * u + m[4] * v + m[12];
    float tv = m[1] * u + m[5] * v + m[13];
    c->curUV[0] = tu;
    c->curUV[1] = tv;
}

void pfTexCoordfv(const PFfloat* v) { pfTexCoord2f(v[0], v[1]); }

void pfNormal3f(PFfloat x, PFfloat y, PFfloat z)
{
    Ctx* c = G;
    if (!c) return;
    c->curNormal[0] = x; c->curNormal[1] = y; c->curNormal[2] = z;
    if (c->state & PF_NORMALIZE) {                           /* pfmVec3Normalize, pfm.h:1183 */
        float sq = x * x + y * y + z * z;
        if (sq == 0.0f) return;
        float inv = 1.0f / sqrtf(sq);
        for (int i = 0; i < 3; i++) c->curNormal[i] = c->curNormal[i] * inv;
    }
}

void pfNormal3fv(const PFfloat* v) { pfNormal3f(v[0], v[1], v[2]); }

/* ------------------------------------------------------------------------ */
/* Vertex arrays (context.c:1161-1575)                                       */
/* ------------------------------------------------------------------------ */

void pfVertexPointer(PFint size, PFenum type, PFsizei stride, const void* pointer)
{
    Ctx* c = G;
    if (!c) return;
    if (size < 2 || size > 4) { c->err = PF_INVALID_VALUE; return; }
    if (!(type == PF_SHORT || type == PF_INT || type == PF_FLOAT || type == PF_DOUBLE)) { c->err = PF_INVALID_ENUM; return; }
    c->aPos = (Attrib){ pointer, stride, size, type };
}

void pfNormalPointer(PFenum type, PFsizei stride, const void* pointer)
{
    Ctx* c = G;
    if (!c) return;
    if (!(type == PF_FLOAT || type == PF_DOUBLE)) { c->err = PF_INVALID_ENUM; return; }
    c->aNrm = (Attrib){ pointer, stride, 3, type };
}

void pfTexCoordPointer(PFenum type, PFsizei stride, const void* pointer)
{
    Ctx* c = G;
    if (!c) return;
    if (!(type == PF_FLOAT || type == PF_DOUBLE)) { c->err = PF_INVALID_ENUM; return; }
    c->aUV = (Attrib){ pointer, stride, 2, type };
}

void pfColorPointer(PFint size, PFenum type, PFsizei stride, const void* pointer)
{
    Ctx* c = G;
    if (!c) return;
    if (size < 3 || size > 4) { c->err = PF_INVALID_VALUE; return; }
    if (!(type == PF_UNSIGNED_BYTE || type == PF_UNSIGNED_SHORT || type == PF_UNSIGNED_INT ||
          type == PF_FLOAT || type == PF_DOUBLE)) { c->err = PF_INVALID_ENUM; return; }
    c->aCol = (Attrib){ pointer, stride, size, type };
}

static PFubyte d2u8(double d)          /* (PFubyte)(double) as gcc emits it on x86-64 */
{
    if (!(d > -2147483649.0 && d < 2147483648.0)) return 0;
    return (PFubyte)(uint32_t)(int32_t)d;
}

/* Component k of element j of a tightly packed array of `size`-wide
 * elements (the reference ignores the stride: context.c:1285, 1318). */
static float attr_f(const Attrib* a, size_t j, int k)
{
    size_t i = j * (size_t)a->size + (size_t)k;
    switch (a->type) {
    case PF_SHORT:  return (float)((const PFshort*)a->buffer)[i];
    case PF_INT:    return (float)((const PFint*)a->buffer)[i];
    case PF_FLOAT:  return ((const PFfloat*)a->buffer)[i];
    case PF_DOUBLE: return (float)((const PFdouble*)a->buffer)[i];
    }
    return 0.0f;
}

static PFubyte attr_u8(const Attrib* a, size_t j, int k)   /* context.c:1356-1392 */
{
    size_t i = j * (size_t)a->size + (size_t)k;
    switch (a->type) {
    case PF_UNSIGNED_BYTE:  return ((const PFubyte*)a->buffer)[i];
    case PF_UNSIGNED_SHORT: return (PFubyte)(((const PFushort*)a->buffer)[i] >> 8);
    case PF_UNSIGNED_INT:   return (PFubyte)(((const PFuint*)a->buffer)[i] >> 24);
    case PF_FLOAT:          return f2u8(((const PFfloat*)a->buffer)[i] * 255);
    case PF_DOUBLE:         return d2u8(((const PFdouble*)a->buffer)[i] * 255);
    }
    return 0;
}

/* The vertex loop shared by pfDrawElements and pfDrawArrays: element j of
 * the enabled arrays becomes the next vertex of the primitive being built
 * (context.c:1245-1404, 1427-1575).  Attributes without an enabled array
 * keep the values the slots were reset to: zero normal and texcoord and the
 * current colour (context.c:1254-1257); array texcoords skip the texture
 * matrix and array normals skip PF_NORMALIZE. */
static void draw_array_begin(Ctx* c, PFdrawmode mode)
{
    uint32_t n = mode_vertex_count(mode);
    for (uint32_t i = 0; i < n && i < 6; i++) {
        memset(&c->vb[i], 0, sizeof c->vb[i]);
        c->vb[i].color = c->curColor;
        c->vb[i].vid = PFF_NO_VERTEX;
    }
    pfBegin(mode);
}

static void draw_array_vertex(Ctx* c, size_t j)
{
    const int useUV = (c->state & PF_TEXTURE_COORD_ARRAY) && c->aUV.buffer;
    const int useN = (c->state & PF_NORMAL_ARRAY) && c->aNrm.buffer;
    const int useC = (c->state & PF_COLOR_ARRAY) && c->aCol.buffer;
    Vert* d = &c->vb[c->vcount++];
    d->position[0] = d->position[1] = d->position[2] = 0.0f;
    d->position[3] = 1.0f;
    for (int k = 0; k < c->aPos.size; k++) d->position[k] = attr_f(&c->aPos, j, k);
    if (useN) for (int k = 0; k < 3; k++) d->normal[k] = attr_f(&c->aNrm, j, k);
    if (useUV) for (int k = 0; k < 2; k++) d->texcoord[k] = attr_f(&c->aUV, j, k);
    if (useC) {
        PFubyte ch8[4] = { 255, 255, 255, 255 };
        for (int k = 0; k < c->aCol.size; k++) ch8[k] = attr_u8(&c->aCol, j, k);
        d->color = (PFcolor){ ch8[0], ch8[1], ch8[2], ch8[3] };
        d->vid = j <= 0xFFFFFFFEu ? (uint32_t)j : PFF_NO_VERTEX;
    }
    if (c->vcount == mode_vertex_count(c->drawMode)) {
        process_primitive(c);
        reset_vertex_buffer(c);
    }
}

static size_t type_bytes(PFenum t)
{
    switch (t) {
    case PF_UNSIGNED_BYTE: return 1;
    case PF_UNSIGNED_SHORT: case PF_SHORT: return 2;
    case PF_DOUBLE: return 8;
    default: return 4;
    }
}

/* Face groups of a draw in a triangle mode under the current culling and
 * polygon modes (ictx.c:94-244): the PFF_PASS_* bits of each group, in the
 * order the reference renders them.  Returns 0 when a point or line
 * polygon mode applies (such draws are recorded vertex by vertex). */
static int face_groups(const Ctx* c, PFdrawmode mode, uint32_t passes[2], uint32_t* groups)
{
    const int culled = (c->state & PF_CULL_FACE) != 0;
    const int face = culled ? !c->cullFace : PF_FRONT_AND_BACK;
    const int polyModes = (mode == PF_TRIANGLES || mode == PF_QUADS);   /* ictx.c:104-136 */
    *groups = 1;
    passes[1] = 0;
    if (face == PF_FRONT_AND_BACK) {
        if (polyModes && (c->polyMode[0] != PF_FILL || c->polyMode[1] != PF_FILL)) return 0;
        if (mode == PF_TRIANGLES) passes[0] = PFF_PASS_FRONT | PFF_PASS_BACK;
        else { *groups = 2; passes[0] = PFF_PASS_FRONT; passes[1] = PFF_PASS_BACK; }
    } else {
        if (polyModes && c->polyMode[face] != PF_FILL) return 0;
        passes[0] = face_pass(face);
    }
    return 1;
}

/* Triangles `count` vertices give per face group in a triangle mode: one
 * primitive per full vertex buffer (context.c:41-74), 1, 2 or 4 triangles
 * each. */
static uint64_t mode_triangles(PFdrawmode mode, uint64_t count)
{
    const uint32_t bufSize = mode_vertex_count(mode);
    const uint32_t keep = (mode == PF_TRIANGLE_FAN || mode == PF_TRIANGLE_STRIP) ? 1u
                        : (mode == PF_QUAD_FAN || mode == PF_QUAD_STRIP) ? 2u : 0u;
    const uint32_t perBuffer = (mode == PF_TRIANGLES) ? 1u : (mode == PF_QUAD_FAN || mode == PF_QUAD_STRIP) ? 4u : 2u;
    const uint64_t fills = count < bufSize ? 0 : 1 + (count - bufSize) / (bufSize - keep);
    return fills * perBuffer;
}

/* Appends an array draw of `tris` triangles (after pfBegin): target, state,
 * triangle slots and draw command.  Returns the record with its element
 * fields still to fill, or NULL after setting the error.  `block` (the
 * element copy, or NULL) is freed with the frame when `owned`. */
static PFFarrays* push_array_draw(Ctx* c, PFdrawmode mode, uint32_t count, uint64_t tris, uint32_t groups,
                                  const uint32_t passes[2], void* block, int owned)
{
    if (!ensure_target(c)) { set_err(PF_INVALID_OPERATION); if (owned) free(block); return NULL; }
    int si = current_state(c);
    Rec* r = &c->rec;
    if (si < 0 || r->ntri + tris > 0xFFFFFF00u
        || !grow((void**)&r->arrOwned, &r->capowned, r->narr + 1, 1)
        || !grow((void**)&r->arr, &r->caparr, r->narr + 1, sizeof(PFFarrays))
        || (r->ntri + tris > r->captri && !rec_grow(r, r->ntri + tris))) {
        if (owned) free(block);
        set_err(PF_ERROR_OUT_OF_MEMORY);
        return NULL;
    }
    PFFarrays* d = &r->arr[r->narr];
    memset(d, 0, sizeof *d);
    r->arrOwned[r->narr] = owned ? 1 : 0;
    d->position = block;
    d->mode = (uint32_t)mode;
    d->count = count;
    d->triFirst = (uint32_t)r->ntri;
    d->triCount = (uint32_t)tris;
    d->numFaces = groups;
    d->facePasses[0] = passes[0];
    d->facePasses[1] = passes[1];
    d->state = (uint32_t)si;
    d->color = pack(c->curColor);
    r->narr++;
    const size_t t = r->ntri;
    r->ntri += (size_t)tris;
    if (r->ncmd && r->cmds[r->ncmd - 1].kind == PFF_CMD_DRAW && r->cmds[r->ncmd - 1].last == t) {
        r->cmds[r->ncmd - 1].last = (uint32_t)r->ntri;
    } else {
        PFFcmd cmd;
        memset(&cmd, 0, sizeof cmd);
        cmd.kind = PFF_CMD_DRAW;
        cmd.first = (uint32_t)t;
        cmd.last = (uint32_t)r->ntri;
        push_cmd(c, &cmd);
    }
    return d;
}

/* A pfDrawElements / pfDrawArrays call whose primitives all become filled
 * triangles is recorded as one array draw (PFFarrays): the elements it
 * reads (0 .. the largest index, or first .. first + count - 1) are copied
 * into the frame now, since the caller may change the arrays afterwards,
 * and the device forms the triangles, in the order the per-vertex loop
 * below would record them.  Returns 0 when the call needs that loop: a
 * list is being recorded, a point or line polygon mode applies, or
 * pfDrawArrays starts before element 0. */
static int record_array_draw(Ctx* c, PFdrawmode mode, PFsizei count, PFdatatype itype, const void* indices,
                             int64_t first)
{
    if (c->curList || count <= 0 || mode < PF_TRIANGLES || mode > PF_QUAD_STRIP || first < 0) return 0;
    uint32_t groups, passes[2];
    if (!face_groups(c, mode, passes, &groups)) return 0;
    const uint64_t tris = mode_triangles(mode, (uint64_t)count) * groups;
    pfBegin(mode);                                   /* matrices, as the loop's pfBegin */
    if (tris == 0) { pfEnd(); return 1; }
    /* Elements read. */
    uint64_t lo = 0, n = 0;
    const size_t isz = indices ? type_bytes(itype) : 0;
    if (indices) {
        uint32_t mx = 0;
        for (PFsizei i = 0; i < count; i++) {
            uint32_t j = itype == PF_UNSIGNED_BYTE ? ((const PFubyte*)indices)[i]
                       : itype == PF_UNSIGNED_SHORT ? ((const PFushort*)indices)[i] : ((const PFuint*)indices)[i];
            if (j > mx) mx = j;
        }
        n = (uint64_t)mx + 1;
    } else {
        lo = (uint64_t)first;
        n = (uint64_t)count;
    }
    const int useUV = (c->state & PF_TEXTURE_COORD_ARRAY) && c->aUV.buffer;
    const int useN = (c->state & PF_NORMAL_ARRAY) && c->aNrm.buffer;
    const int useC = (c->state & PF_COLOR_ARRAY) && c->aCol.buffer;
    const size_t pb = (size_t)c->aPos.size * type_bytes(c->aPos.type);
    const size_t nb = useN ? 3 * type_bytes(c->aNrm.type) : 0;
    const size_t tb = useUV ? 2 * type_bytes(c->aUV.type) : 0;
    const size_t cb = useC ? (size_t)c->aCol.size * type_bytes(c->aCol.type) : 0;
#define ALIGN16(x) (((x) + 15) & ~(size_t)15)
    const size_t total = ALIGN16(n * pb) + ALIGN16(n * nb) + ALIGN16(n * tb) + ALIGN16(n * cb) + (size_t)count * isz + 16;
    uint8_t* blk = (uint8_t*)malloc(total);
    if (!blk) { set_err(PF_ERROR_OUT_OF_MEMORY); pfEnd(); return 1; }
    PFFarrays* d = push_array_draw(c, mode, (uint32_t)count, tris, groups, passes, blk, 1);
    if (!d) { pfEnd(); return 1; }
    uint8_t* q = blk;
    memcpy(q, (const uint8_t*)c->aPos.buffer + lo * pb, n * pb);
    d->posSize = c->aPos.size; d->posType = (uint32_t)c->aPos.type;
    q += ALIGN16(n * pb);
    if (useN) { memcpy(q, (const uint8_t*)c->aNrm.buffer + lo * nb, n * nb); d->normal = q; d->nrmType = (uint32_t)c->aNrm.type; }
    q += ALIGN16(n * nb);
    if (useUV) { memcpy(q, (const uint8_t*)c->aUV.buffer + lo * tb, n * tb); d->texcoord = q; d->uvType = (uint32_t)c->aUV.type; }
    q += ALIGN16(n * tb);
    if (useC) {
        memcpy(q, (const uint8_t*)c->aCol.buffer + lo * cb, n * cb);
        d->colors = q; d->colSize = c->aCol.size; d->colType = (uint32_t)c->aCol.type;
    }
    q += ALIGN16(n * cb);
    if (indices) { memcpy(q, indices, (size_t)count * isz); d->indices = q; d->idxType = (uint32_t)itype; }
#undef ALIGN16
    d->elementBase = (uint32_t)lo;
    d->numElements = (uint32_t)n;
    Rec* r = &c->rec;
    if (useC && !r->anyVid) {                        /* corners before this draw have no index */
        memset(r->vid, 0xFF, (size_t)d->triFirst * 3 * sizeof(uint32_t));
        r->anyVid = 1;
    }
    pfEnd();
    return 1;
}

static int draw_array_ready(Ctx* c, PFdrawmode mode)
{
    if (!(c->state & PF_VERTEX_ARRAY)) { c->err = PF_INVALID_OPERATION; return 0; }
    if (mode > PF_QUAD_STRIP) { c->err = PF_INVALID_ENUM; return 0; }
    /* The reference reads a missing position array through NULL. */
    if (!c->aPos.buffer) { c->err = PF_INVALID_OPERATION; return 0; }
    return 1;
}

void pfDrawElements(PFdrawmode mode, PFsizei count, PFdatatype type, const void* indices)
{
    Ctx* c = G;
    if (!c) return;
    if (!(type == PF_UNSIGNED_BYTE || type == PF_UNSIGNED_SHORT || type == PF_UNSIGNED_INT)) {
        c->err = PF_INVALID_ENUM;
        return;
    }
    if (!draw_array_ready(c, mode)) return;
    if (g_arrayDraws && record_array_draw(c, mode, count, type, indices, 0)) return;
    draw_array_begin(c, mode);
    for (PFsizei i = 0; i < count; i++) {
        size_t j = type == PF_UNSIGNED_BYTE ? ((const PFubyte*)indices)[i]
                 : type == PF_UNSIGNED_SHORT ? ((const PFushort*)indices)[i]
                 : ((const PFuint*)indices)[i];
        draw_array_vertex(c, j);
    }
    pfEnd();
}

void pfDrawArrays(PFdrawmode mode, PFint first, PFsizei count)
{
    Ctx* c = G;
    if (!c) return;
    if (!draw_array_ready(c, mode)) return;
    if (g_arrayDraws && record_array_draw(c, mode, count, PF_UNSIGNED_INT, NULL, first)) return;
    draw_array_begin(c, mode);
    for (PFsizei i = 0; i < count; i++) draw_array_vertex(c, (size_t)((int64_t)first + i));
    pfEnd();
}

/* ------------------------------------------------------------------------ */
/* Getters (getter.c)                                                        */
/* ------------------------------------------------------------------------ */

void pfGetBooleanv(PFenum pname, PFboolean* params)          /* getter.c:24-83 */
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_TEXTURE_2D: case PF_FRAMEBUFFER: case PF_BLEND: case PF_DEPTH_TEST:
    case PF_CULL_FACE: case PF_NORMALIZE: case PF_LIGHTING: case PF_COLOR_MATERIAL:
    case PF_VERTEX_ARRAY: case PF_NORMAL_ARRAY: case PF_COLOR_ARRAY: case PF_TEXTURE_COORD_ARRAY:
        *params = (c->state & pname) != 0;
        break;
    default:
        c->err = PF_INVALID_ENUM;
        break;
    }
}

void pfGetIntegerv(PFenum pname, PFint* params)              /* getter.c:85-190 */
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_VIEWPORT:
        params[0] = c->vpPos[0]; params[1] = c->vpPos[1];
        params[2] = (PFint)(c->vpDim[0] + 1); params[3] = (PFint)(c->vpDim[1] + 1);
        break;
    case PF_COLOR_CLEAR_VALUE:
        params[0] = c->clearColor.r; params[1] = c->clearColor.g;
        params[2] = c->clearColor.b; params[3] = c->clearColor.a;
        break;
    case PF_CULL_FACE_MODE: *params = (PFint)c->cullFace; break;
    case PF_CURRENT_COLOR:
        params[0] = c->curColor.r; params[1] = c->curColor.g;
        params[2] = c->curColor.b; params[3] = c->curColor.a;
        break;
    case PF_CURRENT_RASTER_POSITION:
        params[0] = (PFint)c->rasterPos[0]; params[1] = (PFint)c->rasterPos[1];
        break;
    case PF_POLYGON_MODE: params[0] = (PFint)c->polyMode[0]; params[1] = (PFint)c->polyMode[1]; break;
    case PF_MATRIX_MODE: *params = (PFint)c->matMode; break;
    case PF_MAX_PROJECTION_STACK_DEPTH: *params = MAX_PROJ_STACK; break;
    case PF_MAX_MODELVIEW_STACK_DEPTH: *params = MAX_MV_STACK; break;
    case PF_MAX_TEXTURE_STACK_DEPTH: *params = MAX_TEX_STACK; break;
    case PF_SHADE_MODEL: *params = (PFint)c->shadeModel; break;
    case PF_MAX_LIGHTS: *params = MAX_LIGHTS; break;
    case PF_VERTEX_ARRAY_SIZE: *params = c->aPos.size; break;
    case PF_VERTEX_ARRAY_STRIDE: *params = (PFint)c->aPos.stride; break;
    case PF_VERTEX_ARRAY_TYPE: *params = (PFint)c->aPos.type; break;
    case PF_NORMAL_ARRAY_STRIDE: *params = (PFint)c->aNrm.stride; break;
    case PF_NORMAL_ARRAY_TYPE: *params = (PFint)c->aNrm.type; break;
    case PF_TEXTURE_COORD_ARRAY_STRIDE: *params = (PFint)c->aUV.stride; break;
    case PF_TEXTURE_COORD_ARRAY_TYPE: *params = (PFint)c->aUV.type; break;
    case PF_COLOR_ARRAY_SIZE: *params = c->aCol.size; break;
    case PF_COLOR_ARRAY_STRIDE: *params = (PFint)c->aCol.stride; break;
    case PF_COLOR_ARRAY_TYPE: *params = (PFint)c->aCol.type; break;
    default:
        c->err = PF_INVALID_ENUM;
        break;
    }
}

#define PF_INV_255_D (1.0 / 255.0)                           /* PF_INV_255, context.h */

static void modelview(const Ctx* c, float* out) { m4_mul(out, c->matModel, c->matView); }

void pfGetFloatv(PFenum pname, PFfloat* params)              /* getter.c:192-352 */
{
    Ctx* c = G;
    if (!c) return;
    const float inv = (float)PF_INV_255_D;
    switch (pname) {
    case PF_COLOR_CLEAR_VALUE:
        params[0] = c->clearColor.r * inv; params[1] = c->clearColor.g * inv;
        params[2] = c->clearColor.b * inv; params[3] = c->clearColor.a * inv;
        break;
    case PF_DEPTH_CLEAR_VALUE: *params = c->clearDepth; break;
    case PF_CURRENT_COLOR:
        params[0] = c->curColor.r * inv; params[1] = c->curColor.g * inv;
        params[2] = c->curColor.b * inv; params[3] = c->curColor.a * inv;
        break;
    case PF_CURRENT_NORMAL: memcpy(params, c->curNormal, 12); break;
    case PF_CURRENT_TEXTURE_COORDS: memcpy(params, c->curUV, 8); break;
    case PF_CURRENT_RASTER_POSITION: params[0] = c->rasterPos[0]; params[1] = c->rasterPos[1]; break;
    case PF_POINT_SIZE: *params = c->pointSize; break;
    case PF_LINE_WIDTH: *params = c->lineWidth; break;
    case PF_PROJECTION_MATRIX: memcpy(params, c->matProj, 64); break;
    case PF_MODELVIEW_MATRIX: modelview(c, params); break;
    case PF_TEXTURE_MATRIX: memcpy(params, c->matTex, 64); break;
    case PF_ZOOM_X: *params = c->pixelZoom[0]; break;
    case PF_ZOOM_Y: *params = c->pixelZoom[1]; break;
    default:
        c->err = PF_INVALID_ENUM;
        break;
    }
}

void pfGetDoublev(PFenum pname, PFdouble* params)            /* getter.c:354-476 */
{
    Ctx* c = G;
    if (!c) return;
    float m[16];
    switch (pname) {
    case PF_COLOR_CLEAR_VALUE:
        params[0] = c->clearColor.r * PF_INV_255_D; params[1] = c->clearColor.g * PF_INV_255_D;
        params[2] = c->clearColor.b * PF_INV_255_D; params[3] = c->clearColor.a * PF_INV_255_D;
        break;
    case PF_DEPTH_CLEAR_VALUE: *params = c->clearDepth; break;
    case PF_CURRENT_COLOR:
        params[0] = c->curColor.r * PF_INV_255_D; params[1] = c->curColor.g * PF_INV_255_D;
        params[2] = c->curColor.b * PF_INV_255_D; params[3] = c->curColor.a * PF_INV_255_D;
        break;
    case PF_CURRENT_NORMAL: for (int i = 0; i < 3; i++) params[i] = c->curNormal[i]; break;
    case PF_CURRENT_TEXTURE_COORDS: params[0] = c->curUV[0]; params[1] = c->curUV[1]; break;
    case PF_CURRENT_RASTER_POSITION: params[0] = c->rasterPos[0]; params[1] = c->rasterPos[1]; break;
    case PF_POINT_SIZE: *params = c->pointSize; break;
    case PF_LINE_WIDTH: *params = c->lineWidth; break;
    case PF_PROJECTION_MATRIX: for (int i = 0; i < 16; i++) params[i] = c->matProj[i]; break;
    case PF_MODELVIEW_MATRIX: modelview(c, m); for (int i = 0; i < 16; i++) params[i] = m[i]; break;
    case PF_TEXTURE_MATRIX: for (int i = 0; i < 16; i++) params[i] = c->matTex[i]; break;
    case PF_ZOOM_X: *params = c->pixelZoom[0]; break;
    case PF_ZOOM_Y: *params = c->pixelZoom[1]; break;
    default:
        c->err = PF_INVALID_ENUM;
        break;
    }
}

void pfGetPointerv(PFenum pname, const void** params)        /* getter.c:478-521 */
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_TEXTURE_2D: *params = c->curTex; break;
    case PF_FRAMEBUFFER: *params = c->boundFB; break;
    case PF_BLEND_FUNC: *params = &c->blendMode; break;
    case PF_DEPTH_FUNC: *params = &c->depthFunc; break;
    default:
        c->err = PF_INVALID_ENUM;
        break;
    }
}

/* ------------------------------------------------------------------------ */
/* Outside the GPU triangle path                                             */
/* ------------------------------------------------------------------------ */

void pfLineWidth(PFfloat width)                              /* context.c:611-618 */
{
    if (!G) return;
    if (width <= 0.0f) { G->err = PF_INVALID_VALUE; return; }
    G->lineWidth = width;
}

void pfPointSize(PFfloat size)                               /* context.c:620-627 */
{
    if (!G) return;
    if (size <= 0.0f) { G->err = PF_INVALID_VALUE; return; }
    G->pointSize = size;
}

void pfPixelZoom(PFfloat xfactor, PFfloat yfactor)           /* context.c:2086-2090 */
{
    if (!G) return;
    G->pixelZoom[0] = xfactor;
    G->pixelZoom[1] = yfactor;
}

static void raster_pos(float x, float y, float z, float w)   /* context.c:2091-2155 */
{
    if (!G) return;
    G->rasterPos[0] = x; G->rasterPos[1] = y; G->rasterPos[2] = z; G->rasterPos[3] = w;
}
void pfRasterPos2i(PFint x, PFint y) { raster_pos((float)x, (float)y, 0.0f, 1.0f); }
void pfRasterPos2f(PFfloat x, PFfloat y) { raster_pos(x, y, 0.0f, 1.0f); }
void pfRasterPos2fv(const PFfloat* v) { raster_pos(v[0], v[1], 0.0f, 1.0f); }
void pfRasterPos3i(PFint x, PFint y, PFint z) { raster_pos((float)x, (float)y, (float)z, 1.0f); }
void pfRasterPos3f(PFfloat x, PFfloat y, PFfloat z) { raster_pos(x, y, z, 1.0f); }
void pfRasterPos3fv(const PFfloat* v) { raster_pos(v[0], v[1], v[2], 1.0f); }
void pfRasterPos4i(PFint x, PFint y, PFint z, PFint w) { raster_pos((float)x, (float)y, (float)z, (float)w); }
void pfRasterPos4f(PFfloat x, PFfloat y, PFfloat z, PFfloat w) { raster_pos(x, y, z, w); }
void pfRasterPos4fv(const PFfloat* v) { raster_pos(v[0], v[1], v[2], v[3]); }



/* (PFint)(float) as gcc emits it on x86-64 (cvttss2si): NaN and values out
 * of range give INT32_MIN. */
static int32_t f2i_x86(float f)
{
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}

/* pfDrawPixels (context.c:1987-2080).  The call computes, from the current
 * context, the raster position on screen (the MVP-transformed raster
 * position with no perspective division), the destination rectangle
 * clipped to the viewport, the source steps and the depth / blend state,
 * and records them with a copy of the image as a PFF_CMD_PIXELS command:
 * the pixels are written on the GPU in order with the frame's triangles,
 * through the reference's scalar getter, depth test and blend function. */
void pfDrawPixels(PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type, const void* pixels)
{
    Ctx* c = G;
    if (!c) return;
    if (width == 0 || height == 0) { c->err = PF_INVALID_VALUE; return; }
    const uint32_t source = PFF_SOURCE(format, type);
    const uint32_t bpp = pff_source_texel_bytes(source);
    if (bpp == 0u) { c->err = PF_INVALID_ENUM; return; }              /* context.c:1996-2006 */
    if (!pixels) { c->err = PF_INVALID_VALUE; return; }
    if (!ensure_target(c)) { c->err = PF_INVALID_OPERATION; return; }
    update_matrices(c, 0);
    const float* m = c->matMVP;                                      /* pfmVec4Transform, pfm.h:1996-2006 */
    const float* v = c->rasterPos;
    float rp[3];
    for (int i = 0; i < 3; i++) rp[i] = m[i] * v[0] + m[4 + i] * v[1] + m[8 + i] * v[2] + m[12 + i] * v[3];
    PFFpixels op;
    memset(&op, 0, sizeof op);
    op.width = width;
    op.height = height;
    op.source = source;
    op.xScreen = f2i_x86((float)c->vpPos[0] + (rp[0] + 1.0f) * 0.5f * (float)c->vpDim[0]);
    op.yScreen = f2i_x86((float)c->vpPos[1] + (1.0f - rp[1]) * 0.5f * (float)c->vpDim[1]);
    op.zPos = rp[2];
    /* PF_CLAMP of ints (context.c:2024-2025) and of floats (2026-2027). */
    op.xMin = op.xScreen < c->vpMin[0] ? c->vpMin[0] : (op.xScreen > c->vpMax[0] ? c->vpMax[0] : op.xScreen);
    op.yMin = op.yScreen < c->vpMin[1] ? c->vpMin[1] : (op.yScreen > c->vpMax[1] ? c->vpMax[1] : op.yScreen);
    const float xe = (float)op.xScreen + (float)width * c->pixelZoom[0];
    const float ye = (float)op.yScreen + (float)height * c->pixelZoom[1];
    op.xMax = f2i_x86(xe < (float)c->vpMin[0] ? (float)c->vpMin[0] : (xe > (float)c->vpMax[0] ? (float)c->vpMax[0] : xe));
    op.yMax = f2i_x86(ye < (float)c->vpMin[1] ? (float)c->vpMin[1] : (ye > (float)c->vpMax[1] ? (float)c->vpMax[1] : ye));
    op.invXLen = 1.0f / ((float)width * c->pixelZoom[0]);
    op.invYLen = 1.0f / ((float)height * c->pixelZoom[1]);
    op.flags = ((c->state & PF_DEPTH_TEST) ? PFF_DEPTH_TEST : 0u) | ((c->state & PF_BLEND) ? PFF_BLEND : 0u);
    op.depthFunc = (uint32_t)c->depthFunc;
    op.blendMode = (uint32_t)c->blendMode;
    const size_t bytes = (size_t)width * height * bpp;
    void* copy = malloc(bytes + 4);
    if (!copy) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
    memcpy(copy, pixels, bytes);
    memset((uint8_t*)copy + bytes, 0, 4);
    op.pixels = copy;
    Rec* r = &c->rec;
    if (!grow((void**)&r->pix, &r->cappix, r->npix + 1, sizeof(PFFpixels))) {
        free(copy);
        c->err = PF_ERROR_OUT_OF_MEMORY;
        return;
    }
    r->pix[r->npix] = op;
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_PIXELS;
    cmd.first = (uint32_t)r->npix;
    r->npix++;
    push_cmd(c, &cmd);
}
/* pfRectf (context.c:1938-1978): the corners through the MVP (no division),
 * to screen, sorted and clamped to the viewport; recorded as a pixel
 * command whose 1x1 image is the current colour, with no depth test, depth
 * write or blending, so it fills in order with the draws. */
void pfRectf(PFfloat x1, PFfloat y1, PFfloat x2, PFfloat y2)
{
    Ctx* c = G;
    if (!c) return;
    if (!ensure_target(c)) { c->err = PF_INVALID_OPERATION; return; }
    update_matrices(c, 0);
    const float* m = c->matMVP;
    const float in[2][2] = { { x1, y1 }, { x2, y2 } };
    int32_t ix[2], iy[2];
    for (int k = 0; k < 2; k++) {
        const float v[4] = { in[k][0], in[k][1], 0.0f, 1.0f };
        float t[2];
        for (int i = 0; i < 2; i++) t[i] = m[i] * v[0] + m[4 + i] * v[1] + m[8 + i] * v[2] + m[12 + i] * v[3];
        ix[k] = f2i_x86((float)c->vpPos[0] + (t[0] + 1.0f) * 0.5f * (float)c->vpDim[0]);
        iy[k] = f2i_x86((float)c->vpPos[1] + (1.0f - t[1]) * 0.5f * (float)c->vpDim[1]);
    }
    if (ix[1] < ix[0]) { int32_t t = ix[0]; ix[0] = ix[1]; ix[1] = t; }
    if (iy[1] < iy[0]) { int32_t t = iy[0]; iy[0] = iy[1]; iy[1] = t; }
    for (int k = 0; k < 2; k++) {
        ix[k] = ix[k] < c->vpMin[0] ? c->vpMin[0] : (ix[k] > c->vpMax[0] ? c->vpMax[0] : ix[k]);
        iy[k] = iy[k] < c->vpMin[1] ? c->vpMin[1] : (iy[k] > c->vpMax[1] ? c->vpMax[1] : iy[k]);
    }
    uint32_t* px = (uint32_t*)malloc(8);
    if (!px) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
    px[0] = pack(c->curColor);
    px[1] = 0u;
    PFFpixels op;
    memset(&op, 0, sizeof op);
    op.pixels = px;
    op.width = op.height = 1;
    op.source = PFF_SOURCE(PF_RGBA, PF_UNSIGNED_BYTE);
    op.xScreen = ix[0]; op.yScreen = iy[0];
    op.xMin = ix[0]; op.yMin = iy[0]; op.xMax = ix[1]; op.yMax = iy[1];
    op.flags = PFF_PIXELS_RECT;
    Rec* r = &c->rec;
    if (!grow((void**)&r->pix, &r->cappix, r->npix + 1, sizeof(PFFpixels))) {
        free(px);
        c->err = PF_ERROR_OUT_OF_MEMORY;
        return;
    }
    r->pix[r->npix] = op;
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_PIXELS;
    cmd.first = (uint32_t)r->npix;
    r->npix++;
    push_cmd(c, &cmd);
}

void pfRects(PFshort x1, PFshort y1, PFshort x2, PFshort y2) { pfRectf((float)x1, (float)y1, (float)x2, (float)y2); }
void pfRectsv(const PFshort* v1, const PFshort* v2) { pfRectf((float)v1[0], (float)v1[1], (float)v2[0], (float)v2[1]); }
void pfRectfv(const PFfloat* v1, const PFfloat* v2) { pfRectf(v1[0], v1[1], v2[0], v2[1]); }

/* Fog state setters (context.c:2161-2273), quirks kept: every DENSITY
 * branch stores into the mode (so the density stays 1), pfFogiv and pfFogfv
 * accept any MODE / DENSITY value (their range tests use ||). */
void pfFogi(PFfogparam pname, PFint param)
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_FOG_MODE:
        if (param >= PF_LINEAR && param <= PF_EXP2) c->fogMode = param; else c->err = PF_INVALID_VALUE;
        break;
    case PF_FOG_DENSITY:
        if (param == 0 || param == 1) c->fogMode = param; else c->err = PF_INVALID_VALUE;
        break;
    case PF_FOG_START: c->fogStart = (float)param; break;
    case PF_FOG_END: c->fogEnd = (float)param; break;
    default: c->err = PF_INVALID_ENUM; break;
    }
}

void pfFogf(PFfogparam pname, PFfloat param)
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_FOG_DENSITY:
        if (param >= 0 && param <= 1) c->fogMode = f2i_x86(param); else c->err = PF_INVALID_VALUE;
        break;
    case PF_FOG_START: c->fogStart = param; break;
    case PF_FOG_END: c->fogEnd = param; break;
    default: c->err = PF_INVALID_ENUM; break;
    }
}

void pfFogiv(PFfogparam pname, PFint* param)
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_FOG_MODE: case PF_FOG_DENSITY: c->fogMode = *param; break;
    case PF_FOG_START: c->fogStart = (float)*param; break;
    case PF_FOG_END: c->fogEnd = (float)*param; break;
    case PF_FOG_COLOR:
        c->fogColor = (PFcolor){ (PFubyte)param[0], (PFubyte)param[1], (PFubyte)param[2], (PFubyte)param[3] };
        break;
    default: c->err = PF_INVALID_ENUM; break;
    }
}

void pfFogfv(PFfogparam pname, PFfloat* param)
{
    Ctx* c = G;
    if (!c) return;
    switch (pname) {
    case PF_FOG_DENSITY: c->fogMode = f2i_x86(*param); break;
    case PF_FOG_START: c->fogStart = *param; break;
    case PF_FOG_END: c->fogEnd = *param; break;
    case PF_FOG_COLOR:
        c->fogColor = (PFcolor){ f2u8(255 * param[0]), f2u8(255 * param[1]), f2u8(255 * param[2]), f2u8(255 * param[3]) };
        break;
    default: c->err = PF_INVALID_ENUM; break;
    }
}

/* pfFogProcess (context.c:2275-2343): recorded with the fog state as a
 * PFF_CMD_FOG command; the GPU blends the fog colour into every pixel from
 * its depth after the draws recorded before the call. */
void pfFogProcess(void)
{
    Ctx* c = G;
    if (!c) return;
    if (!ensure_target(c)) { c->err = PF_INVALID_OPERATION; return; }
    PFFpixels op;
    memset(&op, 0, sizeof op);
    op.fogMode = (uint32_t)c->fogMode;
    op.fogDensity = c->fogDensity;
    op.fogStart = c->fogStart;
    op.fogEnd = c->fogEnd;
    op.fogColor = pack(c->fogColor);
    Rec* r = &c->rec;
    if (!grow((void**)&r->pix, &r->cappix, r->npix + 1, sizeof(PFFpixels))) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
    r->pix[r->npix] = op;
    PFFcmd cmd;
    memset(&cmd, 0, sizeof cmd);
    cmd.kind = PFF_CMD_FOG;
    cmd.first = (uint32_t)r->npix;
    r->npix++;
    push_cmd(c, &cmd);
}

static void set_pixel(void* dst, size_t i, PFpixelformat f, PFdatatype t, PFcolor c);
static PFcolor get_pixel(const void* src, size_t i, PFpixelformat f, PFdatatype t);

/* pfPostProcess (context.c:2397-2440): the caller's function is host code,
 * so the recorded frame is executed first and the function then runs over
 * the render target in host memory, pixel by pixel with its depth. */
void pfPostProcess(PFpostprocessfunc fn)
{
    Ctx* c = G;
    if (!c) return;
    if (!fn) { c->err = PF_INVALID_VALUE; return; }
    submit(c);
    Tex* t = fb_tex(c->curFB);
    if (!t || !t->pixels || !c->curFB->zbuffer) { c->err = PF_INVALID_OPERATION; return; }
    const float* z = c->curFB->zbuffer;
    for (int32_t y = 0; y < (int32_t)t->h; y++)
        for (int32_t x = 0; x < (int32_t)t->w; x++) {
            const size_t i = (size_t)y * t->w + (size_t)x;
            set_pixel(t->pixels, i, t->format, t->type, fn(x, y, z[i], get_pixel(t->pixels, i, t->format, t->type)));
        }
}

/* ------------------------------------------------------------------------ */
/* Render lists (renderlist.c:6-97)                                          */
/* ------------------------------------------------------------------------ */

/* A list holds the pfBegin calls made while it was being recorded, each
 * with the materials and texture current at that pfBegin and, per vertex,
 * the position given to pfVertex and the colour, texture coordinate (after
 * the texture matrix) and normal current at that moment.  pfCallList feeds
 * them back through pfBegin / pfColor4ubv / pfTexCoordfv / pfNormal3fv /
 * pfVertex4fv, so replayed triangles are recorded into the frame like any
 * other and the matrices, lights and enables of the moment of the call
 * apply.  Recording and replay each save the colour, texture coordinate,
 * normal, materials, bound texture and enable bits in the context's one
 * backup slot and restore them at pfEndList / the end of pfCallList
 * (internal/context/context.c:26-47); a pfCallList made while recording
 * overwrites the slot that pfEndList restores from, as in the reference. */

static void make_backup(Ctx* c)
{
    memcpy(c->backup.faceMaterial, c->mat, sizeof c->mat);
    memcpy(c->backup.uv, c->curUV, sizeof c->curUV);
    memcpy(c->backup.nrm, c->curNormal, sizeof c->curNormal);
    c->backup.color = c->curColor;
    c->backup.texture = c->curTex;
    c->backup.state = c->state;
}

static void restore_backup(Ctx* c)
{
    memcpy(c->mat, c->backup.faceMaterial, sizeof c->mat);
    memcpy(c->curUV, c->backup.uv, sizeof c->curUV);
    memcpy(c->curNormal, c->backup.nrm, sizeof c->curNormal);
    c->curColor = c->backup.color;
    c->curTex = c->backup.texture;
    c->state = c->backup.state;
    c->stateDirty = 1;
}

static void free_call(ListCall* k)
{
    if (k->packed) free(k->packed);
    else { free(k->pos); free(k->uv); free(k->nrm); free(k->col); }
    if (k->resident) pfdResidentDestroy(k->resident);
    memset(k, 0, sizeof *k);
}

static void store_release(ListStore* st)
{
    if (!st || __atomic_sub_fetch(&st->refs, 1, __ATOMIC_ACQ_REL) != 0) return;
    for (size_t i = 0; i < st->n; i++) free_call(&st->calls[i]);
    free(st->calls);
    free(st);
}

/* pfEndList: the recorded calls become an immutable store.  Each call's
 * vertices are packed into one block (positions, normals, texture
 * coordinates, colours) and copied to the GPU once, so that every later
 * pfCallList references them instead of replaying them (a call keeps no
 * device copy when no GPU is usable). */
static void seal_list(List* L)
{
    ListStore* st = (ListStore*)calloc(1, sizeof *st);
    if (!st) return;                                     /* stays replayable on the host */
    for (size_t i = 0; i < L->n; i++) {
        ListCall* k = &L->calls[i];
        const size_t n = k->n;
#define ALIGN16(x) (((x) + 15) & ~(size_t)15)
        const size_t onrm = ALIGN16(n * 16), ouv = onrm + ALIGN16(n * 12), ocol = ouv + ALIGN16(n * 8);
        const size_t bytes = ocol + n * 4;
#undef ALIGN16
        uint8_t* blk = (uint8_t*)malloc(bytes + 16);
        if (!blk) continue;                              /* this call stays in its own arrays */
        if (n) {
            memcpy(blk, k->pos, n * 16);
            memcpy(blk + onrm, k->nrm, n * 12);
            memcpy(blk + ouv, k->uv, n * 8);
            memcpy(blk + ocol, k->col, n * 4);
        }
        free(k->pos); free(k->uv); free(k->nrm); free(k->col);
        k->packed = blk;
        k->pos = (float*)blk;
        k->nrm = (float*)(blk + onrm);
        k->uv = (float*)(blk + ouv);
        k->col = (PFcolor*)(blk + ocol);
        k->cap = n;
        if (n && g_arrayDraws && pfdResidentCreate(blk, bytes, &k->resident) != 0) k->resident = NULL;
    }
    st->calls = L->calls;
    st->n = L->n;
    st->refs = 1;
    L->store = st;
}

/* Forgets a list's calls: its store is released (pending frames keep their
 * references), unsealed calls are freed. */
static void list_clear(List* L)
{
    if (L->store) store_release(L->store);
    else {
        for (size_t i = 0; i < L->n; i++) free_call(&L->calls[i]);
        free(L->calls);
    }
    L->store = NULL;
    L->calls = NULL;
    L->n = L->cap = 0;
}

/* One sealed call of a render list recorded as an array draw of its
 * vertices, when every primitive becomes a filled triangle and the vertex
 * colours do not drive the material (PF_COLOR_MATERIAL would change the
 * state vertex by vertex): the frame references the list's block and its
 * device copy, and takes a reference on the store.  Returns 0 when the call
 * must be replayed through immediate mode. */
static int call_as_array_draw(Ctx* c, ListStore* st, const ListCall* k)
{
    if (!g_arrayDraws || !k->packed || (c->state & PF_COLOR_MATERIAL) || k->drawMode < PF_TRIANGLES
        || k->drawMode > PF_QUAD_STRIP || k->n > 0xFFFFFF00u)
        return 0;
    uint32_t groups, passes[2];
    if (!face_groups(c, k->drawMode, passes, &groups)) return 0;
    memcpy(c->mat, k->faceMaterial, sizeof c->mat);         /* renderlist.c:82-83 */
    c->stateDirty = 1;
    pfBindTexture((PFtexture)k->texture);
    pfBegin(k->drawMode);
    const uint64_t tris = mode_triangles(k->drawMode, k->n) * groups;
    if (tris) {
        Rec* r = &c->rec;
        int held = 0;
        for (size_t i = 0; i < r->nstores && !held; i++) held = r->stores[i] == st;
        if (!held) {
            if (!grow((void**)&r->stores, &r->capstores, r->nstores + 1, sizeof(ListStore*))) {
                set_err(PF_ERROR_OUT_OF_MEMORY);
                pfEnd();
                return 1;
            }
            __atomic_add_fetch(&st->refs, 1, __ATOMIC_ACQ_REL);
            r->stores[r->nstores++] = st;
        }
        PFFarrays* d = push_array_draw(c, k->drawMode, (uint32_t)k->n, tris, groups, passes, k->packed, 0);
        if (d) {
            d->posSize = 4; d->posType = PFF_FLOAT;
            d->normal = k->nrm; d->nrmType = PFF_FLOAT;
            d->texcoord = k->uv; d->uvType = PFF_FLOAT;
            d->colors = k->col; d->colSize = 4; d->colType = PFF_UBYTE;
            d->numElements = (uint32_t)(k->n ? k->n : 1);
            d->flags = PFF_ARRAYS_LIST | PFF_ARRAYS_TEXMAT | ((c->state & PF_NORMALIZE) ? PFF_ARRAYS_NORMALIZE : 0u);
            const float* m = c->matTex;
            d->uvMatrix[0] = m[0]; d->uvMatrix[1] = m[1]; d->uvMatrix[2] = m[4];
            d->uvMatrix[3] = m[5]; d->uvMatrix[4] = m[12]; d->uvMatrix[5] = m[13];
            d->device = k->resident ? pfdResidentPointer(k->resident) : NULL;
        }
    }
    pfEnd();
    return 1;
}

static void list_begin(Ctx* c, PFdrawmode mode)
{
    List* L = c->curList;
    if (L->n == L->cap) {
        size_t cap = L->cap ? 2 * L->cap : 8;
        ListCall* p = (ListCall*)realloc(L->calls, cap * sizeof *p);
        if (!p) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
        L->calls = p; L->cap = cap;
    }
    ListCall* k = &L->calls[L->n++];
    memset(k, 0, sizeof *k);
    memcpy(k->faceMaterial, c->mat, sizeof c->mat);
    k->texture = c->curTex;
    k->drawMode = mode;                /* checked when the list is called */
}

static void list_vertex(Ctx* c, const float* v)
{
    List* L = c->curList;
    /* A vertex with no pfBegin since pfNewList has no call to join (the
     * reference indexes before its first element). */
    if (L->n == 0) { c->err = PF_INVALID_OPERATION; return; }
    ListCall* k = &L->calls[L->n - 1];
    if (k->n == k->cap) {
        size_t cap = k->cap ? 2 * k->cap : 8;
        float* p = (float*)realloc(k->pos, cap * 4 * sizeof(float));
        if (p) k->pos = p;
        float* u = (float*)realloc(k->uv, cap * 2 * sizeof(float));
        if (u) k->uv = u;
        float* q = (float*)realloc(k->nrm, cap * 3 * sizeof(float));
        if (q) k->nrm = q;
        PFcolor* o = (PFcolor*)realloc(k->col, cap * sizeof(PFcolor));
        if (o) k->col = o;
        if (!p || !u || !q || !o) { c->err = PF_ERROR_OUT_OF_MEMORY; return; }
        k->cap = cap;
    }
    memcpy(k->pos + 4 * k->n, v, 4 * sizeof(float));
    memcpy(k->uv + 2 * k->n, c->curUV, 2 * sizeof(float));
    memcpy(k->nrm + 3 * k->n, c->curNormal, 3 * sizeof(float));
    k->col[k->n] = c->curColor;
    k->n++;
}

PFrenderlist pfGenList(void)                                 /* renderlist.c:6-15 */
{
    return (PFrenderlist)calloc(1, sizeof(List));
}

void pfDeleteList(PFrenderlist* renderList)                  /* renderlist.c:17-36 */
{
    if (!renderList) return;
    List* L = (List*)*renderList;
    if (L) {
        if (G && G->curList == L) G->curList = NULL;
        list_clear(L);
        free(L);
    }
    *renderList = NULL;
}

void pfNewList(PFrenderlist renderList)                      /* renderlist.c:38-57 */
{
    Ctx* c = G;
    if (!c) return;
    if (!renderList) { c->err = PF_INVALID_VALUE; return; }
    List* L = (List*)renderList;
    list_clear(L);
    c->curList = L;
    make_backup(c);
}

void pfEndList(void)                                         /* renderlist.c:59-67 */
{
    Ctx* c = G;
    if (!c) return;
    if (!c->curList) c->err = PF_INVALID_OPERATION;
    else seal_list(c->curList);
    c->curList = NULL;
    restore_backup(c);
}

void pfCallList(const PFrenderlist renderList)              /* renderlist.c:69-97 */
{
    Ctx* c = G;
    if (!c) return;
    /* The reference dereferences a NULL list. */
    if (!renderList) { c->err = PF_INVALID_VALUE; return; }
    List* L = (List*)renderList;
    make_backup(c);
    /* A sealed list called outside recording: each call whose triangles
     * the GPU can form from the list's resident vertices is recorded as an
     * array draw; the others replay through immediate mode below. */
    ListStore* st = (!c->curList && L->store) ? L->store : NULL;
    /* Calls made while recording into this same list append to it: walk the
     * count taken at entry. */
    const size_t ncalls = L->n;
    for (size_t i = 0; i < ncalls; i++) {
        const ListCall* k = &L->calls[i];
        if (st && call_as_array_draw(c, st, k)) continue;
        const size_t nv = k->n;
        memcpy(c->mat, k->faceMaterial, sizeof c->mat);
        c->stateDirty = 1;
        pfBindTexture((PFtexture)k->texture);
        pfBegin(k->drawMode);
        for (size_t j = 0; j < nv; j++) {
            const ListCall* kk = &L->calls[i];                /* the array may move */
            pfColor4ubv((const PFubyte*)&kk->col[j]);
            pfTexCoordfv(kk->uv + 2 * j);
            pfNormal3fv(kk->nrm + 3 * j);
            pfVertex4fv(kk->pos + 4 * j);
        }
        pfEnd();
    }
    restore_backup(c);
}

/* ------------------------------------------------------------------------ */
/* Readback and framebuffers                                                 */
/* ------------------------------------------------------------------------ */

static void finish_target(const void* colorPixels)
{
    if (G && G->rec.ncmd && G->rec.targetColor == colorPixels) submit(G);
}

/* pfmFloatToHalfI (pfm.h:107-126): mantissa rounded half up, results below
 * 2^-14 flushed to signed zero, overflow to infinity, every NaN to 0x7E00. */
static uint16_t float_to_half_ref(float x)
{
    uint32_t ui;
    memcpy(&ui, &x, 4);
    int32_t s = (int32_t)((ui >> 16) & 0x8000u);
    int32_t em = (int32_t)(ui & 0x7FFFFFFFu);
    int32_t h = (em - (112 << 23) + (1 << 12)) >> 13;
    if (em < (113 << 23)) h = 0;
    if (em >= (143 << 23)) h = 0x7C00;
    if (em > (255 << 23)) h = 0x7E00;
    return (uint16_t)(s | h);
}

/* PF_COLOR_GARYSCALE (pixel.h:69-74), in float, left to right. */
static float gray_of(PFcolor c)
{
    const float inv = (float)PF_INV_255_D;
    return c.r * inv * 0.299f + c.g * inv * 0.587f + c.b * inv * 0.114f;
}

/* One pixel through the reference's scalar setter of (format, type)
 * (GC_pixelSetters, pixel.h:815-858; setters pixel.h:78-400).  The
 * normalising macros multiply by (float)(1/255) (pixel.h:39-67) -- the BGRA
 * one does not swap red and blue, so the BGRA 5_5_5_1, 4_4_4_4, HALF and
 * FLOAT setters store RGBA order -- while the single-channel HALF/FLOAT
 * setters divide by 255.0f. */
static void set_pixel(void* dst, size_t i, PFpixelformat f, PFdatatype t, PFcolor c)
{
    const float inv = (float)PF_INV_255_D;
    uint8_t* b = (uint8_t*)dst;
    uint16_t* h = (uint16_t*)dst;
    float* fl = (float*)dst;
    if (f <= PF_ALPHA) {                                                   /* pixel.h:131-196 */
        const uint8_t v = f == PF_RED ? c.r : f == PF_GREEN ? c.g : f == PF_BLUE ? c.b : c.a;
        if (t == PF_UNSIGNED_BYTE) b[i] = v;
        else if (t == PF_HALF_FLOAT) h[i] = float_to_half_ref(v / 255.0f);
        else fl[i] = v / 255.0f;
        return;
    }
    if (f == PF_LUMINANCE) {                                               /* pixel.h:78-97 */
        const float g = gray_of(c);
        if (t == PF_UNSIGNED_BYTE) b[i] = (uint8_t)(255.0f * g);
        else if (t == PF_HALF_FLOAT) h[i] = float_to_half_ref(g);
        else fl[i] = g;
        return;
    }
    if (f == PF_LUMINANCE_ALPHA) {                                         /* pixel.h:102-127 */
        const float g = gray_of(c);
        if (t == PF_UNSIGNED_BYTE) { b[2 * i] = (uint8_t)(255 * g); b[2 * i + 1] = c.a; }
        else if (t == PF_HALF_FLOAT) { h[2 * i] = float_to_half_ref(g); h[2 * i + 1] = float_to_half_ref(c.a * inv); }
        else { fl[2 * i] = g; fl[2 * i + 1] = c.a * inv; }
        return;
    }
    if (f == PF_RGB || f == PF_BGR) {                                      /* pixel.h:206-262 */
        const int swap = f == PF_BGR;
        const uint8_t c0 = swap ? c.b : c.r, c2 = swap ? c.r : c.b;
        const float n0 = c0 * inv, n1 = c.g * inv, n2 = c2 * inv;
        if (t == PF_UNSIGNED_BYTE) { b[3 * i] = c0; b[3 * i + 1] = c.g; b[3 * i + 2] = c2; }
        else if (t == PF_UNSIGNED_SHORT_5_6_5)
            h[i] = (uint16_t)(((uint16_t)(uint8_t)roundf(n0 * 31.0f) << 11) | ((uint16_t)(uint8_t)roundf(n1 * 63.0f) << 5)
                              | (uint16_t)(uint8_t)roundf(n2 * 31.0f));
        else if (t == PF_HALF_FLOAT) { h[3 * i] = float_to_half_ref(n0); h[3 * i + 1] = float_to_half_ref(n1); h[3 * i + 2] = float_to_half_ref(n2); }
        else { fl[3 * i] = n0; fl[3 * i + 1] = n1; fl[3 * i + 2] = n2; }
        return;
    }
    /* RGBA / BGRA (pixel.h:293-400). */
    const float n0 = c.r * inv, n1 = c.g * inv, n2 = c.b * inv, n3 = c.a * inv;
    if (t == PF_UNSIGNED_BYTE) {
        if (f == PF_RGBA) { b[4 * i] = c.r; b[4 * i + 1] = c.g; b[4 * i + 2] = c.b; b[4 * i + 3] = c.a; }
        else { b[4 * i] = c.b; b[4 * i + 1] = c.g; b[4 * i + 2] = c.r; b[4 * i + 3] = c.a; }
    } else if (t == PF_UNSIGNED_SHORT_5_5_5_1) {
        const uint16_t a1 = (double)n3 > (double)(float)50 * PF_INV_255_D ? 1 : 0;   /* config.h:45-47 */
        h[i] = (uint16_t)(((uint16_t)(uint8_t)roundf(n0 * 31.0f) << 11) | ((uint16_t)(uint8_t)roundf(n1 * 31.0f) << 6)
                          | ((uint16_t)(uint8_t)roundf(n2 * 31.0f) << 1) | a1);
    } else if (t == PF_UNSIGNED_SHORT_4_4_4_4) {
        h[i] = (uint16_t)(((uint16_t)(uint8_t)roundf(n0 * 15.0f) << 12) | ((uint16_t)(uint8_t)roundf(n1 * 15.0f) << 8)
                          | ((uint16_t)(uint8_t)roundf(n2 * 15.0f) << 4) | (uint16_t)(uint8_t)roundf(n3 * 15.0f));
    } else if (t == PF_HALF_FLOAT) {
        h[4 * i] = float_to_half_ref(n0); h[4 * i + 1] = float_to_half_ref(n1);
        h[4 * i + 2] = float_to_half_ref(n2); h[4 * i + 3] = float_to_half_ref(n3);
    } else {
        fl[4 * i] = n0; fl[4 * i + 1] = n1; fl[4 * i + 2] = n2; fl[4 * i + 3] = n3;
    }
}

/* pfmHalfToFloatI (pfm.h:128-145): subnormals flush to signed zero,
 * infinities and NaNs keep their payload. */
static float half_to_float_ref(uint16_t h)
{
    const uint32_t s = (uint32_t)(h & 0x8000u) << 16;
    const int32_t em = (int32_t)(h & 0x7FFFu);
    int32_t r = (em + (112 << 10)) << 13;
    if (em < (1 << 10)) r = 0;
    if (em >= (31 << 10)) r += 112 << 23;
    const uint32_t u = s | (uint32_t)r;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

/* One pixel of a render target through the reference's scalar getter of
 * its pair (pixel.h:441-457, 549-759): RGB/BGR bytes; the 16-bit fields
 * scaled by 255.0f/31, /63 or /15 in float and truncated; half and float
 * components as (PFubyte)(v * 255). */
static PFcolor get_pixel(const void* src, size_t i, PFpixelformat f, PFdatatype t)
{
    const uint8_t* b = (const uint8_t*)src;
    if (t == PF_HALF_FLOAT || t == PF_FLOAT) {
        const size_t n = f == PF_LUMINANCE_ALPHA ? 2 : (f == PF_RGB || f == PF_BGR) ? 3 : 4;
        uint8_t c[4] = { 0, 0, 0, 0 };
        for (size_t k = 0; k < n; k++) {
            float v;
            if (t == PF_HALF_FLOAT) { uint16_t h; memcpy(&h, b + 2 * (n * i + k), 2); v = half_to_float_ref(h); }
            else memcpy(&v, b + 4 * (n * i + k), 4);
            c[k] = f2u8(v * 255.0f);
        }
        switch (f) {
        case PF_LUMINANCE_ALPHA: return (PFcolor){ c[0], c[0], c[0], c[1] };   /* pixel.h:441-457 */
        case PF_RGB:  return (PFcolor){ c[0], c[1], c[2], 255 };             /* pixel.h:589-639 */
        case PF_BGR:  return (PFcolor){ c[2], c[1], c[0], 255 };
        case PF_BGRA: return (PFcolor){ c[2], c[1], c[0], c[3] };            /* pixel.h:709-759 */
        default:      return (PFcolor){ c[0], c[1], c[2], c[3] };
        }
    }
    if (t == PF_UNSIGNED_BYTE) {
        if (f == PF_RGB) return (PFcolor){ b[3 * i], b[3 * i + 1], b[3 * i + 2], 255 };
        if (f == PF_BGR) return (PFcolor){ b[3 * i + 2], b[3 * i + 1], b[3 * i], 255 };
        PFcolor c;
        memcpy(&c, b + 4 * i, 4);
        return c;
    }
    const uint32_t x = ((const uint16_t*)src)[i];
    const float s5 = 255.0f / 31, s6 = 255.0f / 63, s4 = 255.0f / 15;
    uint8_t hi, mid, lo, a;
    if (t == PF_UNSIGNED_SHORT_5_6_5) {
        hi = (uint8_t)((float)((x & 0xF800u) >> 11) * s5);
        mid = (uint8_t)((float)((x & 0x07E0u) >> 5) * s6);
        lo = (uint8_t)((float)(x & 0x001Fu) * s5);
        a = 255;
    } else if (t == PF_UNSIGNED_SHORT_5_5_5_1) {
        hi = (uint8_t)((float)((x & 0xF800u) >> 11) * s5);
        mid = (uint8_t)((float)((x & 0x07C0u) >> 6) * s5);
        lo = (uint8_t)((float)((x & 0x003Eu) >> 1) * s5);
        a = (uint8_t)((x & 1u) * 255u);
    } else {
        hi = (uint8_t)((float)((x & 0xF000u) >> 12) * s4);
        mid = (uint8_t)((float)((x & 0x0F00u) >> 8) * s4);
        lo = (uint8_t)((float)((x & 0x00F0u) >> 4) * s4);
        a = (uint8_t)((float)(x & 0x000Fu) * s4);
    }
    return (f == PF_BGR || f == PF_BGRA) ? (PFcolor){ lo, mid, hi, a } : (PFcolor){ hi, mid, lo, a };
}

void pfReadPixels(PFint x, PFint y, PFsizei width, PFsizei height,
                  PFpixelformat format, PFdatatype type, void* pixels)   /* context.c:2349-2395 */
{
    Ctx* c = G;
    if (!c) return;
    /* Every pair with a pixel setter (pixel.h:815-858); the reference checks
     * the same table (context.c:2351-2361). */
    if (pff_source_texel_bytes(PFF_SOURCE(format, type)) == 0u) { c->err = PF_INVALID_ENUM; return; }
    submit(c);
    const Tex* tt = fb_tex(c->curFB);
    if (!(format == PF_RGBA && type == PF_UNSIGNED_BYTE && tt->format == PF_RGBA && tt->type == PF_UNSIGNED_BYTE)) {
        const Tex* t = tt;
        int32_t xMin = x < 0 ? 0 : (x > (int32_t)t->w - 1 ? (int32_t)t->w - 1 : x);
        int32_t yMin = y < 0 ? 0 : (y > (int32_t)t->h - 1 ? (int32_t)t->h - 1 : y);
        int64_t xe = (int64_t)x + (int32_t)width, ye = (int64_t)y + (int32_t)height;
        int32_t xMax = (int32_t)(xe < 0 ? 0 : (xe > t->w ? t->w : xe));
        int32_t yMax = (int32_t)(ye < 0 ? 0 : (ye > t->h ? t->h : ye));
        for (int32_t ys = yMin, yd = 0; ys < yMax; ys++, yd++)
            for (int32_t xs = xMin, xd = 0; xs < xMax; xs++, xd++)
                set_pixel(pixels, (size_t)yd * width + xd, format, type,
                          get_pixel(t->pixels, (size_t)ys * t->w + xs, t->format, t->type));
        return;
    }
    const Tex* t = fb_tex(c->curFB);
    int32_t xMin = x < 0 ? 0 : (x > (int32_t)t->w - 1 ? (int32_t)t->w - 1 : x);
    int32_t yMin = y < 0 ? 0 : (y > (int32_t)t->h - 1 ? (int32_t)t->h - 1 : y);
    int64_t xe = (int64_t)x + (int32_t)width, ye = (int64_t)y + (int32_t)height;
    int32_t xMax = (int32_t)(xe < 0 ? 0 : (xe > t->w ? t->w : xe));
    int32_t yMax = (int32_t)(ye < 0 ? 0 : (ye > t->h ? t->h : ye));
    const uint32_t* src = (const uint32_t*)t->pixels;
    uint32_t* dst = (uint32_t*)pixels;
    for (int32_t ys = yMin, yd = 0; ys < yMax; ys++, yd++)
        for (int32_t xs = xMin, xd = 0; xs < xMax; xs++, xd++)
            dst[(size_t)yd * width + xd] = src[(size_t)ys * t->w + xs];
}

PFframebuffer pfGenFramebuffer(PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type)
{
    PFframebuffer fb = { NULL, NULL };
    const uint32_t kind = pff_target_of(PFF_SOURCE(format, type));
    if (kind == PFF_TARGET_NONE) { set_err(PF_INVALID_ENUM); return fb; }
    size_t n = (size_t)width * height;
    const size_t bytes = (n ? n : 1) * pff_target_bytes(kind);
    void* px = pfdHostAlloc(bytes);                 /* page-locked with a GPU: DMA downloads */
    if (!px) return fb;
    memset(px, 0, bytes);
    PFtexture t = pfGenTexture(px, width, height, format, type);
    if (!t) { pfdHostFree(px); set_err(PF_INVALID_ENUM); return fb; }
    ((Tex*)t)->hostPixels = 1;
    float* z = (float*)pfdHostAlloc((n ? n : 1) * sizeof(float));
    if (!z) { set_err(PF_ERROR_OUT_OF_MEMORY); pfDeleteTexture(&t, PF_TRUE); return fb; }
    for (size_t i = 0; i < n; i++) z[i] = FLT_MAX;
    fb.texture = t;
    fb.zbuffer = z;
    return fb;
}

void pfDeleteFramebuffer(PFframebuffer* fb)
{
    if (!fb) return;
    Tex* t = fb_tex(fb);
    if (t) finish_target(t->pixels);
    if (G && G->boundFB == fb) G->boundFB = NULL;
    if (G && G->curFB == fb) G->curFB = &G->mainFB;
    pfDeleteTexture(&fb->texture, PF_TRUE);
    pfdHostFree(fb->zbuffer);
    fb->zbuffer = NULL;
    fb->texture = NULL;
}

PFboolean pfIsValidFramebuffer(PFframebuffer* fb)
{
    Tex* t = fb ? fb_tex(fb) : NULL;
    return t && fb->zbuffer && t->w > 0 && t->h > 0 && pfIsValidTexture(fb->texture);
}

void pfClearFramebuffer(PFframebuffer* fb, PFcolor color, PFfloat depth)   /* framebuffer.c:86-100 */
{
    Tex* t = fb_tex(fb);
    if (!t) return;
    finish_target(t->pixels);
    size_t n = (size_t)t->w * t->h;
    for (size_t i = 0; i < n; i++) { set_pixel(t->pixels, i, t->format, t->type, color); fb->zbuffer[i] = depth; }
}

PFcolor pfGetFramebufferPixel(const PFframebuffer* fb, PFsizei x, PFsizei y)
{
    const Tex* t = fb_tex(fb);
    finish_target(t->pixels);
    return get_pixel(t->pixels, (size_t)y * t->w + x, t->format, t->type);
}

PFfloat pfGetFramebufferDepth(const PFframebuffer* fb, PFsizei x, PFsizei y)
{
    const Tex* t = fb_tex(fb);
    finish_target(t->pixels);
    return fb->zbuffer[(size_t)y * t->w + x];
}

/* framebuffer.c:115-148.  The scalar depth test: NOTEQUAL is a real !=
 * here, unlike the triangle path (depth.h:35-71). */
void pfSetFramebufferPixelDepthTest(PFframebuffer* fb, PFsizei x, PFsizei y, PFfloat z, PFcolor color,
                                    PFdepthmode depthMode)
{
    if (depthMode > PF_GEQUAL) { set_err(PF_INVALID_ENUM); return; }
    Tex* t = fb_tex(fb);
    finish_target(t->pixels);
    size_t off = (size_t)y * t->w + x;
    float d = fb->zbuffer[off];
    int pass = 0;
    switch (depthMode) {
    case PF_EQUAL: pass = z == d; break;
    case PF_NOTEQUAL: pass = z != d; break;
    case PF_LESS: pass = z < d; break;
    case PF_LEQUAL: pass = z <= d; break;
    case PF_GREATER: pass = z > d; break;
    case PF_GEQUAL: pass = z >= d; break;
    }
    if (pass) {
        set_pixel(t->pixels, off, t->format, t->type, color);
        fb->zbuffer[off] = z;
    }
}

void pfSetFramebufferPixelDepth(PFframebuffer* fb, PFsizei x, PFsizei y, PFfloat z, PFcolor color)
{
    Tex* t = fb_tex(fb);
    finish_target(t->pixels);
    size_t off = (size_t)y * t->w + x;
    set_pixel(t->pixels, off, t->format, t->type, color);
    fb->zbuffer[off] = z;
}

void pfSetFramebufferPixel(PFframebuffer* fb, PFsizei x, PFsizei y, PFcolor color)
{
    Tex* t = fb_tex(fb);
    finish_target(t->pixels);
    set_pixel(t->pixels, (size_t)y * t->w + x, t->format, t->type, color);
}

/* ------------------------------------------------------------------------ */
/* New entry points                                                          */
/* ------------------------------------------------------------------------ */

void pfFinish(void) { if (G) submit(G); }
void pfFlush(void) { if (G) submit(G); }

void pfSubmitBatch(const PFcontext* contexts, PFsizei count)
{
    if (!contexts || count == 0) return;
    PFFframe* frames = (PFFframe*)calloc(count, sizeof(PFFframe));
    Ctx** owners = (Ctx**)calloc(count, sizeof(Ctx*));
    if (!frames || !owners) { free(frames); free(owners); set_err(PF_ERROR_OUT_OF_MEMORY); return; }
    uint32_t n = 0;
    for (PFsizei i = 0; i < count; i++) {
        Ctx* c = (Ctx*)contexts[i];
        if (!c || c->rec.ncmd == 0) continue;
        build_frame(c, &frames[n]);
        owners[n++] = c;
    }
    int rc = 0;
    if (n) {
        if (g_batchExec) {
            rc = g_batchExec(frames, n, g_batchUser);
        } else if (g_exec) {
            for (uint32_t i = 0; i < n && rc == 0; i++) rc = g_exec(&frames[i], g_execUser);
        } else {
            rc = pfdExecuteFrames(frames, n);
            if (rc != 0) fprintf(stderr, "pixelforge_amd: batch execution failed (%d): %s\n", rc, pfdLastError());
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        Ctx* c = owners[i];
        if (rc != 0) c->err = (rc == PFD_E_NOMEM) ? PF_ERROR_OUT_OF_MEMORY : PF_INVALID_OPERATION;
        c->frames++;
        rec_reset(&c->rec);
        c->stateDirty = 1;
    }
    free(frames);
    free(owners);
}

void pfxSetExecutor(PFXexecutor fn, void* user)
{
    g_exec = fn;
    g_execUser = user;
}

void pfxSetBatchExecutor(PFXbatchExecutor fn, void* user)
{
    g_batchExec = fn;
    g_batchUser = user;
}

PFuint pfxFrameCount(void) { return G ? G->frames : 0; }

void pfxSetArrayDraws(PFboolean enabled) { g_arrayDraws = enabled ? 1 : 0; }

/* ------------------------------------------------------------------------ */
/* Differentiable rendering                                                  */
/* ------------------------------------------------------------------------ */

static int diff_params(Ctx* c, const PFdiffparams* p, PFDdiffParams* d)
{
    memset(d, 0, sizeof *d);
    d->texture = -1;
    if (!p) return 1;
    if (p->texture) {
        for (uint32_t i = 0; i < c->diffNumTex; i++)
            if (c->diffTex[i] == (const Tex*)p->texture) d->texture = (int32_t)i;
        if (d->texture < 0) return 0;          /* not used by the frame */
    }
    d->texels = p->texels;
    d->numVertices = p->numVertices;
    d->vertexColors = p->vertexColors;
    d->lightDiffuse = p->lightDiffuse;
    d->lightSpecular = p->lightSpecular;
    return 1;
}

static void diff_fail(Ctx* c, int rc, const char* what)
{
    c->err = (rc == PFD_E_NOMEM) ? PF_ERROR_OUT_OF_MEMORY : PF_INVALID_OPERATION;
    fprintf(stderr, "pixelforge_amd: %s failed (%d): %s\n", what, rc, pfdLastError());
}

void pfDiffRender(const PFdiffparams* params, PFfloat* image)
{
    Ctx* c = G;
    if (!c) return;
    if (c->rec.ncmd) {
        PFFframe f;
        build_frame(c, &f);
        pfdJobDestroy(c->diffJob);
        c->diffJob = NULL;
        free(c->diffTex);
        c->diffTex = NULL;
        c->diffNumTex = 0;
        int rc = pfdDiffCreate(&f, &c->diffJob);
        if (rc == 0) {
            c->diffTex = (const Tex**)malloc((c->rec.ntex ? c->rec.ntex : 1) * sizeof(Tex*));
            if (c->diffTex) {
                memcpy(c->diffTex, c->rec.texKeys, c->rec.ntex * sizeof(Tex*));
                c->diffNumTex = (uint32_t)c->rec.ntex;
            }
        }
        c->frames++;
        rec_reset(&c->rec);
        c->stateDirty = 1;
        if (rc) { diff_fail(c, rc, "pfDiffRender"); return; }
    }
    if (!c->diffJob) { c->err = PF_INVALID_OPERATION; return; }
    PFDdiffParams d;
    if (!diff_params(c, params, &d)) { c->err = PF_INVALID_VALUE; return; }
    int rc = pfdDiffForward(c->diffJob, &d, image);
    if (rc) diff_fail(c, rc, "pfDiffRender");
}

void pfDiffBackward(const PFdiffparams* params, const PFfloat* dImage, PFdiffgrads* grads)
{
    Ctx* c = G;
    if (!c) return;
    if (!c->diffJob || !dImage || !grads) { c->err = PF_INVALID_OPERATION; return; }
    PFDdiffParams d;
    if (!diff_params(c, params, &d)) { c->err = PF_INVALID_VALUE; return; }
    PFDdiffGrads g = { grads->texels, grads->vertexColors, grads->lightDiffuse, grads->lightSpecular };
    int rc = pfdDiffBackward(c->diffJob, &d, dImage, &g);
    if (rc) diff_fail(c, rc, "pfDiffBackward");
}

/* Device-resident timing of the last pfDiffRender frame (for benchmarks). */
int pfxDiffTime(const PFdiffparams* params, PFuint iters, float* rasterMs, float* forwardMs, float* backwardMs)
{
    Ctx* c = G;
    if (!c || !c->diffJob) return PFD_E_INVALID;
    PFDdiffParams d;
    if (!diff_params(c, params, &d)) return PFD_E_INVALID;
    PFDdiffTimes t;
    int rc = pfdDiffTime(c->diffJob, &d, iters, &t);
    if (rc) return rc;
    *rasterMs = t.rasterMs;
    *forwardMs = t.forwardMs;
    *backwardMs = t.backwardMs;
    return 0;
}
