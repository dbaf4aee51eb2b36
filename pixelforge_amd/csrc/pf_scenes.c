/*
 * pf_scenes.c -- seeded synthetic scenes issued through the PixelForge C API.
 *
 * The same scene code drives any library that exports the PixelForge API:
 * libpixelforge_amd.so (GPU) or a CPU build of the reference.  The API is
 * bound at run time with dlsym, so one scene function renders identical call
 * sequences on every backend.  Randomness comes from SplitMix64 (Steele,
 * Lea, Flood 2014) seeded per scene; floats are (u >> 40) * 2^-24.
 *
 * Scene kinds:
 *   PFS_RANDOM_NDC  triangles with vertices uniform in NDC [-1,1]^3 and the
 *                   default matrices (the reference's 2D path), random RGBA8
 *                   vertex colours; optional texture, blending, depth func.
 *   PFS_HEIGHTFIELD jittered height-field grid seen through a perspective
 *                   camera, textured, with a fraction of blended triangles
 *                   drawn after the opaque ones.
 *   PFS_SPHERES     cluster of tessellated spheres with normals, lit per
 *                   fragment (Phong) or per vertex (Gouraud).
 *   PFS_RANDOM_3D   random triangles in a perspective frustum, including
 *                   vertices behind the camera and outside the side planes,
 *                   so that clipping and both culling passes run.
 *   PFS_MESH        indexed grid mesh facing a perspective camera, with
 *                   per-vertex normals, texture coordinates and float
 *                   colours, drawn through pfDrawElements, pfDrawArrays or
 *                   the equivalent immediate-mode calls.
 *   PFS_LIST        spheres recorded once into render lists and called at
 *                   several model-view transforms (pfNewList / pfEndList /
 *                   pfCallList), including a list recorded by calling another.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <time.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pixelforge.h"

#define PFS_RANDOM_NDC   1
#define PFS_HEIGHTFIELD  2
#define PFS_SPHERES      3
#define PFS_RANDOM_3D    4
#define PFS_MESH         5
#define PFS_LIST         6

/* Scene description; zero fields take the defaults noted. */
typedef struct {
    uint32_t kind;
    uint32_t width, height;
    uint32_t ntris;          /* random kinds: triangle count                        */
    uint64_t seed;
    int32_t  cull;           /* -1: disabled, PF_FRONT/PF_BACK: cull that face       */
    int32_t  depthTest;      /* 0/1 */
    int32_t  depthFunc;      /* PFdepthmode */
    int32_t  blend;          /* 0/1 */
    int32_t  blendMode;      /* PFblendmode */
    int32_t  texture;        /* 0 none, 1 RGBA8, 2 RGBA float */
    int32_t  texW, texH;
    int32_t  wrap, filter;
    int32_t  shade;          /* PFshademode */
    int32_t  lighting;       /* 0 none, 1 Gouraud, 2 Phong */
    int32_t  numLights;
    int32_t  spot;           /* 0/1: spot cut-offs on the lights */
    int32_t  attenuation;    /* 0/1: linear/quadratic attenuation */
    float    blendFraction;  /* heightfield: fraction of triangles drawn blended     */
    int32_t  gridX, gridZ;   /* heightfield cells                                    */
    int32_t  spheres;        /* spheres: count                                       */
    int32_t  sphereLat, sphereLon;
    int32_t  clearFirst;     /* 1: pfClear before drawing                            */
    int32_t  colorMaterial;  /* 1: pfColor drives the material                      */
    float    triScale;       /* random kinds: triangle size relative to the screen   */
    int32_t  flushEvery;     /* >0: pfFinish every N triangles (multi-frame)         */
    int32_t  useQuads;       /* 1: PF_QUADS for grids                                */
    int32_t  arrays;         /* mesh: 0 immediate mode, 1 pfDrawElements, 2 pfDrawArrays */
    int32_t  indexType;      /* mesh, pfDrawElements: 0 uint, 1 ubyte, 2 ushort      */
    int32_t  posType;        /* mesh arrays: 0 float, 1 double positions             */
    int32_t  colType;        /* mesh arrays: 0 float, 1 ubyte, 2 ushort, 3 uint, 4 double */
    int32_t  colSize;        /* mesh arrays: 3 or 4 colour components (0: 4)         */
    int32_t  texFormat;      /* texture = 3: PFpixelformat of the texture            */
    int32_t  texType;        /* texture = 3: PFdatatype of the texture               */
    int32_t  texSwizzle;     /* texture = 3, BGRA / UNSIGNED_BYTE: hand the same
                              * texels over as RGBA / UNSIGNED_BYTE instead      */
    int32_t  pixelOps;       /* >0: that many pfDrawPixels calls between the first
                              * and second half of the triangles (random kinds) */
    int32_t  pixFormat;      /* pfDrawPixels image PFpixelformat                  */
    int32_t  pixType;        /* pfDrawPixels image PFdatatype                     */
    int32_t  fog;            /* pfFogProcess after the scene (random kinds: between
                              * the two halves): 1 LINEAR, 2 EXP, 3 EXP2, 4 the
                              * setters' quirks (DENSITY stores the mode, an
                              * unknown mode through pfFogiv)                   */
    float    fogStart, fogEnd;
    int32_t  fogAlpha;       /* fog colour alpha, 0..255                          */
    int32_t  points;         /* >0: that many PF_POINTS vertices between the two
                              * halves of the triangles (random kinds)          */
    float    pointSize;      /* >0: pfPointSize                                   */
    int32_t  polyPoint;      /* bit 0: front faces PF_POINT, bit 1: back faces   */
    int32_t  lines;          /* >0: that many PF_LINES pairs between the two halves
                              * of the triangles (random kinds)                  */
    float    lineWidth;      /* >0: pfLineWidth                                   */
    int32_t  polyLine;       /* bit 0: front faces PF_LINE, bit 1: back faces    */
    int32_t  rects;          /* >0: that many pfRectf fills between the two halves
                              * of the triangles (random kinds)                  */
    float    uvMin, uvMax;   /* random kinds: texture coordinates uniform in
                              * [uvMin, uvMax) when uvMax > uvMin, else [-2, 2) */
    int32_t  listOptions;    /* lists: bit 0 a texture matrix (rotate, scale,
                              * translate) and bit 1 PF_NORMALIZE when they are
                              * called; bit 2 list A re-recorded empty after its
                              * last call (the lists are deleted before pfFinish
                              * in any case)                                    */
    int32_t  arrayMode;      /* mesh arrays: 0 PF_TRIANGLES; a PFdrawmode from
                              * PF_TRIANGLE_FAN to PF_QUAD_STRIP draws each row of
                              * cells as one call in that mode (strip order:
                              * v(i, j), v(i, j+1), v(i+1, j), ...; PF_QUADS: the
                              * cells' four corners)                            */
    int32_t  fbFormat;       /* >0: PFpixelformat of the render target (with
                              * fbType); 0: RGBA / UNSIGNED_BYTE               */
    int32_t  fbType;         /* PFdatatype of the render target               */
    int32_t  fbMain;         /* 1: draw into the context's main buffer, created
                              * in that format, instead of an offscreen
                              * framebuffer (no depth is copied out)          */
} PFSparams;

typedef struct {
    void* h;
    PFcontext (*CreateContext)(void*, PFsizei, PFsizei, PFpixelformat, PFdatatype);
    void (*DeleteContext)(PFcontext);
    void (*MakeCurrent)(PFcontext);
    void (*Enable)(PFstate);
    void (*Disable)(PFstate);
    void (*MatrixMode)(PFmatrixmode);
    void (*LoadIdentity)(void);
    void (*Translatef)(PFfloat, PFfloat, PFfloat);
    void (*Scalef)(PFfloat, PFfloat, PFfloat);
    void (*Rotatef)(PFfloat, PFfloat, PFfloat, PFfloat);
    void (*Frustum)(PFfloat, PFfloat, PFfloat, PFfloat, PFfloat, PFfloat);
    void (*CullFace)(PFface);
    void (*BlendFunc)(PFblendmode);
    void (*DepthFunc)(PFdepthmode);
    void (*ShadeModel)(PFshademode);
    void (*LightModel)(PFlightmode);
    void (*BindTexture)(PFtexture);
    void (*BindFramebuffer)(PFframebuffer*);
    void (*Clear)(PFclearflag);
    void (*ClearColor)(PFubyte, PFubyte, PFubyte, PFubyte);
    void (*EnableLight)(PFsizei);
    void (*DisableLight)(PFsizei);
    void (*Lightf)(PFsizei, PFenum, PFfloat);
    void (*Lightfv)(PFsizei, PFenum, const void*);
    void (*Materialf)(PFface, PFenum, PFfloat);
    void (*Materialfv)(PFface, PFenum, const void*);
    void (*ColorMaterial)(PFface, PFenum);
    void (*Begin)(PFdrawmode);
    void (*End)(void);
    void (*Vertex3f)(PFfloat, PFfloat, PFfloat);
    void (*Color4ub)(PFubyte, PFubyte, PFubyte, PFubyte);
    void (*TexCoord2f)(PFfloat, PFfloat);
    void (*Normal3f)(PFfloat, PFfloat, PFfloat);
    PFframebuffer (*GenFramebuffer)(PFsizei, PFsizei, PFpixelformat, PFdatatype);
    void (*DeleteFramebuffer)(PFframebuffer*);
    PFtexture (*GenTexture)(void*, PFsizei, PFsizei, PFpixelformat, PFdatatype);
    void (*DeleteTexture)(PFtexture*, PFboolean);
    void (*TextureParameter)(PFtexture, PFtexturewrap, PFtexturefilter);
    void* (*GetTexturePixels)(const PFtexture, PFsizei*, PFsizei*, PFpixelformat*, PFdatatype*);
    PFerrcode (*GetError)(void);
    void (*Color4f)(PFfloat, PFfloat, PFfloat, PFfloat);
    void (*VertexPointer)(PFint, PFenum, PFsizei, const void*);
    void (*NormalPointer)(PFenum, PFsizei, const void*);
    void (*TexCoordPointer)(PFenum, PFsizei, const void*);
    void (*ColorPointer)(PFint, PFenum, PFsizei, const void*);
    void (*DrawElements)(PFdrawmode, PFsizei, PFdatatype, const void*);
    void (*DrawArrays)(PFdrawmode, PFint, PFsizei);
    void (*PointSize)(PFfloat);
    void (*LineWidth)(PFfloat);
    void (*Rectf)(PFfloat, PFfloat, PFfloat, PFfloat);
    void (*PolygonMode)(PFface, PFpolygonmode);
    void (*Fogi)(PFfogparam, PFint);
    void (*Fogf)(PFfogparam, PFfloat);
    void (*Fogiv)(PFfogparam, PFint*);
    void (*Fogfv)(PFfogparam, PFfloat*);
    void (*FogProcess)(void);
    void (*RasterPos3f)(PFfloat, PFfloat, PFfloat);
    void (*PixelZoom)(PFfloat, PFfloat);
    void (*DrawPixels)(PFsizei, PFsizei, PFpixelformat, PFdatatype, const void*);
    PFrenderlist (*GenList)(void);
    void (*DeleteList)(PFrenderlist*);
    void (*NewList)(PFrenderlist);
    void (*EndList)(void);
    void (*CallList)(const PFrenderlist);
    void (*Finish)(void);                    /* absent from the reference */
    void (*SubmitBatch)(const PFcontext*, PFsizei);   /* absent from the reference */
} PFSapi;

#define BIND(name) do { a->name = (void*)dlsym(h, "pf" #name); if (!a->name) goto fail; } while (0)

PFSapi* pfs_bind(const char* path)
{
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return NULL;
    PFSapi* a = (PFSapi*)calloc(1, sizeof(PFSapi));
    if (!a) return NULL;
    a->h = h;
    BIND(CreateContext); BIND(DeleteContext); BIND(MakeCurrent); BIND(Enable); BIND(Disable);
    BIND(MatrixMode); BIND(LoadIdentity); BIND(Translatef); BIND(Scalef); BIND(Rotatef); BIND(Frustum);
    BIND(CullFace); BIND(BlendFunc); BIND(DepthFunc); BIND(ShadeModel); BIND(LightModel);
    BIND(BindTexture); BIND(BindFramebuffer); BIND(Clear); BIND(ClearColor); BIND(EnableLight);
    BIND(DisableLight);
    BIND(Lightf); BIND(Lightfv); BIND(Materialf); BIND(Materialfv); BIND(ColorMaterial);
    BIND(Begin); BIND(End); BIND(Vertex3f); BIND(Color4ub); BIND(TexCoord2f); BIND(Normal3f);
    BIND(GenFramebuffer); BIND(DeleteFramebuffer); BIND(GenTexture); BIND(DeleteTexture);
    BIND(TextureParameter); BIND(GetTexturePixels); BIND(GetError); BIND(Color4f);
    BIND(VertexPointer); BIND(NormalPointer); BIND(TexCoordPointer); BIND(ColorPointer);
    BIND(DrawElements); BIND(DrawArrays);
    BIND(RasterPos3f); BIND(PixelZoom); BIND(DrawPixels);
    BIND(Fogi); BIND(Fogf); BIND(Fogiv); BIND(Fogfv); BIND(FogProcess);
    BIND(PointSize); BIND(PolygonMode); BIND(LineWidth); BIND(Rectf);
    BIND(GenList); BIND(DeleteList); BIND(NewList); BIND(EndList); BIND(CallList);
    a->Finish = (void (*)(void))dlsym(h, "pfFinish");
    a->SubmitBatch = (void (*)(const PFcontext*, PFsizei))dlsym(h, "pfSubmitBatch");
    return a;
fail:
    free(a);
    return NULL;
}

const char* pfs_dlerror(void) { return dlerror(); }

/* ------------------------------------------------------------------------ */
/* SplitMix64                                                                */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t s; } Rng;

static uint64_t rnext(Rng* r)
{
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float runit(Rng* r) { return (float)(rnext(r) >> 40) * (1.0f / 16777216.0f); }     /* [0,1) */
static float rrange(Rng* r, float lo, float hi) { return lo + (hi - lo) * runit(r); }
static uint8_t rbyte(Rng* r) { return (uint8_t)(rnext(r) >> 56); }

/* Exported so that harnesses can reproduce the stream. */
uint64_t pfs_splitmix64(uint64_t* state)
{
    Rng r = { *state };
    uint64_t v = rnext(&r);
    *state = r.s;
    return v;
}

/* ------------------------------------------------------------------------ */
/* Scene pieces                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    PFtexture tex;
    void*     pixels;
} SceneTex;

/* Bytes per texel of a (format, type) pair with a texel getter. */
static size_t texel_bytes(int f, int t)
{
    size_t comps = f <= PF_LUMINANCE ? 1 : f == PF_LUMINANCE_ALPHA ? 2 : (f == PF_RGB || f == PF_BGR) ? 3 : 4;
    if (t == PF_UNSIGNED_SHORT_5_6_5 || t == PF_UNSIGNED_SHORT_5_5_5_1 || t == PF_UNSIGNED_SHORT_4_4_4_4) return 2;
    return comps * (t == PF_UNSIGNED_BYTE ? 1 : t == PF_HALF_FLOAT ? 2 : 4);
}

/* Half-float bit patterns: mostly values in [0, 1], one in 16 a special
 * (above 1, negative, infinite, NaN, subnormal, 255). */
static uint16_t rand_half(Rng* r)
{
    static const uint16_t special[8] = { 0x3E00, 0xBC00, 0x7C00, 0x7E00, 0x0001, 0x5BF8, 0x3BFF, 0x8000 };
    uint64_t v = rnext(r);
    if ((v & 15u) == 0) return special[(v >> 4) & 7u];
    return (uint16_t)((v >> 16) % 0x3C01u);
}

static float rand_float(Rng* r)
{
    static const float special[8] = { 1.5f, -0.25f, 300.0f, 1e10f, -1e10f, 0.998f, 1.0f / 510.0f, 0.0f };
    uint64_t v = rnext(r);
    if ((v & 15u) == 0) {
        uint32_t k = (uint32_t)(v >> 4) & 7u;
        if (k == 7u) { uint32_t nan = 0x7FC00000u; float f; memcpy(&f, &nan, 4); return f; }
        return special[k];
    }
    return (float)(v >> 40) * (1.0f / 16777216.0f);
}

/* Random texel bytes of a (format, type) pair for n texels, 64 zero bytes
 * after the last. */
static uint8_t* random_texels(Rng* r, int f, int t, size_t n)
{
    size_t bpp = texel_bytes(f, t);
    uint8_t* px = (uint8_t*)calloc(n * bpp + 64, 1);
    if (!px) return NULL;
    size_t comps = (t == PF_HALF_FLOAT) ? n * bpp / 2 : (t == PF_FLOAT) ? n * bpp / 4 : 0;
    if (t == PF_HALF_FLOAT) for (size_t i = 0; i < comps; i++) ((uint16_t*)px)[i] = rand_half(r);
    else if (t == PF_FLOAT) for (size_t i = 0; i < comps; i++) ((float*)px)[i] = rand_float(r);
    else for (size_t i = 0; i < n * bpp; i++) px[i] = (uint8_t)(rnext(r) >> 56);
    return px;
}

/* pfDrawPixels calls: images of random size in the scene's pixel format at
 * random raster positions; the second call zooms by (2, 1.5) and blends,
 * the third zooms by (0.75, 0.5) without the depth test. */
static void draw_pixel_ops(const PFSapi* a, Rng* r, const PFSparams* p)
{
    for (int k = 0; k < p->pixelOps; k++) {
        uint32_t w = 5 + (uint32_t)(rnext(r) % 24), h = 4 + (uint32_t)(rnext(r) % 18);
        uint8_t* px = random_texels(r, p->pixFormat, p->pixType, (size_t)w * h);
        a->RasterPos3f(rrange(r, -0.9f, 0.5f), rrange(r, -0.5f, 0.9f), rrange(r, -0.6f, 0.6f));
        int kind = k % 3;
        a->PixelZoom(kind == 1 ? 2.0f : kind == 2 ? 0.75f : 1.0f, kind == 1 ? 1.5f : kind == 2 ? 0.5f : 1.0f);
        if (kind == 1) { a->Enable(PF_BLEND); a->BlendFunc((PFblendmode)((p->seed + k) % 8)); }
        if (kind == 2) a->Disable(PF_DEPTH_TEST);
        a->DrawPixels(w, h, (PFpixelformat)p->pixFormat, (PFdatatype)p->pixType, px);
        if (kind == 1) { a->Disable(PF_BLEND); a->BlendFunc((PFblendmode)p->blendMode); }
        if (kind == 2 && p->depthTest) a->Enable(PF_DEPTH_TEST);
        free(px);
    }
    a->PixelZoom(1.0f, 1.0f);
}

/* PF_POINTS: random positions in [-1, 1]^3 of the current matrices, random
 * colours. */
static void draw_points(const PFSapi* a, Rng* r, const PFSparams* p)
{
    a->Begin(PF_POINTS);
    for (int i = 0; i < p->points; i++) {
        a->Color4ub(rbyte(r), rbyte(r), rbyte(r), rbyte(r));
        a->Vertex3f(rrange(r, -1, 1), rrange(r, -1, 1), rrange(r, -1, 1));
    }
    a->End();
}

/* PF_LINES: random end points in [-1.3, 1.3]^3 of the current matrices (some
 * leave the viewport and are clipped), random colours. */
static void draw_lines(const PFSapi* a, Rng* r, const PFSparams* p)
{
    a->Begin(PF_LINES);
    for (int i = 0; i < 2 * p->lines; i++) {
        a->Color4ub(rbyte(r), rbyte(r), rbyte(r), rbyte(r));
        a->Vertex3f(rrange(r, -1.3f, 1.3f), rrange(r, -1.3f, 1.3f), rrange(r, -1.3f, 1.3f));
    }
    a->End();
}

/* pfRectf: corners in [-1.2, 1.2]^2 of the current matrices (some clamped
 * by the viewport), in either order, random colours. */
static void draw_rects(const PFSapi* a, Rng* r, const PFSparams* p)
{
    for (int i = 0; i < p->rects; i++) {
        a->Color4ub(rbyte(r), rbyte(r), rbyte(r), rbyte(r));
        a->Rectf(rrange(r, -1.2f, 1.2f), rrange(r, -1.2f, 1.2f), rrange(r, -1.2f, 1.2f), rrange(r, -1.2f, 1.2f));
    }
}

static void apply_fog(const PFSapi* a, const PFSparams* p)
{
    if (p->fog == 4) {
        a->Fogf(PF_FOG_DENSITY, 0.7f);                /* stores (int)0.7 = LINEAR in the mode */
        PFint m = 7;
        a->Fogiv(PF_FOG_MODE, &m);                    /* accepted: t = 0 for this mode */
    } else {
        a->Fogi(PF_FOG_MODE, p->fog - 1);
    }
    a->Fogf(PF_FOG_START, p->fogStart);
    a->Fogf(PF_FOG_END, p->fogEnd);
    PFfloat col[4] = { 0.8f, 0.7f, 0.9f, (float)p->fogAlpha / 255.0f };
    a->Fogfv(PF_FOG_COLOR, col);
    a->FogProcess();
}

static SceneTex make_texture(const PFSapi* a, Rng* r, const PFSparams* p)
{
    SceneTex st = { NULL, NULL };
    if (!p->texture) return st;
    uint32_t w = p->texW ? (uint32_t)p->texW : 256, h = p->texH ? (uint32_t)p->texH : 256;
    size_t n = (size_t)w * h;
    if (p->texture == 3) {
        /* Any (format, type) pair: random texel bytes, 64 zero bytes after
         * the last texel (the reference's getters read up to 3 past it). */
        int f = p->texFormat, t = p->texType;
        uint8_t* px = random_texels(r, f, t, n);
        if (p->texSwizzle && f == PF_BGRA && t == PF_UNSIGNED_BYTE) {
            for (size_t i = 0; i < n; i++) { uint8_t b = px[4 * i]; px[4 * i] = px[4 * i + 2]; px[4 * i + 2] = b; }
            f = PF_RGBA;
        }
        st.pixels = px;
        st.tex = a->GenTexture(px, w, h, (PFpixelformat)f, (PFdatatype)t);
    } else if (p->texture == 2) {
        float* px = (float*)malloc(n * 4 * sizeof(float) + 64);
        for (size_t i = 0; i < 4 * n; i++) px[i] = runit(r);
        st.pixels = px;
        st.tex = a->GenTexture(px, w, h, PF_RGBA, PF_FLOAT);
    } else {
        uint32_t* px = (uint32_t*)malloc(n * 4 + 64);
        for (size_t i = 0; i < n; i++) px[i] = (uint32_t)(rnext(r) >> 32);
        st.pixels = px;
        st.tex = a->GenTexture(px, w, h, PF_RGBA, PF_UNSIGNED_BYTE);
    }
    a->TextureParameter(st.tex, (PFtexturewrap)p->wrap, (PFtexturefilter)p->filter);
    return st;
}

static void free_texture(const PFSapi* a, SceneTex* st)
{
    if (st->tex) a->DeleteTexture(&st->tex, 0);
    free(st->pixels);
    st->pixels = NULL;
}

static void common_state(const PFSapi* a, const PFSparams* p)
{
    if (p->pointSize > 0) a->PointSize(p->pointSize);
    if (p->polyPoint & 1) a->PolygonMode(PF_FRONT, PF_POINT);
    if (p->polyPoint & 2) a->PolygonMode(PF_BACK, PF_POINT);
    if (p->lineWidth > 0) a->LineWidth(p->lineWidth);
    if (p->polyLine & 1) a->PolygonMode(PF_FRONT, PF_LINE);
    if (p->polyLine & 2) a->PolygonMode(PF_BACK, PF_LINE);
    if (p->cull < 0) a->Disable(PF_CULL_FACE);
    else { a->Enable(PF_CULL_FACE); a->CullFace((PFface)p->cull); }
    if (p->depthTest) { a->Enable(PF_DEPTH_TEST); a->DepthFunc((PFdepthmode)p->depthFunc); }
    a->ShadeModel((PFshademode)p->shade);
    if (p->blend) a->BlendFunc((PFblendmode)p->blendMode);
}

static void setup_lights(const PFSapi* a, Rng* r, const PFSparams* p, float cx, float cy, float cz, float spread)
{
    if (!p->lighting) return;
    a->Enable(PF_LIGHTING);
    a->LightModel(p->lighting == 2 ? PF_PHONG : PF_GOURAUD);
    int n = p->numLights > 0 ? p->numLights : 1;
    for (int i = 0; i < n && i < 8; i++) {
        float pos[3] = { cx + rrange(r, -spread, spread), cy + rrange(r, -spread, spread), cz + rrange(r, -spread, spread) };
        float dif[3] = { runit(r), runit(r), runit(r) };
        float spc[3] = { runit(r), runit(r), runit(r) };
        float amb[3] = { 0.2f * runit(r), 0.2f * runit(r), 0.2f * runit(r) };
        a->Lightfv((PFsizei)i, PF_POSITION, pos);
        a->Lightfv((PFsizei)i, PF_DIFFUSE, dif);
        a->Lightfv((PFsizei)i, PF_SPECULAR, spc);
        a->Lightfv((PFsizei)i, PF_AMBIENT, amb);
        if (p->spot) {
            float dir[3] = { cx - pos[0], cy - pos[1], cz - pos[2] };
            a->Lightfv((PFsizei)i, PF_SPOT_DIRECTION, dir);
            a->Lightf((PFsizei)i, PF_SPOT_INNER_CUTOFF, 20.0f);
            a->Lightf((PFsizei)i, PF_SPOT_OUTER_CUTOFF, 35.0f);
        }
        if (p->attenuation) {
            a->Lightf((PFsizei)i, PF_LINEAR_ATTENUATION, 0.05f + 0.2f * runit(r));
            a->Lightf((PFsizei)i, PF_QUADRATIC_ATTENUATION, 0.02f * runit(r));
        }
        a->EnableLight((PFsizei)i);
    }
    float mspec[3] = { 0.9f, 0.9f, 0.9f };
    a->Materialfv(PF_FRONT_AND_BACK, PF_SPECULAR, mspec);
    a->Materialf(PF_FRONT_AND_BACK, PF_SHININESS, 32.0f);
    if (p->colorMaterial) {
        a->ColorMaterial(PF_FRONT_AND_BACK, PF_AMBIENT_AND_DIFFUSE);
        a->Enable(PF_COLOR_MATERIAL);
    }
}

static void perspective(const PFSapi* a, const PFSparams* p, float fovScale, float n, float f)
{
    float aspect = (float)p->width / (float)p->height;
    a->MatrixMode(PF_PROJECTION);
    a->LoadIdentity();
    a->Frustum(-aspect * fovScale * n, aspect * fovScale * n, -fovScale * n, fovScale * n, n, f);
    a->MatrixMode(PF_MODELVIEW);
    a->LoadIdentity();
}

static void maybe_flush(const PFSapi* a, const PFSparams* p, uint32_t drawn)
{
    if (p->flushEvery > 0 && a->Finish && drawn % (uint32_t)p->flushEvery == 0) a->Finish();
}

/* One texture coordinate of a random-kind vertex (v is drawn before u, the
 * order in which the scenes have always consumed the generator). */
static float rand_uv(Rng* r, const PFSparams* p)
{
    return p->uvMax > p->uvMin ? rrange(r, p->uvMin, p->uvMax) : rrange(r, -2, 2);
}

/* Random triangles with vertices uniform in NDC [-1,1]^3 (2D path). */
static void scene_random_ndc(const PFSapi* a, Rng* r, const PFSparams* p)
{
    float s = p->triScale > 0 ? p->triScale : 1.0f;
    a->Begin(PF_TRIANGLES);
    for (uint32_t t = 0; t < p->ntris; t++) {
        float cx = rrange(r, -1, 1), cy = rrange(r, -1, 1), cz = rrange(r, -1, 1);
        for (int k = 0; k < 3; k++) {
            float x = rrange(r, -1, 1), y = rrange(r, -1, 1), z = rrange(r, -1, 1);
            if (s != 1.0f) { x = cx + s * x; y = cy + s * y; z = cz + s * z; }
            a->Color4ub(rbyte(r), rbyte(r), rbyte(r), rbyte(r));
            if (p->texture) { float v = rand_uv(r, p), u = rand_uv(r, p); a->TexCoord2f(u, v); }
            a->Vertex3f(x, y, z);
        }
        maybe_flush(a, p, t + 1);
    }
    a->End();
}

/* Random triangles inside and around a perspective frustum (3D path). */
static void scene_random_3d(const PFSapi* a, Rng* r, const PFSparams* p)
{
    perspective(a, p, 0.6f, 0.5f, 20.0f);
    float s = p->triScale > 0 ? p->triScale : 1.0f;
    a->Begin(PF_TRIANGLES);
    for (uint32_t t = 0; t < p->ntris; t++) {
        float cz = rrange(r, -12.0f, -0.2f);
        float cx = rrange(r, -1.2f, 1.2f) * -cz * 0.8f, cy = rrange(r, -1.2f, 1.2f) * -cz * 0.6f;
        for (int k = 0; k < 3; k++) {
            float x = cx + s * rrange(r, -2, 2), y = cy + s * rrange(r, -2, 2), z = cz + s * rrange(r, -2, 2);
            a->Color4ub(rbyte(r), rbyte(r), rbyte(r), rbyte(r));
            if (p->texture) { float v = rand_uv(r, p), u = rand_uv(r, p); a->TexCoord2f(u, v); }
            if (p->lighting) a->Normal3f(rrange(r, -1, 1), rrange(r, -1, 1), rrange(r, -1, 1));
            a->Vertex3f(x, y, z);
        }
        maybe_flush(a, p, t + 1);
    }
    a->End();
}

static float hf_height(float x, float z, const float* ph)
{
    return 0.35f * sinf(1.3f * x + ph[0]) * cosf(0.9f * z + ph[1])
         + 0.15f * sinf(3.1f * x + ph[2]) + 0.1f * cosf(2.7f * z + ph[3]);
}

/* Jittered height-field grid under a perspective camera. */
static void scene_heightfield(const PFSapi* a, Rng* r, const PFSparams* p)
{
    int NX = p->gridX > 0 ? p->gridX : 1000, NZ = p->gridZ > 0 ? p->gridZ : 100;
    perspective(a, p, 0.42f, 0.25f, 60.0f);
    a->Rotatef(28.0f, 1.0f, 0.0f, 0.0f);
    a->Translatef(0.0f, -2.2f, -1.0f);
    float ph[4] = { rrange(r, 0, 6.28f), rrange(r, 0, 6.28f), rrange(r, 0, 6.28f), rrange(r, 0, 6.28f) };
    size_t nv = (size_t)(NX + 1) * (NZ + 1);
    float* P = (float*)malloc(nv * 3 * sizeof(float));
    uint32_t* C = (uint32_t*)malloc(nv * sizeof(uint32_t));
    float SX = 7.0f, Z0 = 0.4f, SZ = 9.0f;
    for (int j = 0; j <= NZ; j++)
        for (int i = 0; i <= NX; i++) {
            size_t v = (size_t)j * (NX + 1) + i;
            float jx = (i > 0 && i < NX) ? rrange(r, -0.3f, 0.3f) : 0.0f;
            float jz = (j > 0 && j < NZ) ? rrange(r, -0.3f, 0.3f) : 0.0f;
            float x = ((float)i + jx) / NX * SX - 0.5f * SX;
            float z = -(Z0 + ((float)j + jz) / NZ * SZ);
            P[3 * v] = x;
            P[3 * v + 1] = hf_height(x, z, ph);
            P[3 * v + 2] = z;
            C[v] = (uint32_t)(rnext(r) >> 32);
        }
    size_t ntri = (size_t)NX * NZ * 2;
    uint8_t* blended = (uint8_t*)malloc(ntri);
    for (size_t t = 0; t < ntri; t++) blended[t] = runit(r) < p->blendFraction;
    float ru = 4.0f, rv = 2.0f;      /* texture repeats across the grid: about one texel per pixel up front */
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            if (p->blendFraction <= 0.0f) break;
            a->Enable(PF_BLEND);
            a->BlendFunc((PFblendmode)(p->blend ? p->blendMode : PF_BLEND_ALPHA));
        }
        a->Begin(PF_TRIANGLES);
        uint32_t drawn = 0;
        for (int j = 0; j < NZ; j++)
            for (int i = 0; i < NX; i++)
                for (int h = 0; h < 2; h++) {
                    size_t t = ((size_t)j * NX + i) * 2 + h;
                    if (blended[t] != pass) continue;
                    int ci[3][2];
                    if (h == 0) { ci[0][0] = i; ci[0][1] = j; ci[1][0] = i + 1; ci[1][1] = j; ci[2][0] = i; ci[2][1] = j + 1; }
                    else        { ci[0][0] = i + 1; ci[0][1] = j; ci[1][0] = i + 1; ci[1][1] = j + 1; ci[2][0] = i; ci[2][1] = j + 1; }
                    for (int k = 0; k < 3; k++) {
                        size_t v = (size_t)ci[k][1] * (NX + 1) + ci[k][0];
                        uint32_t c = C[v];
                        a->Color4ub((PFubyte)c, (PFubyte)(c >> 8), (PFubyte)(c >> 16), pass ? 128 : 255);
                        a->TexCoord2f((float)ci[k][0] / NX * ru, (float)ci[k][1] / NZ * rv);
                        a->Vertex3f(P[3 * v], P[3 * v + 1], P[3 * v + 2]);
                    }
                    maybe_flush(a, p, ++drawn);
                }
        a->End();
    }
    free(P); free(C); free(blended);
}

/* Indexed grid mesh (gridX x gridZ cells, two triangles each) spanning the
 * view of a perspective camera: a gently bumped sheet at z ~ -2.4 with
 * analytic normals, texture coordinates [0,1]^2 across the sheet and random
 * float colours in [0.2, 1).  The same vertices go through
 * pfDrawElements (arrays = 1), pfDrawArrays over the expanded corners,
 * offset by `first` (arrays = 2), or pfColor4f / pfTexCoord2f /
 * pfNormal3f / pfVertex3f (arrays = 0).  Values are chosen so that every
 * array type represents them exactly. */
static void store_f(void* arr, int type, size_t i, double v)
{
    if (type == PF_DOUBLE) ((double*)arr)[i] = v; else ((float*)arr)[i] = (float)v;
}

static void store_col(void* arr, int type, size_t i, float v)
{
    switch (type) {
    case PF_UNSIGNED_BYTE:  ((uint8_t*)arr)[i] = (uint8_t)(v * 255.0f); break;
    case PF_UNSIGNED_SHORT: ((uint16_t*)arr)[i] = (uint16_t)((uint32_t)(v * 255.0f) << 8); break;
    case PF_UNSIGNED_INT:   ((uint32_t*)arr)[i] = (uint32_t)(v * 255.0f) << 24; break;
    case PF_DOUBLE:         ((double*)arr)[i] = (double)v; break;
    default:                ((float*)arr)[i] = v; break;
    }
}

static size_t type_size(int type)
{
    switch (type) {
    case PF_UNSIGNED_BYTE: return 1;
    case PF_UNSIGNED_SHORT: return 2;
    case PF_DOUBLE: return 8;
    default: return 4;
    }
}

static void scene_mesh(const PFSapi* a, Rng* r, const PFSparams* p)
{
    int NX = p->gridX > 0 ? p->gridX : 100, NZ = p->gridZ > 0 ? p->gridZ : 100;
    perspective(a, p, 0.5f, 0.5f, 10.0f);
    a->Rotatef(-12.0f, 1.0f, 0.0f, 0.0f);
    float aspect = (float)p->width / (float)p->height;
    size_t nv = (size_t)(NX + 1) * (NZ + 1);
    static const int colTypes[5] = { PF_FLOAT, PF_UNSIGNED_BYTE, PF_UNSIGNED_SHORT, PF_UNSIGNED_INT, PF_DOUBLE };
    static const int idxTypes[3] = { PF_UNSIGNED_INT, PF_UNSIGNED_BYTE, PF_UNSIGNED_SHORT };
    int posType = p->posType == 1 ? PF_DOUBLE : PF_FLOAT;
    int colType = colTypes[(p->colType >= 0 && p->colType < 5) ? p->colType : 0];
    int colSize = p->colSize ? p->colSize : 4;
    float* P = (float*)malloc(nv * 3 * sizeof(float));
    float* N = (float*)malloc(nv * 3 * sizeof(float));
    float* T = (float*)malloc(nv * 2 * sizeof(float));
    float* Cf = (float*)malloc(nv * 4 * sizeof(float));
    float ph0 = rrange(r, 0, 6.28f), ph1 = rrange(r, 0, 6.28f);
    for (int j = 0; j <= NZ; j++)
        for (int i = 0; i <= NX; i++) {
            size_t v = (size_t)j * (NX + 1) + i;
            float x = (2.0f * i / NX - 1.0f) * 1.25f * aspect, y = (2.0f * j / NZ - 1.0f) * 1.25f;
            float h = 0.12f * sinf(2.1f * x + ph0) * cosf(1.7f * y + ph1);
            float dx = 0.12f * 2.1f * cosf(2.1f * x + ph0) * cosf(1.7f * y + ph1);
            float dy = -0.12f * 1.7f * sinf(2.1f * x + ph0) * sinf(1.7f * y + ph1);
            P[3 * v] = x; P[3 * v + 1] = y; P[3 * v + 2] = -2.4f + h;
            float nl = sqrtf(dx * dx + dy * dy + 1.0f);
            N[3 * v] = -dx / nl; N[3 * v + 1] = -dy / nl; N[3 * v + 2] = 1.0f / nl;
            T[2 * v] = (float)i / NX; T[2 * v + 1] = (float)j / NZ;
            for (int k = 0; k < 4; k++) {
                float c = (k == 3) ? 1.0f : 0.2f + 0.8f * runit(r);
                /* Multiples of 1/256: c*255 is exact in float and double, so
                 * every colour array type truncates to the same byte. */
                Cf[4 * v + k] = (float)((int)(c * 256.0f)) / 256.0f;
            }
        }
    const int rowMode = p->arrays && p->arrayMode >= PF_TRIANGLE_FAN && p->arrayMode <= PF_QUAD_STRIP;
    const size_t perRow = !rowMode ? (size_t)NX * 6 : p->arrayMode == PF_QUADS ? (size_t)NX * 4 : (size_t)(NX + 1) * 2;
    size_t ni = perRow * NZ;
    uint32_t* I = (uint32_t*)malloc(ni * sizeof(uint32_t));
    size_t q = 0;
    for (int j = 0; j < NZ; j++) {
        if (rowMode && p->arrayMode != PF_QUADS) {
            for (int i = 0; i <= NX; i++) {
                I[q++] = (uint32_t)(j * (NX + 1) + i);
                I[q++] = (uint32_t)((j + 1) * (NX + 1) + i);
            }
            continue;
        }
        for (int i = 0; i < NX; i++) {
            uint32_t v00 = (uint32_t)(j * (NX + 1) + i), v10 = v00 + 1;
            uint32_t v01 = v00 + (uint32_t)(NX + 1), v11 = v01 + 1;
            if (rowMode) { I[q++] = v00; I[q++] = v10; I[q++] = v11; I[q++] = v01; continue; }
            I[q++] = v00; I[q++] = v10; I[q++] = v11;
            I[q++] = v00; I[q++] = v11; I[q++] = v01;
        }
    }
    const PFdrawmode dm = rowMode ? (PFdrawmode)p->arrayMode : PF_TRIANGLES;
    const size_t calls = rowMode ? (size_t)NZ : 1, perCall = rowMode ? perRow : ni;
    if (p->arrays == 0) {
        a->Begin(PF_TRIANGLES);
        for (size_t k = 0; k < ni; k++) {
            uint32_t v = I[k];
            const float* c = &Cf[4 * v];
            a->Color4f(c[0], c[1], c[2], colSize == 4 ? c[3] : 1.0f);
            a->TexCoord2f(T[2 * v], T[2 * v + 1]);
            a->Normal3f(N[3 * v], N[3 * v + 1], N[3 * v + 2]);
            a->Vertex3f(P[3 * v], P[3 * v + 1], P[3 * v + 2]);
            maybe_flush(a, p, (uint32_t)(k + 1) / 3);
        }
        a->End();
    } else {
        int expand = p->arrays == 2;
        size_t first = expand ? 5 : 0;                  /* DrawArrays from a non-zero first */
        size_t n = expand ? first + ni : nv;
        void* pa = calloc(n * 3, type_size(posType));
        void* na = calloc(n * 3, sizeof(float));
        void* ta = calloc(n * 2, sizeof(float));
        void* ca = calloc(n * (size_t)colSize, type_size(colType));
        for (size_t e = 0; e < n; e++) {
            size_t v = expand ? (e >= first ? I[e - first] : 0) : e;
            for (int k = 0; k < 3; k++) store_f(pa, posType, 3 * e + k, P[3 * v + k]);
            for (int k = 0; k < 3; k++) ((float*)na)[3 * e + k] = N[3 * v + k];
            for (int k = 0; k < 2; k++) ((float*)ta)[2 * e + k] = T[2 * v + k];
            for (int k = 0; k < colSize; k++) store_col(ca, colType, (size_t)colSize * e + k, Cf[4 * v + k]);
        }
        a->Enable(PF_VERTEX_ARRAY | PF_NORMAL_ARRAY | PF_TEXTURE_COORD_ARRAY | PF_COLOR_ARRAY);
        a->VertexPointer(3, (PFenum)posType, 0, pa);
        a->NormalPointer(PF_FLOAT, 0, na);
        a->TexCoordPointer(PF_FLOAT, 0, ta);
        a->ColorPointer(colSize, (PFenum)colType, 0, ca);
        if (expand) {
            for (size_t k = 0; k < calls; k++)
                a->DrawArrays(dm, (PFint)(first + k * perCall), (PFsizei)perCall);
        } else {
            int it = idxTypes[(p->indexType >= 0 && p->indexType < 3) ? p->indexType : 0];
            void* ix = malloc(ni * type_size(it));
            for (size_t k = 0; k < ni; k++) {
                if (it == PF_UNSIGNED_BYTE) ((uint8_t*)ix)[k] = (uint8_t)I[k];
                else if (it == PF_UNSIGNED_SHORT) ((uint16_t*)ix)[k] = (uint16_t)I[k];
                else ((uint32_t*)ix)[k] = I[k];
            }
            for (size_t k = 0; k < calls; k++)
                a->DrawElements(dm, (PFsizei)perCall, (PFdatatype)it, (const uint8_t*)ix + k * perCall * type_size(it));
            free(ix);
        }
        /* The reference reads the arrays inside the draw call; this library
         * copies them into the recorded frame, so both may be freed here. */
        a->Disable(PF_VERTEX_ARRAY | PF_NORMAL_ARRAY | PF_TEXTURE_COORD_ARRAY | PF_COLOR_ARRAY);
        free(pa); free(na); free(ta); free(ca);
    }
    free(P); free(N); free(T); free(Cf); free(I);
}

/* Cluster of tessellated spheres with per-vertex normals. */
static void scene_spheres(const PFSapi* a, Rng* r, const PFSparams* p)
{
    int NS = p->spheres > 0 ? p->spheres : 256;
    int NLAT = p->sphereLat > 0 ? p->sphereLat : 31, NLON = p->sphereLon > 0 ? p->sphereLon : 63;
    perspective(a, p, 0.5f, 0.5f, 30.0f);
    a->Translatef(0.0f, 0.0f, -4.5f);
    setup_lights(a, r, p, 0.0f, 0.0f, 0.0f, 2.5f);
    a->Begin(PF_TRIANGLES);
    uint32_t drawn = 0;
    for (int s = 0; s < NS; s++) {
        float cx = rrange(r, -1.6f, 1.6f), cy = rrange(r, -1.0f, 1.0f), cz = rrange(r, -1.2f, 1.2f);
        float rad = rrange(r, 0.05f, 0.3f);
        uint32_t col = (uint32_t)(rnext(r) >> 32) | 0xFF000000u;
        for (int la = 0; la < NLAT; la++)
            for (int lo = 0; lo < NLON; lo++)
                for (int h = 0; h < 2; h++) {
                    int cc[3][2];
                    if (h == 0) { cc[0][0] = la; cc[0][1] = lo; cc[1][0] = la + 1; cc[1][1] = lo; cc[2][0] = la + 1; cc[2][1] = lo + 1; }
                    else        { cc[0][0] = la; cc[0][1] = lo; cc[1][0] = la + 1; cc[1][1] = lo + 1; cc[2][0] = la; cc[2][1] = lo + 1; }
                    for (int k = 0; k < 3; k++) {
                        float th = (float)cc[k][0] / NLAT * 3.14159265f;
                        float phi = (float)cc[k][1] / NLON * 6.28318531f;
                        float nx = sinf(th) * cosf(phi), ny = cosf(th), nz = sinf(th) * sinf(phi);
                        a->Color4ub((PFubyte)col, (PFubyte)(col >> 8), (PFubyte)(col >> 16), 255);
                        a->Normal3f(nx, ny, nz);
                        if (p->texture) a->TexCoord2f((float)cc[k][1] / NLON, (float)cc[k][0] / NLAT);
                        a->Vertex3f(cx + rad * nx, cy + rad * ny, cz + rad * nz);
                    }
                    maybe_flush(a, p, ++drawn);
                }
    }
    a->End();
}

/* Render lists.  List A holds a unit sphere recorded in immediate mode: its
 * upper half as PF_TRIANGLES with the scene texture bound, its lower half as
 * PF_QUADS with no texture.  While A is recorded the colour, normal, texture
 * coordinate, bound texture and PF_BLEND change; pfEndList restores them.
 * List B is recorded as a colour change followed by pfCallList(A), whose
 * context backup replaces the one pfNewList(B) made, so pfEndList(B) keeps
 * that colour.  The spheres are then A or B called at random model-view
 * transforms, and one immediate-mode triangle in the current colour ends
 * the scene. */
static void scene_list(const PFSapi* a, Rng* r, const PFSparams* p, PFtexture tex)
{
    int NS = p->spheres > 0 ? p->spheres : 8;
    int NLAT = p->sphereLat > 0 ? p->sphereLat : 10, NLON = p->sphereLon > 0 ? p->sphereLon : 20;
    perspective(a, p, 0.5f, 0.5f, 30.0f);
    setup_lights(a, r, p, 0.0f, 0.0f, -7.0f, 3.0f);
    size_t nv = (size_t)(NLAT + 1) * (NLON + 1);
    uint32_t* col = (uint32_t*)malloc(nv * sizeof(uint32_t));
    for (size_t v = 0; v < nv; v++) col[v] = (uint32_t)(rnext(r) >> 32) | 0xFF000000u;
    PFrenderlist A = a->GenList(), B = a->GenList();
    a->Color4ub(200, 100, 50, 255);
    a->NewList(A);
    for (int half = 0; half < 2; half++) {
        a->BindTexture(half == 0 ? tex : NULL);
        if (half == 1) a->Enable(PF_BLEND);
        a->Begin(half == 0 ? PF_TRIANGLES : PF_QUADS);
        for (int la = half ? NLAT / 2 : 0; la < (half ? NLAT : NLAT / 2); la++)
            for (int lo = 0; lo < NLON; lo++) {
                int cc[6][2], n;
                if (half == 0) {
                    int t[6][2] = { { la, lo }, { la + 1, lo }, { la + 1, lo + 1 },
                                    { la, lo }, { la + 1, lo + 1 }, { la, lo + 1 } };
                    memcpy(cc, t, sizeof t); n = 6;
                } else {
                    int q[4][2] = { { la, lo }, { la + 1, lo }, { la + 1, lo + 1 }, { la, lo + 1 } };
                    memcpy(cc, q, sizeof q); n = 4;
                }
                for (int k = 0; k < n; k++) {
                    float th = (float)cc[k][0] / NLAT * 3.14159265f;
                    float phi = (float)cc[k][1] / NLON * 6.28318531f;
                    float nx = sinf(th) * cosf(phi), ny = cosf(th), nz = sinf(th) * sinf(phi);
                    uint32_t c = col[(size_t)cc[k][0] * (NLON + 1) + cc[k][1]];
                    a->Color4ub((PFubyte)c, (PFubyte)(c >> 8), (PFubyte)(c >> 16), 255);
                    a->Normal3f(nx, ny, nz);
                    a->TexCoord2f((float)cc[k][1] / NLON, (float)cc[k][0] / NLAT);
                    a->Vertex3f(nx, ny, nz);
                }
            }
        a->End();
    }
    a->Color4ub(0, 0, 0, 255);
    a->EndList();
    a->NewList(B);
    a->Color4ub(40, 220, 90, 255);
    a->CallList(A);
    a->EndList();
    if (p->listOptions & 1) {
        a->MatrixMode(PF_TEXTURE);
        a->Rotatef(30.0f, 0.0f, 0.0f, 1.0f);
        a->Scalef(1.7f, 0.6f, 1.0f);
        a->Translatef(0.25f, -0.4f, 0.0f);
        a->MatrixMode(PF_MODELVIEW);
    }
    if (p->listOptions & 2) a->Enable(PF_NORMALIZE);
    for (int s = 0; s < NS; s++) {
        a->LoadIdentity();
        a->Translatef(rrange(r, -2.5f, 2.5f), rrange(r, -1.6f, 1.6f), rrange(r, -9.0f, -5.0f));
        a->Rotatef(rrange(r, 0.0f, 360.0f), rrange(r, -1.0f, 1.0f), rrange(r, -1.0f, 1.0f), rrange(r, 0.1f, 1.0f));
        a->CallList(s % 3 == 2 ? B : A);
    }
    a->LoadIdentity();
    a->Begin(PF_TRIANGLES);
    a->Vertex3f(-1.5f, -1.2f, -6.0f);
    a->Vertex3f(0.5f, -1.4f, -6.5f);
    a->Vertex3f(-0.4f, 0.2f, -5.5f);
    a->End();
    if (p->listOptions & 4) {                       /* re-recorded: the old calls must survive */
        a->NewList(A);
        a->EndList();
    }
    a->DeleteList(&A);
    a->DeleteList(&B);
    free(col);
}

/* ------------------------------------------------------------------------ */
/* Entry points                                                              */
/* ------------------------------------------------------------------------ */

static void draw_scene(const PFSapi* a, Rng* r, const PFSparams* p, SceneTex* tex)
{
    common_state(a, p);
    *tex = make_texture(a, r, p);
    if (tex->tex) { a->BindTexture(tex->tex); a->Enable(PF_TEXTURE_2D); }
    if (p->blend && p->kind != PFS_HEIGHTFIELD) a->Enable(PF_BLEND);
    if (p->clearFirst) { a->ClearColor(20, 30, 40, 255); a->Clear(PF_COLOR_BUFFER_BIT | PF_DEPTH_BUFFER_BIT); }
    if ((p->pixelOps > 0 || p->fog || p->points > 0 || p->lines > 0 || p->rects > 0)
        && (p->kind == PFS_RANDOM_NDC || p->kind == PFS_RANDOM_3D)) {
        PFSparams half = *p;
        half.ntris = p->ntris / 2;
        if (p->kind == PFS_RANDOM_3D) setup_lights(a, r, p, 0.0f, 0.0f, -4.0f, 3.0f);
        if (p->kind == PFS_RANDOM_NDC) scene_random_ndc(a, r, &half); else scene_random_3d(a, r, &half);
        draw_pixel_ops(a, r, p);
        if (p->points > 0) draw_points(a, r, p);
        if (p->lines > 0) draw_lines(a, r, p);
        if (p->rects > 0) draw_rects(a, r, p);
        if (p->fog) apply_fog(a, p);
        half.ntris = p->ntris - half.ntris;
        if (p->kind == PFS_RANDOM_NDC) scene_random_ndc(a, r, &half); else scene_random_3d(a, r, &half);
        return;
    }
    switch (p->kind) {
    case PFS_RANDOM_NDC: scene_random_ndc(a, r, p); break;
    case PFS_HEIGHTFIELD: scene_heightfield(a, r, p); break;
    case PFS_SPHERES: scene_spheres(a, r, p); break;
    case PFS_RANDOM_3D:
        setup_lights(a, r, p, 0.0f, 0.0f, -4.0f, 3.0f);
        scene_random_3d(a, r, p);
        break;
    case PFS_MESH:
        setup_lights(a, r, p, 0.0f, 0.0f, -1.2f, 1.0f);
        scene_mesh(a, r, p);
        break;
    case PFS_LIST: scene_list(a, r, p, tex->tex); break;
    default: break;
    }
    if (p->fog) apply_fog(a, p);
}

/* Bytes per pixel of the render-target pairs scenes use (RGBA8 for the
 * others). */
static size_t target_bytes(PFpixelformat f, PFdatatype t)
{
    const size_t comps = f == PF_LUMINANCE_ALPHA ? 2 : (f == PF_RGB || f == PF_BGR) ? 3 : 4;
    if (t == PF_HALF_FLOAT) return 2 * comps;
    if (t == PF_FLOAT) return 4 * comps;
    if (t != PF_UNSIGNED_BYTE) return 2;
    return comps;
}

/* Renders one scene into an offscreen framebuffer of the bound library (or,
 * with fbMain, into the context's own buffer) and copies its colour (W*H
 * pixels of the target's format) and depth into the caller's arrays.  The
 * context's own buffer holds W*H+16 pixels.  Returns the library's last
 * error code, or -1 on setup failure. */
int pfs_render(const PFSapi* a, const PFSparams* p, void* colorOut, float* depthOut)
{
    if (!a || !p || !p->width || !p->height) return -1;
    size_t n = (size_t)p->width * p->height;
    const PFpixelformat ff = p->fbFormat > 0 ? (PFpixelformat)p->fbFormat : PF_RGBA;
    const PFdatatype ft = p->fbFormat > 0 ? (PFdatatype)p->fbType : PF_UNSIGNED_BYTE;
    const size_t bpp = target_bytes(ff, ft);
    void* mainColor = calloc(n + 16, p->fbMain ? bpp : 4);
    if (!mainColor) return -1;
    PFcontext ctx = p->fbMain ? a->CreateContext(mainColor, p->width, p->height, ff, ft)
                              : a->CreateContext(mainColor, p->width, p->height, PF_RGBA, PF_UNSIGNED_BYTE);
    if (!ctx) { free(mainColor); return -1; }
    a->MakeCurrent(ctx);
    PFframebuffer fb = { 0 };
    if (!p->fbMain) {
        fb = a->GenFramebuffer(p->width, p->height, ff, ft);
        if (!fb.texture) {
            int err = (int)a->GetError();
            a->DeleteContext(ctx);
            a->MakeCurrent(NULL);
            free(mainColor);
            return err ? err : -1;
        }
        a->BindFramebuffer(&fb);
        a->Enable(PF_FRAMEBUFFER);
    }
    Rng r = { p->seed };
    SceneTex tex;
    draw_scene(a, &r, p, &tex);
    if (a->Finish) a->Finish();
    int err = (int)a->GetError();
    const void* px = p->fbMain ? mainColor : a->GetTexturePixels(fb.texture, NULL, NULL, NULL, NULL);
    if (colorOut) memcpy(colorOut, px, n * bpp);
    if (depthOut && !p->fbMain) memcpy(depthOut, fb.zbuffer, n * 4);
    if (!p->fbMain) {
        a->Disable(PF_FRAMEBUFFER);
        a->BindFramebuffer(NULL);
        a->DeleteFramebuffer(&fb);
    }
    free_texture(a, &tex);
    a->DeleteContext(ctx);
    a->MakeCurrent(NULL);
    free(mainColor);
    return err;
}

/* Renders `count` scenes, each into its own context and main buffer, and
 * submits them together with pfSubmitBatch when the library has it (one
 * after another otherwise).  colors[i] must hold W_i*H_i+16 pixels. */
int pfs_render_batch(const PFSapi* a, const PFSparams* ps, uint32_t count, uint32_t** colors)
{
    PFcontext* ctxs = (PFcontext*)calloc(count, sizeof(PFcontext));
    SceneTex* texs = (SceneTex*)calloc(count, sizeof(SceneTex));
    if (!ctxs || !texs) { free(ctxs); free(texs); return -1; }
    int err = 0;
    for (uint32_t i = 0; i < count; i++) {
        const PFSparams* p = &ps[i];
        ctxs[i] = a->CreateContext(colors[i], p->width, p->height, PF_RGBA, PF_UNSIGNED_BYTE);
        if (!ctxs[i]) { err = -1; break; }
        a->MakeCurrent(ctxs[i]);
        Rng r = { p->seed };
        draw_scene(a, &r, p, &texs[i]);
        if (!a->SubmitBatch) { int e = (int)a->GetError(); if (e) err = e; }
    }
    if (err == 0 && a->SubmitBatch) {
        a->SubmitBatch(ctxs, count);
        for (uint32_t i = 0; i < count; i++) {
            a->MakeCurrent(ctxs[i]);
            int e = (int)a->GetError();
            if (e) err = e;
        }
    }
    for (uint32_t i = 0; i < count; i++) {
        if (!ctxs[i]) continue;
        a->MakeCurrent(ctxs[i]);
        free_texture(a, &texs[i]);
        a->DeleteContext(ctxs[i]);
    }
    a->MakeCurrent(NULL);
    free(ctxs);
    free(texs);
    return err;
}

/* Records a scene into the current context of the bound library without
 * creating or deleting anything; the caller owns context and target.  The
 * texture is returned through *texOut / *texPixels for the caller to free
 * with pfs_free_texture after the frame has run. */
int pfs_record(const PFSapi* a, const PFSparams* p, PFtexture* texOut, void** texPixels)
{
    Rng r = { p->seed };
    SceneTex tex;
    draw_scene(a, &r, p, &tex);
    if (texOut) *texOut = tex.tex;
    if (texPixels) *texPixels = tex.pixels;
    return (int)a->GetError();
}

void pfs_free_texture(const PFSapi* a, PFtexture tex, void* pixels)
{
    if (tex) a->DeleteTexture(&tex, 0);
    free(pixels);
}

/* ------------------------------------------------------------------------ */
/* Call tapes                                                                */
/* ------------------------------------------------------------------------ */

/* A tape is the sequence of API calls one scene makes, with their
 * arguments, generated once by running the scene code against a recorder.
 * Replaying it through a library issues exactly those calls, so timing a
 * replay measures the library (its API calls and pfFinish) and not the
 * scene generator (random numbers, sinf).  Textures the scene creates are
 * kept as slots on the tape and created in the library once, before the
 * timed frames.  Tapes cover the scene kinds whose calls the recorder
 * knows: random NDC and 3D triangles, height fields and spheres, with no
 * pixel operations, points, lines, rects, fog, flushes or vertex arrays. */

enum {
    T_ENABLE = 1, T_DISABLE, T_CULLFACE, T_DEPTHFUNC, T_SHADEMODEL, T_BLENDFUNC, T_MATRIXMODE, T_CLEAR,
    T_BEGIN, T_LIGHTMODEL, T_ENABLELIGHT, T_LOADIDENTITY, T_END, T_TRANSLATE, T_VERTEX3, T_NORMAL3,
    T_ROTATE, T_FRUSTUM, T_TEXCOORD2, T_COLOR4UB, T_CLEARCOLOR, T_LIGHTFV, T_LIGHTF, T_MATERIALFV,
    T_MATERIALF, T_COLORMATERIAL, T_BINDTEXTURE, T_TEXPARAM, T_DISABLELIGHT
};

#define TAPE_MAX_TEX 4

typedef struct {
    uint32_t* w;                 /* opcodes and arguments, floats as bits */
    size_t    n, cap;
    int       bad;               /* the scene made a call the recorder does not know */
    uint32_t  ntex;
    void*     texPixels[TAPE_MAX_TEX];
    uint32_t  texW[TAPE_MAX_TEX], texH[TAPE_MAX_TEX];
    int32_t   texFormat[TAPE_MAX_TEX], texType[TAPE_MAX_TEX];
    PFtexture bound[TAPE_MAX_TEX];   /* the library's textures while the tape is bound */
    size_t    body;              /* words of the scene's calls; an epilogue follows */
    uint32_t  enabled;           /* PFstate bits the scene enables */
    uint32_t  lights;            /* lights the scene enables, by bit */
    uint32_t  width, height;
    uint64_t  calls;             /* API calls per replay */
} PFStape;

static PFStape* g_tape;          /* the tape being recorded */

static void tput(uint32_t v)
{
    PFStape* t = g_tape;
    if (t->bad) return;
    if (t->n == t->cap) {
        size_t nc = t->cap ? t->cap * 2 : 4096;
        uint32_t* w = (uint32_t*)realloc(t->w, nc * sizeof(uint32_t));
        if (!w) { t->bad = 1; return; }
        t->w = w;
        t->cap = nc;
    }
    t->w[t->n++] = v;
}

static void tputf(float f) { uint32_t v; memcpy(&v, &f, 4); tput(v); }
static void top(uint32_t op) { tput(op); g_tape->calls++; }

static uint32_t tex_slot(PFtexture tex) { return (uint32_t)(uintptr_t)tex; }   /* slot + 1, 0: none */

static void r_Enable(PFstate s) { top(T_ENABLE); tput((uint32_t)s); g_tape->enabled |= (uint32_t)s; }
static void r_Disable(PFstate s) { top(T_DISABLE); tput((uint32_t)s); }
static void r_CullFace(PFface f) { top(T_CULLFACE); tput((uint32_t)f); }
static void r_DepthFunc(PFdepthmode m) { top(T_DEPTHFUNC); tput((uint32_t)m); }
static void r_ShadeModel(PFshademode m) { top(T_SHADEMODEL); tput((uint32_t)m); }
static void r_BlendFunc(PFblendmode m) { top(T_BLENDFUNC); tput((uint32_t)m); }
static void r_MatrixMode(PFmatrixmode m) { top(T_MATRIXMODE); tput((uint32_t)m); }
static void r_Clear(PFclearflag f) { top(T_CLEAR); tput((uint32_t)f); }
static void r_Begin(PFdrawmode m) { top(T_BEGIN); tput((uint32_t)m); }
static void r_LightModel(PFlightmode m) { top(T_LIGHTMODEL); tput((uint32_t)m); }
static void r_EnableLight(PFsizei l) { top(T_ENABLELIGHT); tput((uint32_t)l); if (l < 32) g_tape->lights |= 1u << l; }
static void r_LoadIdentity(void) { top(T_LOADIDENTITY); }
static void r_End(void) { top(T_END); }
static void r_Translatef(PFfloat x, PFfloat y, PFfloat z) { top(T_TRANSLATE); tputf(x); tputf(y); tputf(z); }
static void r_Vertex3f(PFfloat x, PFfloat y, PFfloat z) { top(T_VERTEX3); tputf(x); tputf(y); tputf(z); }
static void r_Normal3f(PFfloat x, PFfloat y, PFfloat z) { top(T_NORMAL3); tputf(x); tputf(y); tputf(z); }
static void r_Rotatef(PFfloat a, PFfloat x, PFfloat y, PFfloat z) { top(T_ROTATE); tputf(a); tputf(x); tputf(y); tputf(z); }
static void r_Frustum(PFfloat l, PFfloat r, PFfloat b, PFfloat t, PFfloat n, PFfloat f)
{
    top(T_FRUSTUM); tputf(l); tputf(r); tputf(b); tputf(t); tputf(n); tputf(f);
}
static void r_TexCoord2f(PFfloat u, PFfloat v) { top(T_TEXCOORD2); tputf(u); tputf(v); }
static void r_Color4ub(PFubyte r, PFubyte g, PFubyte b, PFubyte a)
{
    top(T_COLOR4UB); tput((uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24));
}
static void r_ClearColor(PFubyte r, PFubyte g, PFubyte b, PFubyte a)
{
    top(T_CLEARCOLOR); tput((uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24));
}
static void r_Lightfv(PFsizei l, PFenum p, const void* v)     /* three floats in every scene call */
{
    const float* f = (const float*)v;
    top(T_LIGHTFV); tput((uint32_t)l); tput((uint32_t)p); tputf(f[0]); tputf(f[1]); tputf(f[2]);
}
static void r_Lightf(PFsizei l, PFenum p, PFfloat v) { top(T_LIGHTF); tput((uint32_t)l); tput((uint32_t)p); tputf(v); }
static void r_Materialfv(PFface face, PFenum p, const void* v)
{
    const float* f = (const float*)v;
    top(T_MATERIALFV); tput((uint32_t)face); tput((uint32_t)p); tputf(f[0]); tputf(f[1]); tputf(f[2]);
}
static void r_Materialf(PFface face, PFenum p, PFfloat v) { top(T_MATERIALF); tput((uint32_t)face); tput((uint32_t)p); tputf(v); }
static void r_ColorMaterial(PFface face, PFenum m) { top(T_COLORMATERIAL); tput((uint32_t)face); tput((uint32_t)m); }
static void r_BindTexture(PFtexture t) { top(T_BINDTEXTURE); tput(tex_slot(t)); }
static void r_TextureParameter(PFtexture t, PFtexturewrap w, PFtexturefilter f)
{
    top(T_TEXPARAM); tput(tex_slot(t)); tput((uint32_t)w); tput((uint32_t)f);
}
static PFtexture r_GenTexture(void* px, PFsizei w, PFsizei h, PFpixelformat f, PFdatatype t)
{
    PFStape* tp = g_tape;
    if (tp->ntex == TAPE_MAX_TEX) { tp->bad = 1; return NULL; }
    uint32_t k = tp->ntex++;
    tp->texPixels[k] = px;
    tp->texW[k] = w; tp->texH[k] = h;
    tp->texFormat[k] = f; tp->texType[k] = t;
    return (PFtexture)(uintptr_t)(k + 1);
}
static PFerrcode r_GetError(void) { return PF_NO_ERROR; }

/* Records scene `p` into a new tape; NULL for a scene kind or option the
 * recorder does not cover. */
PFStape* pfs_tape_make(const PFSparams* p)
{
    if (!p || !(p->kind == PFS_RANDOM_NDC || p->kind == PFS_RANDOM_3D || p->kind == PFS_HEIGHTFIELD
                || p->kind == PFS_SPHERES))
        return NULL;
    if (p->pixelOps > 0 || p->fog || p->points > 0 || p->lines > 0 || p->rects > 0 || p->flushEvery > 0
        || p->arrays || p->pointSize > 0 || p->lineWidth > 0 || p->polyPoint || p->polyLine)
        return NULL;
    PFStape* t = (PFStape*)calloc(1, sizeof(PFStape));
    if (!t) return NULL;
    PFSapi r;
    memset(&r, 0, sizeof r);
    r.Enable = r_Enable; r.Disable = r_Disable; r.CullFace = r_CullFace; r.DepthFunc = r_DepthFunc;
    r.ShadeModel = r_ShadeModel; r.BlendFunc = r_BlendFunc; r.MatrixMode = r_MatrixMode; r.Clear = r_Clear;
    r.Begin = r_Begin; r.LightModel = r_LightModel; r.EnableLight = r_EnableLight; r.LoadIdentity = r_LoadIdentity;
    r.End = r_End; r.Translatef = r_Translatef; r.Vertex3f = r_Vertex3f; r.Normal3f = r_Normal3f;
    r.Rotatef = r_Rotatef; r.Frustum = r_Frustum; r.TexCoord2f = r_TexCoord2f; r.Color4ub = r_Color4ub;
    r.ClearColor = r_ClearColor; r.Lightfv = r_Lightfv; r.Lightf = r_Lightf; r.Materialfv = r_Materialfv;
    r.Materialf = r_Materialf; r.ColorMaterial = r_ColorMaterial; r.BindTexture = r_BindTexture;
    r.TextureParameter = r_TextureParameter; r.GenTexture = r_GenTexture; r.GetError = r_GetError;
    g_tape = t;
    Rng rng = { p->seed };
    SceneTex tex;
    draw_scene(&r, &rng, p, &tex);
    /* Epilogue, played after each timed frame: disables every state and
     * light the scene enabled, so that the next frame starts from the state
     * the first one started from (the scene sets the rest itself). */
    t->body = t->n;
    for (uint32_t b = 0; b < 32; b++)
        if (t->enabled & (1u << b)) { tput(T_DISABLE); tput(1u << b); }
    for (uint32_t l = 0; l < 32; l++)
        if (t->lights & (1u << l)) { tput(T_DISABLELIGHT); tput(l); }
    g_tape = NULL;
    t->width = p->width;
    t->height = p->height;
    if (t->bad) {
        for (uint32_t k = 0; k < t->ntex; k++) free(t->texPixels[k]);
        free(t->w);
        free(t);
        return NULL;
    }
    return t;
}

void pfs_tape_free(PFStape* t)
{
    if (!t) return;
    for (uint32_t k = 0; k < t->ntex; k++) free(t->texPixels[k]);
    free(t->w);
    free(t);
}

uint64_t pfs_tape_calls(const PFStape* t) { return t ? t->calls : 0; }

static PFtexture tape_tex(const PFStape* t, uint32_t slot) { return (slot && slot <= t->ntex) ? t->bound[slot - 1] : NULL; }

static float word_f(uint32_t v) { float f; memcpy(&f, &v, 4); return f; }

/* Issues the tape's calls [from, to) (word offsets) through library `a`. */
static void tape_play(const PFSapi* a, const PFStape* t, size_t from, size_t to)
{
    const uint32_t* w = t->w + from;
    const uint32_t* end = t->w + to;
#define F(i) word_f(w[i])
    while (w < end) {
        switch (*w) {
        case T_ENABLE: a->Enable((PFstate)w[1]); w += 2; break;
        case T_DISABLE: a->Disable((PFstate)w[1]); w += 2; break;
        case T_CULLFACE: a->CullFace((PFface)w[1]); w += 2; break;
        case T_DEPTHFUNC: a->DepthFunc((PFdepthmode)w[1]); w += 2; break;
        case T_SHADEMODEL: a->ShadeModel((PFshademode)w[1]); w += 2; break;
        case T_BLENDFUNC: a->BlendFunc((PFblendmode)w[1]); w += 2; break;
        case T_MATRIXMODE: a->MatrixMode((PFmatrixmode)w[1]); w += 2; break;
        case T_CLEAR: a->Clear((PFclearflag)w[1]); w += 2; break;
        case T_BEGIN: a->Begin((PFdrawmode)w[1]); w += 2; break;
        case T_LIGHTMODEL: a->LightModel((PFlightmode)w[1]); w += 2; break;
        case T_ENABLELIGHT: a->EnableLight((PFsizei)w[1]); w += 2; break;
        case T_DISABLELIGHT: a->DisableLight((PFsizei)w[1]); w += 2; break;
        case T_LOADIDENTITY: a->LoadIdentity(); w += 1; break;
        case T_END: a->End(); w += 1; break;
        case T_TRANSLATE: a->Translatef(F(1), F(2), F(3)); w += 4; break;
        case T_VERTEX3: a->Vertex3f(F(1), F(2), F(3)); w += 4; break;
        case T_NORMAL3: a->Normal3f(F(1), F(2), F(3)); w += 4; break;
        case T_ROTATE: a->Rotatef(F(1), F(2), F(3), F(4)); w += 5; break;
        case T_FRUSTUM: a->Frustum(F(1), F(2), F(3), F(4), F(5), F(6)); w += 7; break;
        case T_TEXCOORD2: a->TexCoord2f(F(1), F(2)); w += 3; break;
        case T_COLOR4UB:
            a->Color4ub((PFubyte)w[1], (PFubyte)(w[1] >> 8), (PFubyte)(w[1] >> 16), (PFubyte)(w[1] >> 24));
            w += 2;
            break;
        case T_CLEARCOLOR:
            a->ClearColor((PFubyte)w[1], (PFubyte)(w[1] >> 8), (PFubyte)(w[1] >> 16), (PFubyte)(w[1] >> 24));
            w += 2;
            break;
        case T_LIGHTFV: { float v[3] = { F(3), F(4), F(5) }; a->Lightfv((PFsizei)w[1], (PFenum)w[2], v); w += 6; break; }
        case T_LIGHTF: a->Lightf((PFsizei)w[1], (PFenum)w[2], F(3)); w += 4; break;
        case T_MATERIALFV: { float v[3] = { F(3), F(4), F(5) }; a->Materialfv((PFface)w[1], (PFenum)w[2], v); w += 6; break; }
        case T_MATERIALF: a->Materialf((PFface)w[1], (PFenum)w[2], F(3)); w += 4; break;
        case T_COLORMATERIAL: a->ColorMaterial((PFface)w[1], (PFenum)w[2]); w += 3; break;
        case T_BINDTEXTURE: a->BindTexture(tape_tex(t, w[1])); w += 2; break;
        case T_TEXPARAM:
            a->TextureParameter(tape_tex(t, w[1]), (PFtexturewrap)w[2], (PFtexturefilter)w[3]);
            w += 4;
            break;
        default: return;                     /* not written by the recorder */
        }
    }
#undef F
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* No-op stand-ins for every call a tape makes: replaying a tape through
 * them times the application's own loop (decoding the tape and calling
 * through function pointers) without any library work. */
#define NOOP __attribute__((noinline, noipa))
static NOOP void n_st(PFstate s) { (void)s; }
static NOOP void n_face(PFface f) { (void)f; }
static NOOP void n_dfunc(PFdepthmode m) { (void)m; }
static NOOP void n_shade(PFshademode m) { (void)m; }
static NOOP void n_blend(PFblendmode m) { (void)m; }
static NOOP void n_mat(PFmatrixmode m) { (void)m; }
static NOOP void n_clear(PFclearflag f) { (void)f; }
static NOOP void n_begin(PFdrawmode m) { (void)m; }
static NOOP void n_lmodel(PFlightmode m) { (void)m; }
static NOOP void n_light(PFsizei l) { (void)l; }
static NOOP void n_void(void) { }
static NOOP void n_f3(PFfloat x, PFfloat y, PFfloat z) { (void)x; (void)y; (void)z; }
static NOOP void n_f4(PFfloat x, PFfloat y, PFfloat z, PFfloat w) { (void)x; (void)y; (void)z; (void)w; }
static NOOP void n_f6(PFfloat a, PFfloat b, PFfloat c, PFfloat d, PFfloat e, PFfloat f)
{ (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; }
static NOOP void n_f2(PFfloat u, PFfloat v) { (void)u; (void)v; }
static NOOP void n_ub4(PFubyte r, PFubyte g, PFubyte b, PFubyte a) { (void)r; (void)g; (void)b; (void)a; }
static NOOP void n_lfv(PFsizei l, PFenum p, const void* v) { (void)l; (void)p; (void)v; }
static NOOP void n_lf(PFsizei l, PFenum p, PFfloat v) { (void)l; (void)p; (void)v; }
static NOOP void n_mfv(PFface f, PFenum p, const void* v) { (void)f; (void)p; (void)v; }
static NOOP void n_mf(PFface f, PFenum p, PFfloat v) { (void)f; (void)p; (void)v; }
static NOOP void n_cmat(PFface f, PFenum m) { (void)f; (void)m; }
static NOOP void n_tex(PFtexture t) { (void)t; }
static NOOP void n_texp(PFtexture t, PFtexturewrap w, PFtexturefilter f) { (void)t; (void)w; (void)f; }
#undef NOOP

/* Wall time per frame (ms[i]) of replaying the tape's frame calls through
 * the no-op stand-ins: the part of pfs_tape_time's frame time that is the
 * application's loop rather than the library. */
int pfs_tape_time_loop(PFStape* t, uint32_t frames, double* ms)
{
    if (!t) return -1;
    PFSapi a;
    memset(&a, 0, sizeof a);
    a.Enable = n_st; a.Disable = n_st; a.CullFace = n_face; a.DepthFunc = n_dfunc; a.ShadeModel = n_shade;
    a.BlendFunc = n_blend; a.MatrixMode = n_mat; a.Clear = n_clear; a.Begin = n_begin; a.LightModel = n_lmodel;
    a.EnableLight = n_light; a.DisableLight = n_light; a.LoadIdentity = n_void; a.End = n_void;
    a.Translatef = n_f3; a.Vertex3f = n_f3; a.Normal3f = n_f3; a.Rotatef = n_f4; a.Frustum = n_f6;
    a.TexCoord2f = n_f2; a.Color4ub = n_ub4; a.ClearColor = n_ub4; a.Lightfv = n_lfv; a.Lightf = n_lf;
    a.Materialfv = n_mfv; a.Materialf = n_mf; a.ColorMaterial = n_cmat; a.BindTexture = n_tex;
    a.TextureParameter = n_texp;
    const PFSapi* volatile ap = &a;           /* the calls stay indirect, as through a loaded library */
    for (uint32_t f = 0; f < frames; f++) {
        double t0 = now_ms();
        tape_play(ap, t, 0, t->body);
        if (ms) ms[f] = now_ms() - t0;
        tape_play(ap, t, t->body, t->n);
    }
    return 0;
}

/* Renders the tape's scene `frames` times as an application would: one
 * context of library `a` over the caller's colour buffer (W*H+16 pixels),
 * the tape's textures created in the library once, then per frame the
 * scene's calls (pfClear and every draw) followed by pfFinish where the
 * library has one, and the tape's epilogue (untimed), which undoes what
 * the scene enabled so that every frame starts from the same state.
 * ms[i] receives frame i's wall time (CLOCK_MONOTONIC) from its first call
 * to the return of pfFinish (to the return of the last draw for the
 * reference, which draws inside the calls).  The buffer holds the last
 * frame.  Returns the library's error code, -1 on setup failure. */
int pfs_tape_time(const PFSapi* a, PFStape* t, uint32_t frames, uint32_t* color, double* ms)
{
    if (!a || !t || !color) return -1;
    PFcontext ctx = a->CreateContext(color, t->width, t->height, PF_RGBA, PF_UNSIGNED_BYTE);
    if (!ctx) return -1;
    a->MakeCurrent(ctx);
    for (uint32_t k = 0; k < t->ntex; k++)
        t->bound[k] = a->GenTexture(t->texPixels[k], t->texW[k], t->texH[k], (PFpixelformat)t->texFormat[k],
                                    (PFdatatype)t->texType[k]);
    int err = 0;
    for (uint32_t f = 0; f < frames && err == 0; f++) {
        double t0 = now_ms();
        tape_play(a, t, 0, t->body);
        if (a->Finish) a->Finish();
        if (ms) ms[f] = now_ms() - t0;
        tape_play(a, t, t->body, t->n);
        err = (int)a->GetError();
    }
    for (uint32_t k = 0; k < t->ntex; k++) { a->DeleteTexture(&t->bound[k], 0); t->bound[k] = NULL; }
    a->DeleteContext(ctx);
    a->MakeCurrent(NULL);
    return err;
}
