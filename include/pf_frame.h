/*
 * pf_frame.h -- the recorded frame: what the host library hands to the GPU.
 *
 * The reference rasterises each triangle synchronously inside pfVertex*
 * (context.c:1671-1674 -> internal/context/context.c:94-244).  This library
 * instead records, per render target, a frame made of
 *   - triangle streams: three vertices per recorded triangle (object-space
 *     position, RGBA8 colour, texture coordinate after the texture matrix,
 *     normal after PF_NORMALIZE), plus one word per triangle naming its
 *     state snapshot and which culling passes it runs (FRONT, BACK);
 *   - state snapshots: everything that pfiProcessRasterize_TRIANGLE_IMPL and
 *     Rasterize_Triangle read from the context (triangles.c:62-116, 287-558);
 *   - textures referenced by the snapshots;
 *   - an ordered command list of CLEAR and DRAW(range) commands.
 * All types are plain C structs; every pointer is a host pointer.
 */
#ifndef PF_FRAME_H
#define PF_FRAME_H

#include <stdint.h>

#define PFF_MAX_LIGHTS 8            /* config.h:37-39 */

/* Bits of PFFstate.flags; same values as PFstate in pixelforge.h. */
#define PFF_TEXTURE_2D   0x0001u
#define PFF_BLEND        0x0004u
#define PFF_DEPTH_TEST   0x0008u
#define PFF_LIGHTING     0x0040u    /* lighting enabled and >= 1 active light */

/* Bits of the per-triangle word (PFFtriangles.info). */
#define PFF_STATE_MASK   0x00FFFFFFu
#define PFF_PASS_FRONT   0x01000000u /* run the FRONT pass (triangles.c:303-305) */
#define PFF_PASS_BACK    0x02000000u /* run the BACK pass; after FRONT if both  */

#define PFF_NO_VERTEX    0xFFFFFFFFu /* corner without a colour-array index */

/* Texture pixel formats accepted by the pipeline. */
#define PFF_TEX_RGBA8    0u         /* pixel.h:696-700, 2909-2913 */
#define PFF_TEX_RGBAF32  1u         /* pixel.h:735-746, 2983-3008 */
#define PFF_TEX_SOURCE   2u         /* any other (PFpixelformat, PFdatatype)
                                     * pair of the reference's texel getter
                                     * table (pixel.h:3050-3095), named by
                                     * PFFtexture.source; its texels are read
                                     * as the AVX2 getters read them
                                     * (pixel.h:2250-3040) */

/* PFFtexture.source: PFpixelformat in bits 0-7, PFdatatype in bits 8-15. */
#define PFF_SOURCE(format, type) ((uint32_t)(format) | ((uint32_t)(type) << 8))

/* Bytes per texel of a source pair (pfiGetPixelBytes, pixel.h:3155-3212);
 * 0 for a pair the getter table lacks.  Formats: RED 0, GREEN 1, BLUE 2,
 * ALPHA 3, LUMINANCE 4, LUMINANCE_ALPHA 5, RGB 6, RGBA 7, BGR 8, BGRA 9;
 * types: UNSIGNED_BYTE 0, UNSIGNED_SHORT_5_6_5 2, _5_5_5_1 3, _4_4_4_4 4,
 * HALF_FLOAT 9, FLOAT 10. */
static inline uint32_t pff_source_texel_bytes(uint32_t source)
{
    uint32_t f = source & 0xFFu, t = (source >> 8) & 0xFFu;
    if (f > 9u) return 0u;
    uint32_t comps = f <= 4u ? 1u : f == 5u ? 2u : (f == 6u || f == 8u) ? 3u : 4u;
    switch (t) {
    case 0u:  return comps;
    case 9u:  return 2u * comps;
    case 10u: return 4u * comps;
    case 2u:  return (f == 6u || f == 8u) ? 2u : 0u;          /* RGB / BGR only  */
    case 3u: case 4u: return (f == 7u || f == 9u) ? 2u : 0u;  /* RGBA / BGRA only */
    default:  return 0u;
    }
}

typedef struct {
    float    position[3];
    float    direction[3];
    float    innerCutOff, outerCutOff;      /* cosines; >= pi disables the spot */
    float    attConstant, attLinear, attQuadratic;
    uint32_t ambient, diffuse, specular;    /* packed RGBA8, r in the low byte */
    uint32_t id;                            /* light number (pfEnableLight)    */
    uint32_t pad[2];
} PFFlight;                                 /* 80 bytes; context.h:216-229 */

typedef struct {
    uint32_t ambient, diffuse, specular, emission; /* packed RGBA8 */
    float    shininess;
    uint32_t pad[3];
} PFFmaterial;                              /* 32 bytes; context.h:234-240 */

typedef struct {
    float    mvp[16];          /* column-major, as pfmVec4Transform reads it  */
    float    matNormal[16];    /* transpose(inverse(model)) or identity       */
    float    viewRow2[4];      /* matView[8..10]: Gouraud N.V test (triangles.c:96) */
    float    viewPos[4];       /* translation of inverse(matView) (triangles.c:84-86) */
    int32_t  vpPos[2];         /* context.h:340 */
    uint32_t vpDim[2];         /* viewport size minus one */
    int32_t  vpMin[2];
    int32_t  vpMax[2];
    uint32_t flags;            /* PFF_* state bits */
    uint32_t depthFunc;        /* PFdepthmode */
    uint32_t blendMode;        /* PFblendmode */
    uint32_t shadeModel;       /* PFshademode */
    uint32_t lightMode;        /* PFlightmode */
    int32_t  texture;          /* index into PFFframe.textures, -1 for none   */
    uint32_t numLights;        /* active lights in enable order               */
    uint32_t pad0;
    PFFmaterial material[2];   /* [PF_FRONT], [PF_BACK] */
    PFFlight    lights[PFF_MAX_LIGHTS];
} PFFstate;

typedef struct {
    const void* pixels;        /* host memory, read at execution time         */
    uint32_t    width, height;
    uint32_t    format;        /* PFF_TEX_* */
    uint32_t    wrap;          /* PFtexturewrap */
    uint32_t    filter;        /* PFtexturefilter */
    uint32_t    source;        /* PFF_TEX_SOURCE: PFF_SOURCE(format, type) */
} PFFtexture;

#define PFF_CMD_CLEAR  1u
#define PFF_CMD_DRAW   2u
#define PFF_CMD_PIXELS 3u      /* pfDrawPixels: PFFframe.pixelOps[first]      */
#define PFF_CMD_FOG    4u      /* pfFogProcess: PFFframe.pixelOps[first]      */
#define PFF_CMD_POINTS 5u      /* points [first, last) of PFFframe.points      */
#define PFF_CMD_LINES  6u      /* lines [first, last) of PFFframe.lines        */

/* One line (a pair of PF_LINES vertices, or an edge of a triangle or quad
 * in PF_LINE polygon mode; lines.c:56-105), projected, clipped and
 * rasterised on the device (lines.c:107-532).  The edges of one polygon
 * form a group: an edge the clipper rejects ends its group (lines.c:93). */
typedef struct {
    float    position[2][4];   /* object space */
    uint32_t color[2];         /* packed RGBA8 */
    uint32_t state;            /* index into PFFframe.states */
    float    width;            /* pfLineWidth */
    uint32_t groupStart;       /* 1 for the first line of a group */
    uint32_t pad;
} PFFline;

/* One point (pfVertex in PF_POINTS mode, or a corner of a triangle or quad
 * in PF_POINT polygon mode; points.c:35-57), projected and rasterised on
 * the device (points.c:59-183). */
typedef struct {
    float    position[4];      /* object space */
    uint32_t color;            /* packed RGBA8 */
    uint32_t state;            /* index into PFFframe.states (MVP, viewport,
                                  depth test and blending) */
    float    size;             /* pfPointSize */
    uint32_t pad;
} PFFpoint;

/* PFFpixels.flags: a pfRect* fill (context.c:1938-1978) -- the 1x1 image
 * holds the current colour, both steps are 0, no depth is written. */
#define PFF_PIXELS_RECT  0x80000000u

/* One pfDrawPixels call (context.c:1987-2080), with everything the call
 * computes from the context before its pixel loop; or, for PFF_CMD_FOG,
 * one pfFogProcess call (context.c:2275-2343) and the fog state it reads
 * (image fields unused). */
typedef struct {
    const void* pixels;        /* copy of the caller's image: width*height
                                  texels of `source`                          */
    uint32_t    width, height;
    uint32_t    source;        /* PFF_SOURCE(format, type)                    */
    int32_t     xScreen, yScreen;          /* raster position on screen       */
    int32_t     xMin, yMin, xMax, yMax;    /* destination rectangle, inclusive */
    float       zPos;          /* depth written and tested                     */
    float       invXLen, invYLen;          /* 1 / (width * zoom), 1 / (height * zoom) */
    uint32_t    flags;         /* PFF_DEPTH_TEST, PFF_BLEND, PFF_PIXELS_RECT   */
    uint32_t    depthFunc;     /* PFdepthmode */
    uint32_t    blendMode;     /* PFblendmode */
    uint32_t    fogMode;       /* PFfogmode (any value; unknown ones give t = 0) */
    float       fogDensity, fogStart, fogEnd;
    uint32_t    fogColor;      /* packed RGBA8 */
} PFFpixels;

typedef struct {
    uint32_t kind;             /* PFF_CMD_* */
    uint32_t first, last;      /* DRAW: triangle range [first, last)          */
    uint32_t clearColor;       /* CLEAR: packed RGBA8                         */
    float    clearDepth;       /* CLEAR */
    uint32_t pad[3];
} PFFcmd;

typedef struct {
    uint32_t        count;
    const float*    position;  /* [count*3][4] object space, w included      */
    const uint32_t* color;     /* [count*3] packed RGBA8                      */
    const float*    texcoord;  /* [count*3][2]; NULL when no state textures   */
    const float*    normal;    /* [count*3][3]; NULL when no state lights     */
    const uint32_t* info;      /* [count] state index | PFF_PASS_* bits       */
    const uint32_t* vertexId;  /* [count*3] colour-array index of each corner
                                  (pfDrawArrays / pfDrawElements with a colour
                                  array), PFF_NO_VERTEX otherwise; NULL when
                                  no corner has one                          */
} PFFtriangles;

/* PFdatatype values of the array element types (pixelforge.h). */
#define PFF_UBYTE   0u
#define PFF_USHORT  1u
#define PFF_UINT    5u
#define PFF_SHORT   7u
#define PFF_INT     8u
#define PFF_FLOAT  10u
#define PFF_DOUBLE 11u

/* One pfDrawElements / pfDrawArrays call whose primitives all become filled
 * triangles (context.c:1231-1575 with ictx.c:94-244): the elements it reads
 * are copied into the frame when it is called, and its triangles fill the
 * slots [triFirst, triFirst + triCount) of PFFtriangles, which the host
 * leaves unwritten.  The vertex buffer of `mode` is filled with one element
 * per index; each time it is full (3 vertices for PF_TRIANGLES, 4 for
 * PF_TRIANGLE_FAN / _STRIP and PF_QUADS, 6 for PF_QUAD_FAN / _STRIP) it
 * gives, for each face group in order, the mode's triangles (1, 2, 2 or 4)
 * with that group's PFF_PASS_* bits, and the buffer keeps its last 1 (fans,
 * strips) or 2 (quad fans, quad strips) vertices.  A corner takes position
 * (x, y, z, w) from `position` (missing components 0, w 1), normal and
 * texture coordinate from their arrays or zero (no texture matrix, no
 * PF_NORMALIZE), and colour from `colors` (missing alpha 255, colour-array
 * index = element + elementBase) or `color` (index PFF_NO_VERTEX). */
typedef struct {
    uint32_t    mode;          /* PFdrawmode: PF_TRIANGLES .. PF_QUAD_STRIP    */
    uint32_t    count;         /* vertices the call issues                     */
    uint32_t    triFirst;      /* first triangle slot of the call              */
    uint32_t    triCount;
    uint32_t    numFaces;      /* face groups per full buffer, 1 or 2          */
    uint32_t    facePasses[2]; /* PFF_PASS_* bits of each group, in order      */
    uint32_t    state;         /* index into PFFframe.states                   */
    uint32_t    color;         /* packed RGBA8: corners without a colour array */
    uint32_t    elementBase;   /* colour-array index of element 0 (pfDrawArrays' first) */
    uint32_t    numElements;   /* elements in each array copy                  */
    const void* position;  int32_t posSize; uint32_t posType;   /* PF_SHORT, PF_INT, PF_FLOAT, PF_DOUBLE */
    const void* normal;    uint32_t nrmType;                    /* PF_FLOAT, PF_DOUBLE; NULL: none */
    const void* texcoord;  uint32_t uvType;                     /* PF_FLOAT, PF_DOUBLE; NULL: none */
    const void* colors;    int32_t colSize; uint32_t colType;   /* PF_UNSIGNED_BYTE/SHORT/INT, PF_FLOAT,
                                                                   PF_DOUBLE; NULL: none */
    const void* indices;   uint32_t idxType;                    /* PF_UNSIGNED_BYTE/SHORT/INT; NULL:
                                                                   vertex i is element i */
    uint32_t    flags;         /* PFF_ARRAYS_* */
    float       uvMatrix[6];   /* PFF_ARRAYS_TEXMAT: u' = (m0 u + m2 v) + m4,
                                  v' = (m1 u + m3 v) + m5 (pfmVec2Transform) */
    const void* device;        /* non-NULL: a device copy of the elements, the
                                  arrays at the same offsets from it as from
                                  `position` on the host */
} PFFarrays;

/* PFFarrays.flags.  A call of a render list (renderlist.c:69-97) is an
 * array draw of the list's recorded vertices whose colours come through
 * pfColor4ubv (no colour-array index), texture coordinates through
 * pfTexCoordfv (the current texture matrix applies) and normals through
 * pfNormal3fv (normalised under PF_NORMALIZE, unless of zero length). */
#define PFF_ARRAYS_LIST       0x1u  /* colours carry no colour-array index */
#define PFF_ARRAYS_TEXMAT     0x2u  /* texture coordinates through uvMatrix */
#define PFF_ARRAYS_NORMALIZE  0x4u  /* normals normalised */

/* Pixel formats of a render target's colour buffer (PFFframe.target): the
 * (format, type) pairs whose reference accessors read and write only their
 * own pixel -- the SIMD ones of the triangle raster and pfClear
 * (pixel.h:1013-1090, 1492-2250, 2328-2383, 2536-3040) and the scalar ones
 * of points, lines, pfDrawPixels, rects and fog (pixel.h:78-400,
 * 405-760).  BGRA / UNSIGNED_BYTE (its SIMD accessors copy one pixel's
 * bytes into its neighbours, pixel.h:2070-2078, 2916-2920), the
 * single-channel and luminance pairs (their SIMD getters read a
 * neighbour's byte, or their SIMD setters rewrite unmasked neighbours
 * through 8 bits, pixel.h:873-965, 1092-1487, 2253-2530) are not render
 * targets here.  Kinds from PFF_TARGET_LA_HALF on are "wide": more than 4
 * bytes, or not held as one stored word by the raster. */
#define PFF_TARGET_RGBA8     0u     /* RGBA / UNSIGNED_BYTE           */
#define PFF_TARGET_RGB8      1u     /* RGB / UNSIGNED_BYTE (3 bytes)  */
#define PFF_TARGET_BGR8      2u     /* BGR / UNSIGNED_BYTE (3 bytes)  */
#define PFF_TARGET_RGB565    3u     /* RGB / UNSIGNED_SHORT_5_6_5     */
#define PFF_TARGET_BGR565    4u     /* BGR / UNSIGNED_SHORT_5_6_5     */
#define PFF_TARGET_RGBA5551  5u     /* RGBA / UNSIGNED_SHORT_5_5_5_1  */
#define PFF_TARGET_BGRA5551  6u     /* BGRA / UNSIGNED_SHORT_5_5_5_1  */
#define PFF_TARGET_RGBA4444  7u     /* RGBA / UNSIGNED_SHORT_4_4_4_4  */
#define PFF_TARGET_BGRA4444  8u     /* BGRA / UNSIGNED_SHORT_4_4_4_4  */
#define PFF_TARGET_LA_HALF   9u     /* LUMINANCE_ALPHA / HALF_FLOAT   */
#define PFF_TARGET_LA_FLOAT  10u    /* LUMINANCE_ALPHA / FLOAT        */
#define PFF_TARGET_RGB_HALF  11u    /* RGB / HALF_FLOAT               */
#define PFF_TARGET_BGR_HALF  12u    /* BGR / HALF_FLOAT               */
#define PFF_TARGET_RGB_FLOAT 13u    /* RGB / FLOAT                    */
#define PFF_TARGET_BGR_FLOAT 14u    /* BGR / FLOAT                    */
#define PFF_TARGET_RGBA_HALF 15u    /* RGBA / HALF_FLOAT              */
#define PFF_TARGET_BGRA_HALF 16u    /* BGRA / HALF_FLOAT              */
#define PFF_TARGET_RGBA_FLOAT 17u   /* RGBA / FLOAT                   */
#define PFF_TARGET_BGRA_FLOAT 18u   /* BGRA / FLOAT                   */
#define PFF_TARGET_FIRST_WIDE PFF_TARGET_LA_HALF
#define PFF_TARGET_NONE      0xFFFFFFFFu

/* The target of a PFF_SOURCE pair, or PFF_TARGET_NONE. */
static inline uint32_t pff_target_of(uint32_t source)
{
    switch (source) {
    case PFF_SOURCE(7, 0): return PFF_TARGET_RGBA8;
    case PFF_SOURCE(6, 0): return PFF_TARGET_RGB8;
    case PFF_SOURCE(8, 0): return PFF_TARGET_BGR8;
    case PFF_SOURCE(6, 2): return PFF_TARGET_RGB565;
    case PFF_SOURCE(8, 2): return PFF_TARGET_BGR565;
    case PFF_SOURCE(7, 3): return PFF_TARGET_RGBA5551;
    case PFF_SOURCE(9, 3): return PFF_TARGET_BGRA5551;
    case PFF_SOURCE(7, 4): return PFF_TARGET_RGBA4444;
    case PFF_SOURCE(9, 4): return PFF_TARGET_BGRA4444;
    case PFF_SOURCE(5, 9): return PFF_TARGET_LA_HALF;
    case PFF_SOURCE(5, 10): return PFF_TARGET_LA_FLOAT;
    case PFF_SOURCE(6, 9): return PFF_TARGET_RGB_HALF;
    case PFF_SOURCE(8, 9): return PFF_TARGET_BGR_HALF;
    case PFF_SOURCE(6, 10): return PFF_TARGET_RGB_FLOAT;
    case PFF_SOURCE(8, 10): return PFF_TARGET_BGR_FLOAT;
    case PFF_SOURCE(7, 9): return PFF_TARGET_RGBA_HALF;
    case PFF_SOURCE(9, 9): return PFF_TARGET_BGRA_HALF;
    case PFF_SOURCE(7, 10): return PFF_TARGET_RGBA_FLOAT;
    case PFF_SOURCE(9, 10): return PFF_TARGET_BGRA_FLOAT;
    default: return PFF_TARGET_NONE;
    }
}

/* PFF_SOURCE pair of a target (PFF_TARGET_NONE for an unknown one). */
static inline uint32_t pff_target_source(uint32_t target)
{
    static const uint32_t src[19] = {
        PFF_SOURCE(7, 0), PFF_SOURCE(6, 0), PFF_SOURCE(8, 0), PFF_SOURCE(6, 2), PFF_SOURCE(8, 2),
        PFF_SOURCE(7, 3), PFF_SOURCE(9, 3), PFF_SOURCE(7, 4), PFF_SOURCE(9, 4), PFF_SOURCE(5, 9),
        PFF_SOURCE(5, 10), PFF_SOURCE(6, 9), PFF_SOURCE(8, 9), PFF_SOURCE(6, 10), PFF_SOURCE(8, 10),
        PFF_SOURCE(7, 9), PFF_SOURCE(9, 9), PFF_SOURCE(7, 10), PFF_SOURCE(9, 10) };
    return target < 19u ? src[target] : PFF_TARGET_NONE;
}

/* Bytes per pixel of a target (0 for an unknown one). */
static inline uint32_t pff_target_bytes(uint32_t target)
{
    return target < 19u ? pff_source_texel_bytes(pff_target_source(target)) : 0u;
}

typedef struct {
    uint32_t           width, height;    /* render target + float32 depth        */
    void*              color;            /* host buffers of the render target:   */
    float*             depth;            /* width*height pixels of `target`      */
    PFFtriangles       tris;
    const PFFstate*    states;   uint32_t numStates;
    const PFFtexture*  textures; uint32_t numTextures;
    const PFFcmd*      cmds;     uint32_t numCmds;
    const PFFpixels*   pixelOps; uint32_t numPixelOps;
    const PFFpoint*    points;   uint32_t numPoints;
    const PFFline*     lines;    uint32_t numLines;
    const PFFarrays*   arrays;   uint32_t numArrays;   /* in triFirst order */
    uint32_t           target;   /* PFF_TARGET_*: pixel format of `color`  */
} PFFframe;

#endif /* PF_FRAME_H */
