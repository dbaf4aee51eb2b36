/*
 * pixelforge.h -- public C API of the MI355X-native PixelForge fragment pipeline.
 *
 * This header is the drop-in boundary.  Every type, enum value and function
 * signature below matches the reference header `src/pixelforge.h` so that an
 * application written against the reference compiles and links unchanged
 * against libpixelforge_amd.so.  Each declaration cites the reference
 * declaration it replaces as `pixelforge.h:LINE` (and the implementation as
 * `context.c:LINE`, `texture.c:LINE`, `framebuffer.c:LINE`).
 *
 * What differs from the reference, by design:
 *   - Triangles are recorded, not rasterised, when their third vertex
 *     arrives.  Recorded work runs on the GPU at a synchronisation point:
 *     pfFinish(), pfReadPixels(), pfGetFramebufferPixel/Depth(),
 *     pfSwapBuffers(), pfSetMainBuffer(), pfDeleteContext(), and any change
 *     of render target.  Code that reads a colour buffer directly must call
 *     pfFinish() first (new entry point, no reference counterpart).
 *   - Only RGBA / UNSIGNED_BYTE render targets are accepted.  Textures may
 *     use every (format, type) pair that has a texel getter in the reference
 *     (pixel.h:3050-3095); pairs without one are PF_INVALID_ENUM.
 *   - Points, lines, rects, pfDrawPixels and pfFogProcess are drawn on the
 *     GPU in order with the triangles; pfPostProcess runs the caller's
 *     function on the host after the frame.  pfFogProcess runs on the GPU in order
 *     with the draws; pfPostProcess runs the caller's function on the host
 *     after the frame has been drawn.
 *
 * Errors follow the reference convention: a sticky last-error code that
 * pfGetError() reads and clears (context.c:380-385).  GPU failures map to
 * PF_ERROR_OUT_OF_MEMORY (allocation) or PF_INVALID_OPERATION (anything
 * else) and are also reported on stderr.
 */
#ifndef PIXELFORGE_AMD_H
#define PIXELFORGE_AMD_H

#include <stdint.h>
#include <stdbool.h>

#ifndef PF_API
#   define PF_API __attribute__((visibility("default")))
#endif

/* ------------------------------------------------------------------------ */
/* Scalar types (pixelforge.h:120-134)                                       */
/* ------------------------------------------------------------------------ */

typedef bool        PFboolean;
#define PF_FALSE 0
#define PF_TRUE  1

typedef int8_t      PFbyte;
typedef uint8_t     PFubyte;
typedef int16_t     PFshort;
typedef uint16_t    PFushort;
typedef int32_t     PFint;
typedef uint32_t    PFuint;
typedef int64_t     PFint64;
typedef uint64_t    PFuint64;
typedef uint32_t    PFsizei;
typedef uint32_t    PFenum;
typedef intptr_t    PFintptr;
typedef uintptr_t   PFsizeiptr;
typedef float       PFfloat;
typedef double      PFdouble;

/* Pixel data types (pixelforge.h:140-153). */
typedef enum {
    PF_UNSIGNED_BYTE,
    PF_UNSIGNED_SHORT,
    PF_UNSIGNED_SHORT_5_6_5,
    PF_UNSIGNED_SHORT_5_5_5_1,
    PF_UNSIGNED_SHORT_4_4_4_4,
    PF_UNSIGNED_INT,
    PF_BYTE,
    PF_SHORT,
    PF_INT,
    PF_HALF_FLOAT,
    PF_FLOAT,
    PF_DOUBLE
} PFdatatype;

/* Opaque handles (pixelforge.h:168, 420, 424). */
typedef void* PFcontext;
typedef void* PFtexture;
typedef void* PFrenderlist;

/* State bits for pfEnable/pfDisable (pixelforge.h:170-183). */
typedef enum {
    PF_TEXTURE_2D           = 0x0001,
    PF_FRAMEBUFFER          = 0x0002,
    PF_BLEND                = 0x0004,
    PF_DEPTH_TEST           = 0x0008,
    PF_CULL_FACE            = 0x0010,
    PF_NORMALIZE            = 0x0020,
    PF_LIGHTING             = 0x0040,
    PF_COLOR_MATERIAL       = 0x0080,
    PF_VERTEX_ARRAY         = 0x0100,
    PF_NORMAL_ARRAY         = 0x0200,
    PF_COLOR_ARRAY          = 0x0400,
    PF_TEXTURE_COORD_ARRAY  = 0x0800
} PFstate;

/* Error codes (pixelforge.h:251-271; release-build values). */
typedef enum {
    PF_NO_ERROR,
    PF_INVALID_ENUM,
    PF_INVALID_VALUE,
    PF_STACK_OVERFLOW,
    PF_STACK_UNDERFLOW,
    PF_INVALID_OPERATION,
    PF_ERROR_OUT_OF_MEMORY
} PFerrcode;

/* pixelforge.h:275-278 */
typedef enum {
    PF_COLOR_BUFFER_BIT = 0x01,
    PF_DEPTH_BUFFER_BIT = 0x02
} PFclearflag;

/* pixelforge.h:280-284 */
typedef enum {
    PF_MODELVIEW,
    PF_PROJECTION,
    PF_TEXTURE
} PFmatrixmode;

/* pixelforge.h:286-295 */
typedef enum {
    PF_POINTS,
    PF_LINES,
    PF_TRIANGLES,
    PF_TRIANGLE_FAN,
    PF_TRIANGLE_STRIP,
    PF_QUADS,
    PF_QUAD_FAN,
    PF_QUAD_STRIP
} PFdrawmode;

/* pixelforge.h:297-306 */
typedef enum {
    PF_BLEND_AVERAGE,
    PF_BLEND_ALPHA,
    PF_BLEND_ADD,
    PF_BLEND_SUB,
    PF_BLEND_MUL,
    PF_BLEND_SCREEN,
    PF_BLEND_LIGHTEN,
    PF_BLEND_DARKEN
} PFblendmode;

/* pixelforge.h:308-315 */
typedef enum {
    PF_EQUAL,
    PF_NOTEQUAL,
    PF_LESS,
    PF_LEQUAL,
    PF_GREATER,
    PF_GEQUAL
} PFdepthmode;

/* pixelforge.h:317-321 */
typedef enum { PF_POINT, PF_LINE, PF_FILL } PFpolygonmode;
/* pixelforge.h:323-326 */
typedef enum { PF_FLAT, PF_SMOOTH } PFshademode;
/* pixelforge.h:328-331 */
typedef enum { PF_GOURAUD, PF_PHONG } PFlightmode;
/* pixelforge.h:333-337; PF_FRONT and PF_BACK double as indices. */
typedef enum { PF_FRONT = 0, PF_BACK = 1, PF_FRONT_AND_BACK } PFface;
/* pixelforge.h:339-343 */
typedef enum { PF_REPEAT, PF_MIRRORED_REPEAT, PF_CLAMP_TO_EDGE } PFtexturewrap;
/* pixelforge.h:345-348 */
typedef enum { PF_NEAREST, PF_BILINEAR } PFtexturefilter;

/* pixelforge.h:350-360 */
typedef enum {
    PF_LIGHT0 = 0, PF_LIGHT1, PF_LIGHT2, PF_LIGHT3,
    PF_LIGHT4, PF_LIGHT5, PF_LIGHT6, PF_LIGHT7, PF_LIGHT8
} PFlights;

/* pixelforge.h:362-367 */
typedef enum {
    PF_AMBIENT_AND_DIFFUSE      = 1,
    PF_AMBIENT                  = 2,
    PF_DIFFUSE                  = 3,
    PF_SPECULAR                 = 4
} PFrendercolor;

/* pixelforge.h:369-372 */
typedef enum {
    PF_EMISSION                 = 5,
    PF_SHININESS                = 6
} PFmaterialparam;

/* pixelforge.h:374-382 */
typedef enum {
    PF_POSITION                 = 7,
    PF_SPOT_DIRECTION           = 8,
    PF_SPOT_INNER_CUTOFF        = 10,
    PF_SPOT_OUTER_CUTOFF        = 11,
    PF_CONSTANT_ATTENUATION     = 12,
    PF_LINEAR_ATTENUATION       = 13,
    PF_QUADRATIC_ATTENUATION    = 14
} PFlightparam;

/* Queries of pfGet*v (pixelforge.h:185-247). */
typedef enum {
    PF_VIEWPORT                         = 10000,
    PF_COLOR_CLEAR_VALUE,
    PF_DEPTH_CLEAR_VALUE,
    PF_CULL_FACE_MODE,
    PF_CURRENT_COLOR,
    PF_CURRENT_NORMAL,
    PF_CURRENT_TEXTURE_COORDS,
    PF_CURRENT_RASTER_POSITION,
    PF_BLEND_FUNC,
    PF_DEPTH_FUNC,
    PF_POLYGON_MODE,
    PF_POINT_SIZE,
    PF_LINE_WIDTH,
    PF_MATRIX_MODE,
    PF_PROJECTION_MATRIX,
    PF_MODELVIEW_MATRIX,
    PF_TEXTURE_MATRIX,
    PF_MAX_PROJECTION_STACK_DEPTH,
    PF_MAX_MODELVIEW_STACK_DEPTH,
    PF_MAX_TEXTURE_STACK_DEPTH,
    PF_SHADE_MODEL,
    PF_MAX_LIGHTS,
    PF_VERTEX_ARRAY_SIZE,
    PF_VERTEX_ARRAY_STRIDE,
    PF_VERTEX_ARRAY_TYPE,
    PF_NORMAL_ARRAY_STRIDE,
    PF_NORMAL_ARRAY_TYPE,
    PF_TEXTURE_COORD_ARRAY_STRIDE,
    PF_TEXTURE_COORD_ARRAY_TYPE,
    PF_COLOR_ARRAY_SIZE,
    PF_COLOR_ARRAY_STRIDE,
    PF_COLOR_ARRAY_TYPE,
    PF_ZOOM_X,
    PF_ZOOM_Y
} PFgettable;

/* pixelforge.h:385-397 */
typedef enum { PF_FOG_MODE, PF_FOG_DENSITY, PF_FOG_START, PF_FOG_END, PF_FOG_COLOR } PFfogparam;
typedef enum { PF_LINEAR, PF_EXP, PF_EXP2 } PFfogmode;

/* pixelforge.h:398-400: bytes r,g,b,a in memory order. */
typedef struct { PFubyte r, g, b, a; } PFcolor;

/* pixelforge.h:406-417 */
typedef enum {
    PF_RED, PF_GREEN, PF_BLUE, PF_ALPHA,
    PF_LUMINANCE, PF_LUMINANCE_ALPHA,
    PF_RGB, PF_RGBA, PF_BGR, PF_BGRA
} PFpixelformat;

/* pixelforge.h:403 */
typedef PFcolor (*PFpostprocessfunc)(PFint, PFint, PFfloat, PFcolor);

/* pixelforge.h:428-431: callers may dereference zbuffer directly. */
typedef struct {
    PFtexture texture;
    PFfloat  *zbuffer;
} PFframebuffer;

#if defined(__cplusplus)
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Context (pixelforge.h:452-627; context.c:115-385)                         */
/* ------------------------------------------------------------------------ */

PF_API PFcontext pfCreateContext(void* targetBuffer, PFsizei width, PFsizei height,
                                 PFpixelformat format, PFdatatype type);   /* pixelforge.h:452 */
PF_API void      pfDeleteContext(PFcontext ctx);                          /* pixelforge.h:465 */
PF_API void      pfSetMainBuffer(void* targetBuffer, PFsizei width, PFsizei height,
                                 PFpixelformat format, PFdatatype type);   /* pixelforge.h:486 */
PF_API void      pfSetAuxBuffer(void* auxFramebuffer);                     /* pixelforge.h:499 */
PF_API void      pfSwapBuffers(void);                                      /* pixelforge.h:507 */
PF_API PFcontext pfGetCurrentContext(void);                                /* pixelforge.h:517 */
PF_API void      pfMakeCurrent(PFcontext ctx);                             /* pixelforge.h:527 */
PF_API PFboolean pfIsEnabled(PFstate state);                               /* pixelforge.h:539 */
PF_API void      pfEnable(PFstate state);                                  /* pixelforge.h:549 */
PF_API void      pfDisable(PFstate state);                                 /* pixelforge.h:559 */
PF_API PFerrcode pfGetError(void);                                         /* pixelforge.h:627 */
PF_API void      pfGetBooleanv(PFenum pname, PFboolean* params);           /* pixelforge.h:574 */
PF_API void      pfGetIntegerv(PFenum pname, PFint* params);               /* pixelforge.h:585 */
PF_API void      pfGetFloatv(PFenum pname, PFfloat* params);               /* pixelforge.h:596 */
PF_API void      pfGetDoublev(PFenum pname, PFdouble* params);             /* pixelforge.h:607 */
PF_API void      pfGetPointerv(PFenum pname, const void** params);         /* pixelforge.h:618 */

/* ------------------------------------------------------------------------ */
/* Matrices (pixelforge.h:640-743; context.c:395-548)                        */
/* ------------------------------------------------------------------------ */

PF_API void pfMatrixMode(PFmatrixmode mode);                               /* pixelforge.h:640 */
PF_API void pfPushMatrix(void);                                            /* pixelforge.h:648 */
PF_API void pfPopMatrix(void);                                             /* pixelforge.h:656 */
PF_API void pfLoadIdentity(void);                                          /* pixelforge.h:664 */
PF_API void pfTranslatef(PFfloat x, PFfloat y, PFfloat z);                 /* pixelforge.h:676 */
PF_API void pfRotatef(PFfloat angle, PFfloat x, PFfloat y, PFfloat z);     /* pixelforge.h:689 */
PF_API void pfScalef(PFfloat x, PFfloat y, PFfloat z);                     /* pixelforge.h:701 */
PF_API void pfMultMatrixf(const PFfloat* mat);                             /* pixelforge.h:711 */
PF_API void pfFrustum(PFfloat left, PFfloat right, PFfloat bottom, PFfloat top,
                      PFfloat znear, PFfloat zfar);                        /* pixelforge.h:726 */
PF_API void pfOrtho(PFfloat left, PFfloat right, PFfloat bottom, PFfloat top,
                    PFfloat znear, PFfloat zfar);                          /* pixelforge.h:743 */

/* ------------------------------------------------------------------------ */
/* Render state (pixelforge.h:761-921; context.c:553-797)                    */
/* ------------------------------------------------------------------------ */

PF_API void pfViewport(PFint x, PFint y, PFsizei width, PFsizei height);   /* pixelforge.h:761 */
PF_API void pfPolygonMode(PFface face, PFpolygonmode mode);                /* pixelforge.h:776 */
PF_API void pfShadeModel(PFshademode mode);                                /* pixelforge.h:790 */
PF_API void pfLightModel(PFlightmode mode);                                /* pixelforge.h:812 */
PF_API void pfCullFace(PFface face);                                       /* pixelforge.h:842 */
PF_API void pfBlendFunc(PFblendmode mode);                                 /* pixelforge.h:852 */
PF_API void pfDepthFunc(PFdepthmode mode);                                 /* pixelforge.h:862 */
PF_API void pfBindFramebuffer(PFframebuffer* framebuffer);                 /* pixelforge.h:875 */
PF_API void pfBindTexture(PFtexture texture);                              /* pixelforge.h:888 */
PF_API void pfClear(PFclearflag flag);                                     /* pixelforge.h:898 */
PF_API void pfClearDepth(PFfloat depth);                                   /* pixelforge.h:908 */
PF_API void pfClearColor(PFubyte r, PFubyte g, PFubyte b, PFubyte a);      /* pixelforge.h:921 */

/* ------------------------------------------------------------------------ */
/* Lights and materials (pixelforge.h:934-1014; context.c:802-1156)          */
/* ------------------------------------------------------------------------ */

PF_API void      pfEnableLight(PFsizei light);                             /* pixelforge.h:934 */
PF_API void      pfDisableLight(PFsizei light);                            /* pixelforge.h:944 */
PF_API PFboolean pfIsEnabledLight(PFsizei light);                          /* pixelforge.h:955 */
PF_API void      pfLightf(PFsizei light, PFenum param, PFfloat value);     /* pixelforge.h:967 */
PF_API void      pfLightfv(PFsizei light, PFenum param, const void* value);/* pixelforge.h:979 */
PF_API void      pfMaterialf(PFface face, PFenum param, PFfloat value);    /* pixelforge.h:991 */
PF_API void      pfMaterialfv(PFface face, PFenum param, const void* value);/* pixelforge.h:1003 */
PF_API void      pfColorMaterial(PFface face, PFenum mode);                /* pixelforge.h:1014 */

/* ------------------------------------------------------------------------ */
/* Immediate mode (pixelforge.h:1107-1460; context.c:1580-1913)              */
/* ------------------------------------------------------------------------ */

PF_API void pfBegin(PFdrawmode mode);                                      /* pixelforge.h:1107 */
PF_API void pfEnd(void);                                                   /* pixelforge.h:1115 */
PF_API void pfVertex2i(PFint x, PFint y);                                  /* pixelforge.h:1126 */
PF_API void pfVertex2f(PFfloat x, PFfloat y);                              /* pixelforge.h:1137 */
PF_API void pfVertex2fv(const PFfloat* v);                                 /* pixelforge.h:1147 */
PF_API void pfVertex3i(PFint x, PFint y, PFint z);                         /* pixelforge.h:1159 */
PF_API void pfVertex3f(PFfloat x, PFfloat y, PFfloat z);                   /* pixelforge.h:1171 */
PF_API void pfVertex3fv(const PFfloat* v);                                 /* pixelforge.h:1181 */
PF_API void pfVertex4i(PFint x, PFint y, PFint z, PFint w);                /* pixelforge.h:1194 */
PF_API void pfVertex4f(PFfloat x, PFfloat y, PFfloat z, PFfloat w);        /* pixelforge.h:1207 */
PF_API void pfVertex4fv(const PFfloat* v);                                 /* pixelforge.h:1217 */
PF_API void pfColor(PFcolor color);                                        /* pixelforge.h:1227 */
PF_API void pfColor1ui(PFuint color);                                      /* pixelforge.h:1237 */
PF_API void pfColor3ub(PFubyte r, PFubyte g, PFubyte b);                   /* pixelforge.h:1249 */
PF_API void pfColor3ubv(const PFubyte* v);                                 /* pixelforge.h:1259 */
PF_API void pfColor3us(PFushort r, PFushort g, PFushort b);                /* pixelforge.h:1271 */
PF_API void pfColor3usv(const PFushort* v);                                /* pixelforge.h:1281 */
PF_API void pfColor3ui(PFuint r, PFuint g, PFuint b);                      /* pixelforge.h:1293 */
PF_API void pfColor3uiv(const PFuint* v);                                  /* pixelforge.h:1303 */
PF_API void pfColor3f(PFfloat r, PFfloat g, PFfloat b);                    /* pixelforge.h:1315 */
PF_API void pfColor3fv(const PFfloat* v);                                  /* pixelforge.h:1325 */
PF_API void pfColor4ub(PFubyte r, PFubyte g, PFubyte b, PFubyte a);        /* pixelforge.h:1338 */
PF_API void pfColor4ubv(const PFubyte* v);                                 /* pixelforge.h:1348 */
PF_API void pfColor4us(PFushort r, PFushort g, PFushort b, PFushort a);    /* pixelforge.h:1361 */
PF_API void pfColor4usv(const PFushort* v);                                /* pixelforge.h:1371 */
PF_API void pfColor4ui(PFuint r, PFuint g, PFuint b, PFuint a);            /* pixelforge.h:1384 */
PF_API void pfColor4uiv(const PFuint* v);                                  /* pixelforge.h:1394 */
PF_API void pfColor4f(PFfloat r, PFfloat g, PFfloat b, PFfloat a);         /* pixelforge.h:1407 */
PF_API void pfColor4fv(const PFfloat* v);                                  /* pixelforge.h:1417 */
PF_API void pfTexCoord2f(PFfloat u, PFfloat v);                            /* pixelforge.h:1428 */
PF_API void pfTexCoordfv(const PFfloat* v);                                /* pixelforge.h:1438 */
PF_API void pfNormal3f(PFfloat x, PFfloat y, PFfloat z);                   /* pixelforge.h:1450 */
PF_API void pfNormal3fv(const PFfloat* v);                                 /* pixelforge.h:1460 */

/* ------------------------------------------------------------------------ */
/* Vertex arrays (pixelforge.h:1031-1095; context.c:1161-1575).  As in the   */
/* reference, `stride` is stored but every array is read tightly packed.     */
/* ------------------------------------------------------------------------ */

PF_API void pfVertexPointer(PFint size, PFenum type, PFsizei stride, const void* pointer); /* pixelforge.h:1031 */
PF_API void pfNormalPointer(PFenum type, PFsizei stride, const void* pointer);             /* pixelforge.h:1044 */
PF_API void pfTexCoordPointer(PFenum type, PFsizei stride, const void* pointer);           /* pixelforge.h:1056 */
PF_API void pfColorPointer(PFint size, PFenum type, PFsizei stride, const void* pointer);  /* pixelforge.h:1069 */
PF_API void pfDrawElements(PFdrawmode mode, PFsizei count, PFdatatype type, const void* indices); /* pixelforge.h:1082 */
PF_API void pfDrawArrays(PFdrawmode mode, PFint first, PFsizei count);                     /* pixelforge.h:1095 */

/* ------------------------------------------------------------------------ */
/* Readback (pixelforge.h:1717; context.c:2349-2395)                         */
/* ------------------------------------------------------------------------ */

PF_API void pfReadPixels(PFint x, PFint y, PFsizei width, PFsizei height,
                         PFpixelformat format, PFdatatype type, void* pixels); /* pixelforge.h:1717 */

/* ------------------------------------------------------------------------ */
/* Framebuffers (pixelforge.h:1886-2002; framebuffer.c:29-148)               */
/* ------------------------------------------------------------------------ */

PF_API PFframebuffer pfGenFramebuffer(PFsizei width, PFsizei height,
                                      PFpixelformat format, PFdatatype type); /* pixelforge.h:1886 */
PF_API void      pfDeleteFramebuffer(PFframebuffer* framebuffer);          /* pixelforge.h:1898 */
PF_API PFboolean pfIsValidFramebuffer(PFframebuffer* framebuffer);         /* pixelforge.h:1910 */
PF_API void      pfClearFramebuffer(PFframebuffer* framebuffer, PFcolor color, PFfloat depth); /* pixelforge.h:1923 */
PF_API PFcolor   pfGetFramebufferPixel(const PFframebuffer* framebuffer, PFsizei x, PFsizei y); /* pixelforge.h:1937 */
PF_API PFfloat   pfGetFramebufferDepth(const PFframebuffer* framebuffer, PFsizei x, PFsizei y); /* pixelforge.h:1951 */
PF_API void      pfSetFramebufferPixelDepthTest(PFframebuffer* framebuffer, PFsizei x, PFsizei y, PFfloat z,
                                                PFcolor color, PFdepthmode depthMode); /* pixelforge.h:1969 */
PF_API void      pfSetFramebufferPixelDepth(PFframebuffer* framebuffer, PFsizei x, PFsizei y, PFfloat z,
                                            PFcolor color);                /* pixelforge.h:1986 */
PF_API void      pfSetFramebufferPixel(PFframebuffer* framebuffer, PFsizei x, PFsizei y, PFcolor color); /* pixelforge.h:2003 */

/* ------------------------------------------------------------------------ */
/* Textures (pixelforge.h:2024-2085; texture.c:30-123)                       */
/* ------------------------------------------------------------------------ */

PF_API PFtexture pfGenTexture(void* pixels, PFsizei width, PFsizei height,
                              PFpixelformat format, PFdatatype type);     /* pixelforge.h:2024 */
PF_API void      pfDeleteTexture(PFtexture* texture, PFboolean freeBuffer);/* pixelforge.h:2039 */
PF_API PFboolean pfIsValidTexture(const PFtexture texture);                /* pixelforge.h:2052 */
PF_API void      pfTextureParameter(PFtexture texture, PFtexturewrap wrapMode,
                                    PFtexturefilter filterMode);           /* pixelforge.h:2066 */
PF_API void*     pfGetTexturePixels(const PFtexture texture, PFsizei* width, PFsizei* height,
                                    PFpixelformat* format, PFdatatype* type); /* pixelforge.h:2085 */

/* ------------------------------------------------------------------------ */
/* Beside the triangle path: pfLineWidth and pfPointSize apply to lines and  */
/* points; rectangles fill with the current colour.  pfDrawPixels (with the  */
/* raster position and zoom) and pfFogProcess are recorded and run on the    */
/* GPU in order with the frame's triangles; pfPostProcess runs on the host   */
/* after the frame.                                                          */
/* ------------------------------------------------------------------------ */

PF_API void pfLineWidth(PFfloat width);                                    /* pixelforge.h:823 */
PF_API void pfPointSize(PFfloat size);                                     /* pixelforge.h:833 */
PF_API void pfRects(PFshort x1, PFshort y1, PFshort x2, PFshort y2);       /* pixelforge.h:1477 */
PF_API void pfRectsv(const PFshort* v1, const PFshort* v2);                /* pixelforge.h:1488 */
PF_API void pfRectf(PFfloat x1, PFfloat y1, PFfloat x2, PFfloat y2);       /* pixelforge.h:1501 */
PF_API void pfRectfv(const PFfloat* v1, const PFfloat* v2);                /* pixelforge.h:1511 */
PF_API void pfDrawPixels(PFsizei width, PFsizei height, PFpixelformat format, PFdatatype type,
                         const void* pixels);                              /* pixelforge.h:1527 */
PF_API void pfPixelZoom(PFfloat xfactor, PFfloat yfactor);                 /* pixelforge.h:1540 */
PF_API void pfRasterPos2i(PFint x, PFint y);                               /* pixelforge.h:1551 */
PF_API void pfRasterPos2f(PFfloat x, PFfloat y);                           /* pixelforge.h:1562 */
PF_API void pfRasterPos2fv(const PFfloat* v);                              /* pixelforge.h:1572 */
PF_API void pfRasterPos3i(PFint x, PFint y, PFint z);                      /* pixelforge.h:1584 */
PF_API void pfRasterPos3f(PFfloat x, PFfloat y, PFfloat z);                /* pixelforge.h:1596 */
PF_API void pfRasterPos3fv(const PFfloat* v);                              /* pixelforge.h:1606 */
PF_API void pfRasterPos4i(PFint x, PFint y, PFint z, PFint w);             /* pixelforge.h:1619 */
PF_API void pfRasterPos4f(PFfloat x, PFfloat y, PFfloat z, PFfloat w);     /* pixelforge.h:1632 */
PF_API void pfRasterPos4fv(const PFfloat* v);                              /* pixelforge.h:1642 */
PF_API void pfFogi(PFfogparam pname, PFint param);                         /* pixelforge.h:1656 */
PF_API void pfFogf(PFfogparam pname, PFfloat param);                       /* pixelforge.h:1667 */
PF_API void pfFogiv(PFfogparam pname, PFint* param);                       /* pixelforge.h:1678 */
PF_API void pfFogfv(PFfogparam pname, PFfloat* param);                     /* pixelforge.h:1689 */
PF_API void pfFogProcess(void);                                            /* pixelforge.h:1697 */
PF_API void pfPostProcess(PFpostprocessfunc postProcessFunction);          /* pixelforge.h:1735 */
/* ------------------------------------------------------------------------ */
/* Render lists.  Recording stores each pfBegin's draw mode, materials and   */
/* texture and each vertex's position, colour, texture coordinate and       */
/* normal; pfCallList replays them through immediate mode, so the triangles  */
/* go to the GPU with the frame.  pfNewList and pfCallList save colour,      */
/* normal, texture coordinate, materials, texture and enable bits in one     */
/* backup slot, restored by pfEndList and at the end of pfCallList.          */
/* ------------------------------------------------------------------------ */

PF_API PFrenderlist pfGenList(void);                                       /* pixelforge.h:1768 */
PF_API void pfDeleteList(PFrenderlist* renderList);                        /* pixelforge.h:1793 */
PF_API void pfNewList(PFrenderlist renderList);                            /* pixelforge.h:1817 */
PF_API void pfEndList(void);                                               /* pixelforge.h:1838 */
PF_API void pfCallList(const PFrenderlist renderList);                     /* pixelforge.h:1867 */

/* ------------------------------------------------------------------------ */
/* New entry points (no reference counterpart)                               */
/* ------------------------------------------------------------------------ */

/* Runs every recorded command of the current context on the GPU and copies
 * the render target back into the caller's colour and depth buffers.
 * Needed because the reference writes pixels before each call returns
 * (context.c:1671-1674) while this library defers them. */
PF_API void pfFinish(void);

/* Submits recorded work to the GPU without waiting for it. */
PF_API void pfFlush(void);

/* Renders the pending work of `count` contexts in one batched launch and
 * waits for it; each context's buffers hold its result on return. */
PF_API void pfSubmitBatch(const PFcontext* contexts, PFsizei count);

/* Differentiable rendering.  pfDiffRender runs the current context's
 * recorded frame once through the GPU pipeline (the target receives the
 * usual RGBA8 image) and keeps, for every pixel, the fragment that was
 * written last; it then writes a float image [H][W][4] of those fragments
 * shaded without quantisation from the given parameters.  Called again with
 * nothing recorded in between, it re-shades the same fragments with new
 * parameter values (coverage stays fixed).  pfDiffBackward returns the
 * gradients of a loss for dL/dimage [H][W][4].  Vertex colour parameter i
 * is element i of the colour array of pfDrawElements / pfDrawArrays; light
 * parameters are RGB per light number.  Frames with blending, flat shading,
 * Gouraud lighting or a clear after the first command set
 * PF_INVALID_OPERATION. */
typedef struct {
    PFtexture      texture;        /* texture whose texels are parameters, NULL: none        */
    const PFfloat* texels;         /* [h][w][4] RGBA; NULL: that texture's RGBA-float pixels */
    PFsizei        numVertices;
    const PFfloat* vertexColors;   /* [numVertices][4]; NULL: recorded colours, no gradient  */
    const PFfloat* lightDiffuse;   /* [8][3]; NULL: the lights' own colours                  */
    const PFfloat* lightSpecular;  /* [8][3] */
} PFdiffparams;

typedef struct {                   /* outputs; NULL members are skipped */
    PFfloat* texels;
    PFfloat* vertexColors;
    PFfloat* lightDiffuse;
    PFfloat* lightSpecular;
} PFdiffgrads;

PF_API void pfDiffRender(const PFdiffparams* params, PFfloat* image);
PF_API void pfDiffBackward(const PFdiffparams* params, const PFfloat* dImage, PFdiffgrads* grads);

#if defined(__cplusplus)
}
#endif

#endif /* PIXELFORGE_AMD_H */
