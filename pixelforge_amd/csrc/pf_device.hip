// pf_device.hip -- HIP kernels for gfx950 (MI355X) and the C ABI of
// include/pf_device.h.
//
// Pipeline for one pass of a job (a job holds one or more recorded frames;
// a pass holds at most one clear+draw segment per frame):
//   k_setup<0>       one thread per recorded triangle: projection, the
//                    clip tests, fan split, cull; counts the raster triangles
//                    and tile pairs each emits and lists the triangles that
//                    need clipping                    (triangles.c:62-116, 157-349)
//   k_clip_*<0>      the listed triangles, clipped: their counts
//   k_scan_local     exclusive scan of the counts -> record and pair slots
//                    (submission order)
//   k_setup<1>       same work with lighting; writes the raster-triangle
//   (+ k_clip_*<1>)  records and the (tile, raster triangle) pairs; the
//                    clipped triangles are written beside it on a second
//                    stream
//   radix sort       stable LSD sort of the pairs by tile, 8 bits per pass
//                    (histogram, column scan, scatter), so each tile's list
//                    stays in submission order
//   k_tile_bounds    per-tile [begin, end) of the sorted pairs
//   k_raster         one workgroup per tile; the tile's colour and depth
//                    live in VGPRs (one pixel per lane per 8x8 block),
//                    initialised by the fused clear or loaded once, updated
//                    triangle by triangle in list order, written back once
//                                                          (triangles.c:287-558)
// All arithmetic that decides coverage, depth and colour follows the
// reference's AVX2 path operation by operation; the file is compiled with
// -ffp-contract=off and uses correctly rounded division and square root, so
// it is bit-identical to the CPU oracle's restatement.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#include <math.h>
#include <vector>
#include <algorithm>
#include <type_traits>
#include <mutex>
#include <atomic>
#include <condition_variable>
#include <thread>

#include "pf_frame.h"
#include "pf_device.h"
#define PF_X86_TABLE_QUAL static __constant__ const
#include "pf_x86_approx.h"

#define PF_PI_D 3.14159265358979323846

namespace {

// ---------------------------------------------------------------------------
// Device-side tables
// ---------------------------------------------------------------------------

struct DTex {
    const void* pixels;     // device copy
    uint32_t    w, h, format, wrap, filter, pad;
};

struct DFrame {
    uint8_t*  color;        // device render target: W*H pixels of `kind`
    float*    depth;
    uint32_t  W, H;
    uint32_t  tilesX, tilesY, tileBase, kind;   // kind: PFF_TARGET_*
    uint32_t  bpp, pad[3];                       // bytes per colour pixel
};

// ---------------------------------------------------------------------------
// Render-target pixels.  A pixel is held as the word of its stored bytes
// (byte i of the pixel in bits 8i..8i+7: RGBA8 words, RGB/BGR byte triples,
// 16-bit words).  The triangle raster and pfClear go through the
// reference's SIMD accessors (pixel.h:1492-2067 setters, 2536-2913
// getters); points, lines, pfDrawPixels, rects and fog through its scalar
// ones (tgt_get_scalar / tgt_set_scalar below).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t tgt_load(const uint8_t* p, uint32_t kind, size_t i) {
    if (kind == PFF_TARGET_RGBA8) return ((const uint32_t*)p)[i];
    if (kind <= PFF_TARGET_BGR8) {
        const uint8_t* q = p + 3 * i;
        return (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16);
    }
    return ((const uint16_t*)p)[i];
}

__device__ __forceinline__ void tgt_store(uint8_t* p, uint32_t kind, size_t i, uint32_t v) {
    if (kind == PFF_TARGET_RGBA8) { ((uint32_t*)p)[i] = v; return; }
    if (kind <= PFF_TARGET_BGR8) {
        uint8_t* q = p + 3 * i;
        q[0] = (uint8_t)v;
        q[1] = (uint8_t)(v >> 8);
        q[2] = (uint8_t)(v >> 16);
        return;
    }
    ((uint16_t*)p)[i] = (uint16_t)v;
}

// The SIMD setter's encoding of RGBA8 colour c.  Channels scale by
// (x * (2^n - 1)) >> 8; the 5_5_5_1 alpha bit is a > 127.
__host__ __device__ __forceinline__ uint32_t tgt_set_simd(uint32_t kind, uint32_t c) {
    const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u, a = c >> 24;
    switch (kind) {
    case PFF_TARGET_RGB8:     return c & 0xFFFFFFu;                                  // :1590-1638
    case PFF_TARGET_BGR8:     return b | (g << 8) | (r << 16);                       // :1641-1689
    case PFF_TARGET_RGB565:   return (((r * 31u) >> 8) << 11) | (((g * 63u) >> 8) << 5) | ((b * 31u) >> 8);  // :1492-1538
    case PFF_TARGET_BGR565:   return (((b * 31u) >> 8) << 11) | (((g * 63u) >> 8) << 5) | ((r * 31u) >> 8);  // :1541-1587
    case PFF_TARGET_RGBA5551: return (((r * 31u) >> 8) << 11) | (((g * 31u) >> 8) << 6) | (((b * 31u) >> 8) << 1)
                                     | (a > 127u ? 1u : 0u);                         // :1851-1900
    case PFF_TARGET_BGRA5551: return (((b * 31u) >> 8) << 11) | (((g * 31u) >> 8) << 6) | (((r * 31u) >> 8) << 1)
                                     | (a > 127u ? 1u : 0u);                         // :1903-1952
    case PFF_TARGET_RGBA4444: return (((r * 15u) >> 8) << 12) | (((g * 15u) >> 8) << 8) | (((b * 15u) >> 8) << 4)
                                     | ((a * 15u) >> 8);                             // :1955-2005
    case PFF_TARGET_BGRA4444: return (((b * 15u) >> 8) << 12) | (((g * 15u) >> 8) << 8) | (((r * 15u) >> 8) << 4)
                                     | ((a * 15u) >> 8);                             // :2008-2058
    default:                  return c;                                              // :2061-2067
    }
}

// The SIMD getter's RGBA8 reading of stored pixel v: 16-bit fields widened
// by shifts (5_6_5) or integer products (x 8 for 5 bits, x 17 for 4 bits).
__device__ __forceinline__ uint32_t tgt_get_simd(uint32_t kind, uint32_t v) {
    switch (kind) {
    case PFF_TARGET_RGB8:     return v | 0xFF000000u;                                // :2604-2632
    case PFF_TARGET_BGR8:     return ((v >> 16) & 255u) | (v & 0xFF00u) | ((v & 255u) << 16) | 0xFF000000u;  // :2635-2663
    case PFF_TARGET_RGB565: case PFF_TARGET_BGR565: {                                // :2536-2601
        const uint32_t hi = ((v & 0xF800u) >> 11) << 3, g = ((v & 0x07E0u) >> 5) << 2, lo = (v & 0x1Fu) << 3;
        return kind == PFF_TARGET_RGB565 ? (0xFF000000u | (lo << 16) | (g << 8) | hi)
                                         : (0xFF000000u | (hi << 16) | (g << 8) | lo);
    }
    case PFF_TARGET_RGBA5551: case PFF_TARGET_BGRA5551: {                            // :2794-2849
        const uint32_t c0 = ((v >> 11) & 0x1Fu) * 8u, c1 = ((v >> 6) & 0x1Fu) * 8u, c2 = ((v >> 1) & 0x1Fu) * 8u;
        const uint32_t a = (v & 1u) * 255u;
        return kind == PFF_TARGET_RGBA5551 ? ((a << 24) | (c2 << 16) | (c1 << 8) | c0)
                                           : ((a << 24) | (c0 << 16) | (c1 << 8) | c2);
    }
    case PFF_TARGET_RGBA4444: case PFF_TARGET_BGRA4444: {                            // :2852-2907
        const uint32_t c0 = ((v >> 12) & 0xFu) * 17u, c1 = ((v >> 8) & 0xFu) * 17u, c2 = ((v >> 4) & 0xFu) * 17u;
        const uint32_t a = (v & 0xFu) * 17u;
        return kind == PFF_TARGET_RGBA4444 ? ((a << 24) | (c2 << 16) | (c1 << 8) | c0)
                                           : ((a << 24) | (c0 << 16) | (c1 << 8) | c2);
    }
    default:                  return v;                                              // :2910-2913
    }
}

// What the raster kernel needs of a tile's frame in one pass: the frame
// and its segment's clear, in one record (one scalar load per wave).
struct DTileFrame {
    DFrame   F;
    uint32_t active;        // the frame has a segment in this pass
    uint32_t clear;
    uint32_t clearColor;    // as the target stores it (tgt_set_simd)
    float    clearDepth;
    uint32_t pad[2];
};

struct DSeg {               // the part of one frame a pass renders
    uint32_t frame;
    uint32_t triFirst, triLast;   // global recorded-triangle range
    uint32_t passTriBase;         // index of triFirst within the pass
    uint32_t clear;               // 1: a clear precedes the draws
    uint32_t clearColor;
    float    clearDepth;
    uint32_t pixOp;               // 1 + index into the job's pfDrawPixels calls, 0 for none
};

// Raster triangle: what the tile kernel needs for coverage, depth and
// attribute interpolation.  w_k(x, y) = A_k*x + B_k*y + C_k (mod 2^32)
// equals the reference's incremental edge functions (triangles.c:324-349).
struct __attribute__((aligned(16))) RTri {
    int32_t  A[3], B[3], C[3];
    int32_t  xMin, yMin, xMax, yMax;
    float    wInvSum;
    float    z[3];
    uint32_t c[3];
    uint32_t meta;          // see META_* below
    uint32_t state;         // global state index
    float    uv[6];
    uint32_t tex[3];        // texture of META_TEX triangles: address bits 0-31; bits 32-47 |
                            // format << 16 | wrap << 17 | filter << 19; width | height << 16
};
static_assert(sizeof(RTri) == 128, "RTri layout");
static_assert(offsetof(RTri, xMin) == 36 && offsetof(RTri, yMax) == 48, "bbox read as dwords 9..12");
static_assert(offsetof(RTri, meta) == 80, "meta read as dword 20");

struct __attribute__((aligned(16))) RTriL {   // per-fragment lighting inputs
    float n[9];
    float p[9];
    float pad[2];
};
static_assert(sizeof(RTriL) == 80, "RTriL layout");

// Float-domain companion of a raster triangle (differentiable jobs only):
// corner k of the raster triangle is sum_j M[3k+j] * corner j of recorded
// triangle `src` (the identity unless clipping made new vertices).
struct __attribute__((aligned(16))) RTriD {
    float    M[9];
    uint32_t src;
    uint32_t face;
    uint32_t pad;
};
static_assert(sizeof(RTriD) == 48, "RTriD layout");

enum : uint32_t {
    META_BACK      = 1u << 0,
    META_3D        = 1u << 1,
    META_SMOOTH    = 1u << 2,
    META_DEPTH     = 1u << 3,
    META_BLEND     = 1u << 4,
    META_TEX       = 1u << 5,
    META_PHONG     = 1u << 6,
    META_SMALL     = 1u << 7,   // every A_k, B_k fits in 24 signed bits
    META_DFUNC_SH  = 8,     // 3 bits
    META_BMODE_SH  = 12,    // 3 bits
    META_NOWRAP    = 1u << 16,  // no int32 wrap of any w_k within 64 px of the bbox
    META_EXACT     = 1u << 17,  // NOWRAP and an unclamped (3D) bbox: covered pixels lie
                                // inside the bbox, so only x < xMax needs testing
};

// Smallest and largest true (64-bit) value of w = A*x + B*y + C over the
// pixels [x0, x1] x [y0, y1] (a linear function: extremes at corners).
__device__ __forceinline__ void edge_range(int32_t A, int32_t B, int32_t C, int64_t x0, int64_t x1, int64_t y0,
                                           int64_t y1, int64_t& lo, int64_t& hi) {
    const int64_t ax0 = (int64_t)A * x0, ax1 = (int64_t)A * x1, by0 = (int64_t)B * y0, by1 = (int64_t)B * y1;
    lo = (ax0 < ax1 ? ax0 : ax1) + (by0 < by1 ? by0 : by1) + C;
    hi = (ax0 < ax1 ? ax1 : ax0) + (by0 < by1 ? by1 : by0) + C;
}

// ---------------------------------------------------------------------------
// Numeric helpers (x86 semantics of the reference build)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t cvtt_i32(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}
__device__ __forceinline__ int32_t cvt_rne_i32(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)__builtin_rintf(f);
}
// cvt_rne_i32 for a value known to lie in int32 range (not NaN): a clamped
// colour channel times 255, a clamped texture coordinate times the size.
__device__ __forceinline__ int32_t cvt_rne_inrange(float f) { return (int32_t)__builtin_rintf(f); }
__device__ __forceinline__ uint8_t f2u8(float f) { return (uint8_t)(uint32_t)cvtt_i32(f); }
__device__ __forceinline__ float mn_ps(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx_ps(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clamp_ps(float x, float lo, float hi) { return mn_ps(mx_ps(x, lo), hi); }
__device__ __forceinline__ float sqrt_cr(float x) { return __fsqrt_rn(x); }

// _mm256_rcp_ps / _mm256_rsqrt_ps of the reference build (simd.h:1217-1245):
// table approximations measured on Intel Xeon (include/pf_x86_approx.h).
// Branch-free: the table result is formed for every input (its index is
// always in range) and the special cases replace it by selects, in
// reverse order of precedence (NaN / infinity, then zero and denormals,
// then the rest).
__device__ __forceinline__ float rcp_x86(float xf) {
    uint32_t x = __float_as_uint(xf), s = x & 0x80000000u, E = (x >> 23) & 0xffu, m = x & 0x7fffffu;
    const int oe = 253 - (int)E;
    uint32_t r = s | ((uint32_t)oe << 23) | ((uint32_t)pf_rcp_table[m >> 12] << 11);
    r = oe <= 0 ? s : r;                                        // results below the normal range
    r = E == 0 ? (s | 0x7f800000u) : r;
    r = E == 0xffu ? (m ? (x | 0x00400000u) : s) : r;
    return __uint_as_float(r);
}
__device__ __forceinline__ float rsqrt_x86(float xf) {
    uint32_t x = __float_as_uint(xf), s = x & 0x80000000u, E = (x >> 23) & 0xffu, m = x & 0x7fffffu;
    const int odd = (int)(E & 1u);
    const int oe = odd ? 126 - ((int)E - 127) / 2 : 127 - ((int)E - 126) / 2;
    uint32_t r = ((uint32_t)oe << 23) | ((uint32_t)pf_rsqrt_table[odd][m >> 13] << 11);
    r = s ? 0xffc00000u : r;
    r = E == 0 ? (s | 0x7f800000u) : r;
    r = E == 0xffu ? (m ? (x | 0x00400000u) : (s ? 0xffc00000u : 0u)) : r;
    return __uint_as_float(r);
}
__device__ __forceinline__ int ch(uint32_t c, int i) { return (int)((c >> (8 * i)) & 0xFFu); }
__device__ __forceinline__ uint8_t min255(int n) { return (uint8_t)(n | ((255 - n) >> 31)); }

// Cephes single-precision logf / expf as vectorised in Pommier's
// sse_mathfun; pow(x, e) = exp(log(x) * e) (simd.h:183-304, 1157-1169).
__device__ float cephes_logf(float x) {
    const float one = 1.0f;
    uint32_t invalid = (x <= 0.0f) ? 0xFFFFFFFFu : 0u;
    x = mx_ps(x, __uint_as_float(0x00800000u));
    int32_t e_i = (int32_t)(__float_as_uint(x) >> 23);
    x = __uint_as_float((__float_as_uint(x) & ~0x7f800000u) | __float_as_uint(0.5f));
    e_i -= 0x7f;
    float e = (float)e_i;
    e = e + one;
    bool lt = x < 0.707106781186547524f;
    float tmp = lt ? x : 0.0f;
    x = x - one;
    e = e - (lt ? one : 0.0f);
    x = x + tmp;
    float z = x * x;
    float y = 7.0376836292E-2f;
    y = y * x; y = y + -1.1514610310E-1f;
    y = y * x; y = y + 1.1676998740E-1f;
    y = y * x; y = y + -1.2420140846E-1f;
    y = y * x; y = y + 1.4249322787E-1f;
    y = y * x; y = y + -1.6668057665E-1f;
    y = y * x; y = y + 2.0000714765E-1f;
    y = y * x; y = y + -2.4999993993E-1f;
    y = y * x; y = y + 3.3333331174E-1f;
    y = y * x;
    y = y * z;
    tmp = e * -2.12194440e-4f;
    y = y + tmp;
    tmp = z * 0.5f;
    y = y - tmp;
    tmp = e * 0.693359375f;
    x = x + y;
    x = x + tmp;
    return __uint_as_float(__float_as_uint(x) | invalid);
}

__device__ float cephes_expf(float x) {
    const float one = 1.0f;
    x = mn_ps(x, 88.3762626647949f);
    x = mx_ps(x, -88.3762626647949f);
    float fx = x * 1.44269504088896341f;
    fx = fx + 0.5f;
    float tmp = floorf(fx);
    float mask = (tmp > fx) ? one : 0.0f;
    fx = tmp - mask;
    tmp = fx * 0.693359375f;
    float z = fx * -2.12194440e-4f;
    x = x - tmp;
    x = x - z;
    z = x * x;
    float y = 1.9875691500E-4f;
    y = y * x; y = y + 1.3981999507E-3f;
    y = y * x; y = y + 8.3334519073E-3f;
    y = y * x; y = y + 4.1665795894E-2f;
    y = y * x; y = y + 1.6666665459E-1f;
    y = y * x; y = y + 5.0000001201E-1f;
    y = y * z;
    y = y + x;
    y = y + one;
    int32_t n = cvtt_i32(fx) + 0x7f;
    float pow2n = __uint_as_float((uint32_t)n << 23);
    return y * pow2n;
}

__device__ __forceinline__ float pow_cephes(float b, float e) { return cephes_expf(cephes_logf(b) * e); }

// ---------------------------------------------------------------------------
// Colour operations (color.h, blend.h)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t pack_f3(float r, float g, float b) {      // color.h:110-121, vecSize 3
    uint32_t p = (uint32_t)cvt_rne_inrange(clamp_ps(r, 0.0f, 1.0f) * 255.0f);
    p |= (uint32_t)cvt_rne_inrange(clamp_ps(g, 0.0f, 1.0f) * 255.0f) << 8;
    p |= (uint32_t)cvt_rne_inrange(clamp_ps(b, 0.0f, 1.0f) * 255.0f) << 16;
    return p;
}

__device__ __forceinline__ uint32_t bary_smooth(uint32_t c1, uint32_t c2, uint32_t c3,
                                                float b1, float b2, float b3) {  // color.h:153-181
    uint32_t u1 = (uint32_t)cvt_rne_i32(b1 * 255.0f);
    uint32_t u2 = (uint32_t)cvt_rne_i32(b2 * 255.0f);
    uint32_t u3 = (uint32_t)cvt_rne_i32(b3 * 255.0f);
    uint32_t p = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t s = u1 * (uint32_t)ch(c1, i) + u2 * (uint32_t)ch(c2, i);
        s = s + u3 * (uint32_t)ch(c3, i);
        p |= ((s * 257u) >> 16) << (8 * i);
    }
    return p;
}

__device__ __forceinline__ uint32_t bary_flat(uint32_t c1, uint32_t c2, uint32_t c3,
                                              float b1, float b2, float b3) {    // color.h:183-203
    float m = mx_ps(b1, mx_ps(b2, b3));
    return ((m == b1) ? c1 : 0u) | (((m == b2) ? c2 : 0u) | ((m == b3) ? c3 : 0u));
}

// The mode is chosen once per fragment, outside the channel loop (inside
// it, it was a branch tree per channel), alpha blending first.
template <typename F>
__device__ __forceinline__ uint32_t blend_channels(uint32_t s, uint32_t d, F f) {
    uint32_t p = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) p |= (uint32_t)f(i, ch(s, i), ch(d, i)) << (8 * i);
    return p;
}
__device__ __forceinline__ uint32_t blend_op(uint32_t mode, uint32_t s, uint32_t d) {  // blend.h:137-274
    const int sa = ch(s, 3) + 1;
    if (mode == 1)
        return blend_channels(s, d, [&](int i, int sc, int dc) { return ((i == 3 ? 255 : sc) * sa + dc * (256 - sa)) >> 8; });
    switch (mode) {
    case 0: return blend_channels(s, d, [](int, int sc, int dc) { return (sc + dc) >> 1; });
    case 2: return blend_channels(s, d, [](int, int sc, int dc) { return min(sc + dc, 255); });
    case 3: return blend_channels(s, d, [](int, int sc, int dc) { return max(sc + dc, 0); });
    case 4: return blend_channels(s, d, [](int, int sc, int dc) { return (sc * dc) >> 8; });
    case 5: return blend_channels(s, d, [](int, int sc, int dc) { return min(((dc * (255 - sc)) >> 8) + sc, 255); });
    case 6: return blend_channels(s, d, [](int, int sc, int dc) { return max(sc, dc); });
    default: return blend_channels(s, d, [](int, int sc, int dc) { return min(sc, dc); });
    }
}

__device__ __forceinline__ uint32_t modulate(uint32_t t, uint32_t f) {       // blend.h:199-212
    uint32_t p = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) p |= (uint32_t)((ch(t, i) * ch(f, i)) >> 8) << (8 * i);
    return p;
}

// ---------------------------------------------------------------------------
// Texture sampling (sampler.h:202-410; pixel.h:2909-3008)
// ---------------------------------------------------------------------------

// Texel at offset off (clamped into the texture) of an RGBA8 (RGBA8 true)
// or RGBA float texture, as RGBA8.
template <bool RGBA8>
__device__ __forceinline__ uint32_t texel_at_f(const DTex& t, int32_t off) {
    int64_t n = (int64_t)t.w * t.h;
    int64_t o = off;
    o = o < 0 ? 0 : (o >= n ? n - 1 : o);
    if (RGBA8) return ((const uint32_t*)t.pixels)[o];
    const float4 p = ((const float4*)t.pixels)[o];
    int32_t r = cvt_rne_i32(p.x * 255.0f), g = cvt_rne_i32(p.y * 255.0f);
    int32_t b = cvt_rne_i32(p.z * 255.0f), a = cvt_rne_i32(p.w * 255.0f);
    return ((uint32_t)a << 24) | ((uint32_t)b << 16) | ((uint32_t)g << 8) | (uint32_t)r;
}
__device__ __forceinline__ uint32_t texel_at(const DTex& t, int32_t off) {
    return t.format == PFF_TEX_RGBA8 ? texel_at_f<true>(t, off) : texel_at_f<false>(t, off);
}

// Texel coordinate of u in [0, size) for wrap mode WRAP (0 repeat, 1
// mirrored repeat, else clamp).
template <int WRAP>
__device__ __forceinline__ int32_t wrap_coord_m(float u, uint32_t size) {
    float tw = (float)(size - 1);
    if (WRAP == 0) {
        int32_t x = cvt_rne_i32((u - truncf(u)) * tw);
        return x < 0 ? (int32_t)(0u - (uint32_t)x) : x;
    }
    if (WRAP == 1) {
        float a = fabsf(u);
        float m = a - floorf(__fdiv_rn(a, 2.0f)) * 2.0f;
        float mirror = 1.0f - (m - 1.0f);
        if (m > 1.0f) m = mirror;
        return cvt_rne_i32(m * tw + 0.5f);
    }
    return cvt_rne_inrange(clamp_ps(u, 0.0f, 1.0f) * tw + 0.5f);
}
__device__ __forceinline__ int32_t wrap_coord(uint32_t wrap, float u, uint32_t size) {
    return wrap == 0 ? wrap_coord_m<0>(u, size) : wrap == 1 ? wrap_coord_m<1>(u, size) : wrap_coord_m<2>(u, size);
}

__device__ __forceinline__ uint32_t lerp_color(uint32_t a, uint32_t b, float t) {   // color.h:137-144
    const float inv255 = 1.0f / 255.0f;
    uint32_t p = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float fa = (float)ch(a, i) * inv255;
        float fb = (float)ch(b, i) * inv255;
        float r = fa + t * (fb - fa);
        p |= (uint32_t)cvt_rne_inrange(clamp_ps(r, 0.0f, 1.0f) * 255.0f) << (8 * i);
    }
    return p;
}

// 1 / w correctly rounded for an integer 0 <= w <= 65535, as __fdiv_rn(1.0f,
// (float)w) but cheaper: the hardware reciprocal (within 1 ulp) and one
// fused Newton step.  Checked for every w >= 1 and every start within
// 1 ulp of the rounded reciprocal (tests/test_rcp_newton.py); w = 0 gives
// +inf as the division does.
__device__ __forceinline__ float rcp_int_rn(uint32_t w) {
    const float fw = (float)w;
    const float r0 = __builtin_amdgcn_rcpf(fw);
    const float e = __builtin_fmaf(-fw, r0, 1.0f);
    return w ? __builtin_fmaf(e, r0, r0) : __builtin_inff();
}

__device__ uint32_t sample_tex(const DTex& t, float u, float v) {
    uint32_t w = t.w;
    if (t.filter == 0) {
        int32_t x, y;
        if (t.wrap == 0) { x = wrap_coord_m<0>(u, t.w); y = wrap_coord_m<0>(v, t.h); }
        else if (t.wrap == 1) { x = wrap_coord_m<1>(u, t.w); y = wrap_coord_m<1>(v, t.h); }
        else { x = wrap_coord_m<2>(u, t.w); y = wrap_coord_m<2>(v, t.h); }
        return texel_at(t, (int32_t)((uint32_t)y * w + (uint32_t)x));
    }
    float tx = rcp_int_rn(t.w), ty = rcp_int_rn(t.h);   // texture sides are at most 65535
    // The wrap mode and the texel format are chosen once for the four
    // taps (per tap they were a branch tree each).
    int32_t x0, y0, x1, y1;
    const auto coords = [&](auto m) {
        constexpr int M = decltype(m)::value;
        x0 = wrap_coord_m<M>(u, t.w);
        y0 = wrap_coord_m<M>(v, t.h);
        x1 = wrap_coord_m<M>(u + tx, t.w);
        y1 = wrap_coord_m<M>(v + ty, t.h);
    };
    if (t.wrap == 0) coords(std::integral_constant<int, 0>());
    else if (t.wrap == 1) coords(std::integral_constant<int, 1>());
    else coords(std::integral_constant<int, 2>());
    float fx = clamp_ps(u * (float)t.w - (float)x0, 0.0f, 1.0f);
    float fy = clamp_ps(v * (float)t.h - (float)y0, 0.0f, 1.0f);
    const int32_t o00 = (int32_t)((uint32_t)y0 * w + (uint32_t)x0), o10 = (int32_t)((uint32_t)y0 * w + (uint32_t)x1);
    const int32_t o01 = (int32_t)((uint32_t)y1 * w + (uint32_t)x0), o11 = (int32_t)((uint32_t)y1 * w + (uint32_t)x1);
    uint32_t c00, c10, c01, c11;
    if (t.format == PFF_TEX_RGBA8) {
        c00 = texel_at_f<true>(t, o00); c10 = texel_at_f<true>(t, o10);
        c01 = texel_at_f<true>(t, o01); c11 = texel_at_f<true>(t, o11);
    } else {
        c00 = texel_at_f<false>(t, o00); c10 = texel_at_f<false>(t, o10);
        c01 = texel_at_f<false>(t, o01); c11 = texel_at_f<false>(t, o11);
    }
    uint32_t c0 = lerp_color(c00, c10, fx);
    uint32_t c1 = lerp_color(c01, c11, fx);
    return lerp_color(c0, c1, fy);
}

// ---------------------------------------------------------------------------
// Lighting
// ---------------------------------------------------------------------------

__device__ __forceinline__ float dot3(const float* a, const float* b) {      // pfm.h:1227
    float d = 0.0f;
    d += a[0] * b[0];
    d += a[1] * b[1];
    d += a[2] * b[2];
    return d;
}

// pfiLightingProcess (lighting.c:23-144), per vertex.
__device__ uint32_t light_vertex(const PFFstate& st, const PFFmaterial& mat, uint32_t diffuse,
                                 const float* viewPos, const float* fragPos, const float* N) {
    uint8_t R = (uint8_t)ch(mat.emission, 0), G = (uint8_t)ch(mat.emission, 1), B = (uint8_t)ch(mat.emission, 2);
    uint8_t aC[3];
    for (int i = 0; i < 3; i++) aC[i] = (uint8_t)((ch(mat.ambient, i) * ch(diffuse, i)) / 255);
    float V[3];
    float lsq = 0.0f;
    for (int i = 0; i < 3; i++) { V[i] = viewPos[i] - fragPos[i]; lsq += V[i] * V[i]; }
    float inv = __fdiv_rn(1.0f, sqrt_cr(lsq));
    for (int i = 0; i < 3; i++) V[i] = V[i] * inv;
    for (uint32_t li = 0; li < st.numLights; li++) {
        const PFFlight& Lt = st.lights[li];
        uint8_t lC[3] = {0, 0, 0};
        float L[3];
        for (int i = 0; i < 3; i++) L[i] = Lt.position[i] - fragPos[i];
        float distSq = L[0] * L[0] + L[1] * L[1] + L[2] * L[2];
        float dist = 0.0f;
        uint8_t intensity = 255, att = 255;
        if (distSq != 0.0f) {
            dist = sqrt_cr(distSq);
            float invMag = __fdiv_rn(1.0f, dist);
            for (int i = 0; i < 3; i++) L[i] *= invMag;
        }
        bool skip = false;
        if ((double)Lt.innerCutOff < PF_PI_D) {
            float nd[3] = { -Lt.direction[0], -Lt.direction[1], -Lt.direction[2] };
            float theta = dot3(L, nd);
            float eps = Lt.innerCutOff - Lt.outerCutOff;
            int32_t iv = cvtt_i32(__fdiv_rn(255.0f * (theta - Lt.outerCutOff), eps));
            intensity = (uint8_t)(iv < 0 ? 0 : (iv > 255 ? 255 : iv));
            if (intensity == 0) skip = true;
        }
        if (!skip && (Lt.attLinear != 0.0f || Lt.attQuadratic != 0.0f)) {
            att = f2u8(__fdiv_rn(255.0f, Lt.attConstant + Lt.attLinear * dist + Lt.attQuadratic * distSq));
            if (att == 0) skip = true;
        }
        if (!skip) {
            uint8_t factor = (uint8_t)((intensity * att) / 255);
            int32_t dv = cvtt_i32(255.0f * dot3(N, L));
            uint8_t diff = (uint8_t)(dv > 0 ? dv : 0);
            for (int i = 0; i < 3; i++)
                lC[i] = min255(lC[i] + (ch(diffuse, i) * ch(Lt.diffuse, i) * diff) / (255 * 255));
            float H[3];
            for (int i = 0; i < 3; i++) H[i] = L[i] + V[i];
            float hs = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
            if (hs != 0.0f) {
                float hi = __fdiv_rn(1.0f, sqrt_cr(hs));
                for (int i = 0; i < 3; i++) H[i] = H[i] * hi;
            }
            float nh = fmaxf(dot3(N, H), 0.0f);
            uint8_t spec = f2u8(255.0f * (float)pow((double)nh, (double)mat.shininess));
            for (int i = 0; i < 3; i++)
                lC[i] = min255(lC[i] + (ch(mat.specular, i) * ch(Lt.specular, i) * spec) / (255 * 255));
            for (int i = 0; i < 3; i++) lC[i] = (uint8_t)((lC[i] * factor) / 255);
        }
        R = min255(R + lC[0] + (aC[0] * ch(Lt.ambient, 0)) / 255);
        G = min255(G + lC[1] + (aC[1] * ch(Lt.ambient, 1)) / 255);
        B = min255(B + lC[2] + (aC[2] * ch(Lt.ambient, 2)) / 255);
    }
    return (uint32_t)R | ((uint32_t)G << 8) | ((uint32_t)B << 16) | (diffuse & 0xFF000000u);
}

__device__ __forceinline__ void direction_simd(float* d, const float* a, const float* b) {  // simd.h:2469-2490
    for (int i = 0; i < 3; i++) d[i] = a[i] - b[i];
    float l = d[0] * d[0] + d[1] * d[1];
    l = l + d[2] * d[2];
    l = mx_ps(l, 1e-5f);
    float inv = rsqrt_x86(l);
    for (int i = 0; i < 3; i++) d[i] = d[i] * inv;
}

__device__ __forceinline__ float dot3_simd(const float* a, const float* b) {  // simd.h:2356-2363
    float d = a[0] * b[0] + a[1] * b[1];
    return d + a[2] * b[2];
}

// pfiSimdLightingProcess (lighting.c:148-258), per fragment.
__device__ uint32_t light_fragment(const PFFstate* __restrict__ st, const PFFmaterial& mat, uint32_t frag,
                                   const float* P, const float* N) {
    const float inv255 = 1.0f / 255.0f;
    float D[3], A[3], S[3];
    for (int i = 0; i < 3; i++) {
        D[i] = (float)ch(frag, i) * inv255;
        A[i] = (float)ch(mat.ambient, i) * inv255;
        S[i] = (float)ch(mat.specular, i) * inv255;
        A[i] = A[i] * D[i];
    }
    float V[3];
    direction_simd(V, st->viewPos, P);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t nl = st->numLights;
    for (uint32_t li = 0; li < nl; li++) {
        const PFFlight& Lt = st->lights[li];
        float la[3], ld[3], ls[3];
        for (int i = 0; i < 3; i++) {
            la[i] = (float)ch(Lt.ambient, i) * inv255;
            ld[i] = (float)ch(Lt.diffuse, i) * inv255;
            ls[i] = (float)ch(Lt.specular, i) * inv255;
        }
        float L[3];
        direction_simd(L, Lt.position, P);
        float amb[3], dif[3], spc[3];
        for (int i = 0; i < 3; i++) amb[i] = la[i] * A[i];
        float diff = mx_ps(dot3_simd(N, L), 0.0f);
        for (int i = 0; i < 3; i++) { dif[i] = ld[i] * diff; dif[i] = dif[i] * D[i]; }
        float H[3];
        for (int i = 0; i < 3; i++) H[i] = L[i] + V[i];
        float hs = H[0] * H[0] + H[1] * H[1];
        hs = hs + H[2] * H[2];
        float hinv = rsqrt_x86(hs);
        for (int i = 0; i < 3; i++) H[i] = H[i] * hinv;
        float spec = mx_ps(dot3_simd(N, H), 0.0f);
        spec = pow_cephes(spec, mat.shininess);
        for (int i = 0; i < 3; i++) { spc[i] = ls[i] * spec; spc[i] = spc[i] * S[i]; }
        if ((double)Lt.innerCutOff < PF_PI_D) {
            float nd[3] = { 0.0f - Lt.direction[0], 0.0f - Lt.direction[1], 0.0f - Lt.direction[2] };
            float theta = dot3_simd(L, nd);
            float eps = Lt.innerCutOff - Lt.outerCutOff;
            float I = clamp_ps(__fdiv_rn(theta - Lt.outerCutOff, eps), 0.0f, 1.0f);
            for (int i = 0; i < 3; i++) { dif[i] = dif[i] * I; spc[i] = spc[i] * I; }
        }
        if (Lt.attLinear != 0.0f || Lt.attQuadratic != 0.0f) {
            float d0 = Lt.position[0] - P[0];
            float d1 = Lt.position[1] - P[1];
            float d2 = Lt.position[1] - P[1];          // simd.h:2429 reads y twice
            float dsq = d0 * d0 + (d1 * d1 + d2 * d2);
            float dist = sqrt_cr(dsq);
            float att = rcp_x86(Lt.attConstant + (Lt.attLinear * dist + Lt.attQuadratic * dsq));
            for (int i = 0; i < 3; i++) { amb[i] *= att; dif[i] *= att; spc[i] *= att; }
        }
        for (int i = 0; i < 3; i++) {
            acc[i] = acc[i] + amb[i];
            acc[i] = acc[i] + dif[i];
            acc[i] = acc[i] + spc[i];
        }
    }
    return pack_f3(acc[0], acc[1], acc[2]);
}

// ---------------------------------------------------------------------------
// Setup: projection, clipping, fan split, cull (triangles.c:62-349)
// ---------------------------------------------------------------------------

struct Vtx {
    float hom[4];
    float scr[2];
    float pos[4];
    float nrm[3];
    float uv[2];
    uint32_t col;
    float bw[3];      // weights of the recorded triangle's corners (float domain)
};

__device__ __forceinline__ Vtx lerp_vtx(const Vtx& s, const Vtx& e, float t) {   // ictx.c:67-90
    Vtx r;
    uint8_t uT = f2u8(255.0f * t);
    uint32_t col = 0;
    for (int i = 0; i < 4; i++) {
        r.hom[i] = s.hom[i] + t * (e.hom[i] - s.hom[i]);
        r.pos[i] = s.pos[i] + t * (e.pos[i] - s.pos[i]);
        int cs = ch(s.col, i), ce = ch(e.col, i);
        col |= (uint32_t)(uint8_t)(cs + ((int)uT * (ce - cs)) / 255) << (8 * i);
    }
    for (int i = 0; i < 2; i++) r.uv[i] = s.uv[i] + t * (e.uv[i] - s.uv[i]);
    for (int i = 0; i < 3; i++) r.nrm[i] = s.nrm[i] + t * (e.nrm[i] - s.nrm[i]);
    for (int i = 0; i < 3; i++) r.bw[i] = s.bw[i] + t * (e.bw[i] - s.bw[i]);
    r.scr[0] = r.scr[1] = 0.0f;
    r.col = col;
    return r;
}

// pfiHomogeneousToScreen (ictx.c:51-65); vp = {vpPos.x, vpPos.y, vpDim.x, vpDim.y} as floats.
__device__ __forceinline__ void to_screen(const float* vp, Vtx& v) {
    v.scr[0] = (vp[0] + ((v.hom[0] + 1.0f) * 0.5f) * vp[2]) + 0.5f;
    v.scr[1] = (vp[1] + ((1.0f - v.hom[1]) * 0.5f) * vp[3]) + 0.5f;
}

// Forms of a clip vertex in LDS, by what the rest of the setup reads:
// VF_POS the clip position alone (the counting pass), VF_SLIM position, uv
// and colour (no Phong record, no corner weights), VF_FULL everything.  The
// clip position comes first in each form.  Attributes a form leaves out
// load as zero; screen coordinates are recomputed from the position.
constexpr int VF_POS = 0, VF_SLIM = 1, VF_FULL = 2;
template <int VF> struct CVtx;
template <> struct CVtx<VF_POS>  { float hom[4]; };
template <> struct CVtx<VF_SLIM> { float hom[4]; float uv[2]; uint32_t col; };
template <> struct CVtx<VF_FULL> { Vtx v; };

__device__ __forceinline__ const float* cv_hom(const CVtx<VF_FULL>& c) { return c.v.hom; }
template <int VF> __device__ __forceinline__ const float* cv_hom(const CVtx<VF>& c) { return c.hom; }

__device__ __forceinline__ Vtx cv_load(const CVtx<VF_FULL>& c) { return c.v; }
__device__ __forceinline__ void cv_store(CVtx<VF_FULL>& c, const Vtx& v) { c.v = v; }
template <int VF>
__device__ __forceinline__ Vtx cv_load(const CVtx<VF>& c) {
    Vtx v;
    for (int i = 0; i < 4; i++) { v.hom[i] = c.hom[i]; v.pos[i] = 0.0f; }
    for (int i = 0; i < 3; i++) { v.nrm[i] = 0.0f; v.bw[i] = 0.0f; }
    v.scr[0] = v.scr[1] = 0.0f;
    if constexpr (VF == VF_SLIM) { v.uv[0] = c.uv[0]; v.uv[1] = c.uv[1]; v.col = c.col; }
    else { v.uv[0] = v.uv[1] = 0.0f; v.col = 0u; }
    return v;
}
template <int VF>
__device__ __forceinline__ void cv_store(CVtx<VF>& c, const Vtx& v) {
    for (int i = 0; i < 4; i++) c.hom[i] = v.hom[i];
    if constexpr (VF == VF_SLIM) { c.uv[0] = v.uv[0]; c.uv[1] = v.uv[1]; c.col = v.col; }
}

// Sutherland-Hodgman against w >= eps then +-x, +-y, +-z <= w
// (triangles.c:157-244), in the caller's two LDS buffers of 12 vertices
// (ping-pong; the result is returned through *out).  A plane with every
// vertex inside leaves the polygon as it is, so it is skipped: the result
// is the same, without the copy the reference makes for it.
template <int VF>
__device__ __forceinline__ int clip_polygon(CVtx<VF>* buf0, CVtx<VF>* buf1, int n, CVtx<VF>** out) {
    CVtx<VF>* in = buf0;
    CVtx<VF>* o = buf1;
    for (int pl = 0; pl < 7; pl++) {
        const int ax = (pl - 1) >> 1, side = (pl - 1) & 1;
        bool all = true;
        for (int i = 0; i < n; i++) {
            const float w = cv_hom(in[i])[3];
            const bool ins = (pl == 0) ? !(w < 1e-5f) : ((side ? -cv_hom(in[i])[ax] : cv_hom(in[i])[ax]) <= w);
            all = all && ins;
        }
        if (all) continue;
        int m = 0;
        int prev = n - 1;
        if (pl == 0) {
            int pd = (cv_hom(in[prev])[3] < 1e-5f) ? -1 : 1;
            for (int i = 0; i < n; i++) {
                int cd = (cv_hom(in[i])[3] < 1e-5f) ? -1 : 1;
                if (pd * cd < 0 && m < 12)
                    cv_store(o[m++], lerp_vtx(cv_load(in[prev]), cv_load(in[i]), __fdiv_rn(1e-5f - cv_hom(in[prev])[3], cv_hom(in[i])[3] - cv_hom(in[prev])[3])));
                if (cd > 0 && m < 12) o[m++] = in[i];
                pd = cd;
                prev = i;
            }
        } else {
            float pa = side ? -cv_hom(in[prev])[ax] : cv_hom(in[prev])[ax];
            int pd = (pa <= cv_hom(in[prev])[3]) ? 1 : -1;
            for (int i = 0; i < n; i++) {
                float ca = side ? -cv_hom(in[i])[ax] : cv_hom(in[i])[ax];
                int cd = (ca <= cv_hom(in[i])[3]) ? 1 : -1;
                if (pd * cd <= 0 && m < 12) {
                    float num, den;
                    if (!side) { num = cv_hom(in[prev])[3] - cv_hom(in[prev])[ax]; den = num - (cv_hom(in[i])[3] - cv_hom(in[i])[ax]); }
                    else       { num = cv_hom(in[prev])[3] + cv_hom(in[prev])[ax]; den = num - (cv_hom(in[i])[3] + cv_hom(in[i])[ax]); }
                    cv_store(o[m++], lerp_vtx(cv_load(in[prev]), cv_load(in[i]), __fdiv_rn(num, den)));
                }
                if (cd > 0 && m < 12) o[m++] = in[i];
                pd = cd;
                prev = i;
            }
        }
        n = m;
        CVtx<VF>* t = in; in = o; o = t;
        if (n == 0) break;
    }
    *out = in;
    return n;
}

// Kernel-argument pointers are read through the constant address space
// when the index is wavefront-uniform, so that the loads become scalar.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}
typedef const __attribute__((address_space(4))) PFFstate CState;
// Four words loaded as one vector (a uint4 struct may be split into
// narrower loads).
typedef uint32_t Words4 __attribute__((ext_vector_type(4)));
// emit_raster reads flags .. numLights as two aligned 16-byte vectors.
static_assert(sizeof(PFFstate) % 16 == 0 && offsetof(PFFstate, flags) % 16 == 0, "state raster fields alignment");
static_assert(offsetof(PFFstate, depthFunc) == offsetof(PFFstate, flags) + 4 &&
              offsetof(PFFstate, blendMode) == offsetof(PFFstate, flags) + 8 &&
              offsetof(PFFstate, shadeModel) == offsetof(PFFstate, flags) + 12 &&
              offsetof(PFFstate, lightMode) == offsetof(PFFstate, flags) + 16 &&
              offsetof(PFFstate, texture) == offsetof(PFFstate, flags) + 20 &&
              offsetof(PFFstate, numLights) == offsetof(PFFstate, flags) + 24, "state raster fields order");

struct SetupArgs {
    const float4*   pos;
    const uint32_t* col;
    const float2*   uv;
    const float*    nrm;
    const uint32_t* info;
    const PFFstate* states;
    const DTex*     texs;
    const DSeg*     segs;
    const DFrame*   frames;
    uint32_t        numSegs;
    const uint32_t* blockSeg;       // per 256-thread block: segment of its first thread
    uint32_t        numPassTris;
    uint32_t        tileShift;
    uint32_t        numTiles;
    uint32_t*       cntRT;          // [numPassTris] raster triangles per recorded triangle
    uint32_t*       cntPairs;       // [numPassTris] tile pairs per recorded triangle
    uint32_t*       blockRT;        // per block: sum (count pass), then exclusive scan
    uint32_t*       blockPairs;
    uint32_t*       tileBegin;      // zeroed by the count pass
    uint32_t*       tileEnd;
    RTri*           rts;
    RTriL*          rtl;
    RTriD*          rtd;            // float-domain records, or null
    uint32_t*       keys;           // tile of each pair
    uint32_t*       vals;           // raster triangle of each pair
    uint32_t        capRT;
    uint32_t        capPairs;
    uint4*          clipList;       // triangles to clip {gid, t, info, frame} (PHASE 0 appends)
    uint32_t*       clipCount;
    uint32_t        clipCap;        // entries clipList holds
    uint32_t        clipKnown;      // listed triangles when the pass is sized (the clip
                                    // kernels then skip reading clipCount), else ~0u
    uint32_t        clipBlocks;     // k_setup<1, .., CLIPW>: leading workgroups that clip
    uint32_t*       clearSnap;      // per frame, 8 words: colour word and depth bits of pixel 0 when the
                                    // pass began, then (wide targets) its bytes
    uint32_t*       pairOff;        // non-null: PHASE 1 writes, per raster triangle, its first pair
    uint4*          pairRect;       // slot and {tile0, w | h << 16, tilesX}; k_expand_pairs writes
                                    // the pairs
};

// Tiles [tx0..tx1] x [ty0..ty1] overlapped by the covered pixels of a
// triangle: columns [xMin, xMax) and rows [yMin, yMax] clipped to the target.
struct TileSpan { uint32_t tx0, tx1, ty0, ty1; };

// The tile rectangle of one raster triangle's pairs: w x h tiles from tile0
// (rows tilesX apart), kept by the clipper so that the wavefront writes the
// pairs of all its lanes' triangles together.
struct PairRect {
    uint32_t tile0, wh;                 // wh = w | h << 16
    uint32_t q;                         // quadrant bounds (pair_quads), 0: none
};

// A tile pair's value: its raster triangle, and in the top 4 bits the 8x8
// quadrants of its 16x16 tile that the triangle's covered-pixel bounding
// box meets (bit 2 * qy + qx; all four at other tile sizes).  The raster
// kernel's blocks at 16x16 tiles are those quadrants, and a block's bbox
// test passes exactly for the pairs whose bit it holds, so a wave skips
// the others before fetching their records.
constexpr uint32_t PAIR_RT_MASK = 0x0FFFFFFFu;
constexpr uint32_t PAIR_QUAD_SH = 28;

// The quadrant bounds of a pair rectangle whose covered pixels span
// [cx0, cx1] x [cy0, cy1] at 16x16 tiles (their positions inside the first
// and last tile of each axis); 0 at other tile sizes.
__device__ __forceinline__ uint32_t pair_quads(uint32_t ts, int32_t cx0, int32_t cx1, int32_t cy0, int32_t cy1) {
    if (ts != 4u) return 0u;
    return 0x10000u | ((uint32_t)cx0 & 15u) | (((uint32_t)cx1 & 15u) << 4) | (((uint32_t)cy0 & 15u) << 8)
         | (((uint32_t)cy1 & 15u) << 12);
}

// The value of the pair at column i (of w) and row j (of h) of a pair
// rectangle with quadrant bounds q.
__device__ __forceinline__ uint32_t pair_value(uint32_t rt, uint32_t i, uint32_t w, uint32_t j, uint32_t h, uint32_t q) {
    uint32_t m = 15u;
    if (q) {
        const uint32_t x0 = i == 0u ? (q & 15u) : 0u, x1 = i + 1u == w ? ((q >> 4) & 15u) : 15u;
        const uint32_t y0 = j == 0u ? ((q >> 8) & 15u) : 0u, y1 = j + 1u == h ? ((q >> 12) & 15u) : 15u;
        const uint32_t cols = (x0 < 8u ? 1u : 0u) | (x1 >= 8u ? 2u : 0u);
        const uint32_t rows = (y0 < 8u ? 1u : 0u) | (y1 >= 8u ? 2u : 0u);
        m = ((rows & 1u) ? cols : 0u) | ((rows & 2u) ? cols << 2 : 0u);
    }
    return rt | (m << PAIR_QUAD_SH);
}
constexpr int MAX_FAN_RECTS = 10;       // fan triangles of a polygon of at most 12 vertices
// The pair rectangles of the (at most two) raster triangles one lane of the
// setup emits for an unclipped triangle, in emission order.
struct LaneRects {
    PairRect a, b;
    // Field by field: a store through a pointer chosen at run time would
    // keep the pair in scratch memory.
    __device__ __forceinline__ void put(bool first, const PairRect& r) {
        a.tile0 = first ? r.tile0 : a.tile0; a.wh = first ? r.wh : a.wh; a.q = first ? r.q : a.q;
        b.tile0 = first ? b.tile0 : r.tile0; b.wh = first ? b.wh : r.wh; b.q = first ? b.q : r.q;
    }
};

// Builds (or only counts) the raster triangle of screen vertices a, b, c for
// one culling pass (Rasterize_Triangle setup, triangles.c:287-349) and its
// tile pairs.  Returns false when it is culled or covers no pixel.  EXTRA
// false: the job has no lit state and no float-domain records (no Phong
// record, no RTriD), so the corners' positions, normals and weights are
// not needed.
template <int PHASE, bool EXTRA = true, bool PRELOAD = false>
__device__ __forceinline__ bool emit_raster(const SetupArgs& A, CState* st, const DFrame& F,
                                            uint32_t stateIdx, uint32_t src, int face, bool is3D,
                                            const Vtx& a, const Vtx& b, const Vtx& c,
                                            uint32_t rtSlot, uint32_t pairSlot, uint32_t& pairsOut,
                                            PairRect* rect = nullptr) {
    int32_t x1 = cvtt_i32(a.scr[0]), y1 = cvtt_i32(a.scr[1]);
    int32_t x2 = cvtt_i32(b.scr[0]), y2 = cvtt_i32(b.scr[1]);
    int32_t x3 = cvtt_i32(c.scr[0]), y3 = cvtt_i32(c.scr[1]);
    uint32_t ux1 = x1, uy1 = y1, ux2 = x2, uy2 = y2, ux3 = x3, uy3 = y3;
    float area = (float)(int32_t)((ux2 - ux1) * (uy3 - uy1) - (ux3 - ux1) * (uy2 - uy1));
    if ((face == 0 && area >= 0.0f) || (face == 1 && area <= 0.0f)) return false;
    int32_t xMin = min(x1, min(x2, x3)), yMin = min(y1, min(y2, y3));
    int32_t xMax = max(x1, max(x2, x3)), yMax = max(y1, max(y2, y3));
    if (!is3D) {
        xMin = min(max(xMin, st->vpMin[0]), st->vpMax[0]);
        yMin = min(max(yMin, st->vpMin[1]), st->vpMax[1]);
        xMax = min(max(xMax, st->vpMin[0]), st->vpMax[0]);
        yMax = min(max(yMax, st->vpMin[1]), st->vpMax[1]);
    }
    // Unsigned loops of the reference (triangles.c:404, 409) run nothing
    // when the minimum is negative and the maximum is not.
    if ((uint32_t)xMin > (uint32_t)xMax || (uint32_t)yMin > (uint32_t)yMax) return false;
    // Pixels outside the render target are never written.
    int32_t cx0 = max(xMin, 0), cy0 = max(yMin, 0);
    int32_t cx1 = min(xMax - 1, (int32_t)F.W - 1), cy1 = min(yMax, (int32_t)F.H - 1);
    if (cx0 > cx1 || cy0 > cy1) return false;
    const uint32_t ts = A.tileShift;
    const uint32_t tx0 = (uint32_t)cx0 >> ts, tx1 = (uint32_t)cx1 >> ts;
    const uint32_t ty0 = (uint32_t)cy0 >> ts, ty1 = (uint32_t)cy1 >> ts;

    uint32_t Ax[3] = { uy3 - uy2, uy1 - uy3, uy2 - uy1 };
    uint32_t By[3] = { ux2 - ux3, ux3 - ux1, ux1 - ux2 };
    if (face == 1) for (int k = 0; k < 3; k++) { Ax[k] = 0u - Ax[k]; By[k] = 0u - By[k]; }
    uint32_t ox[3] = { ux2, ux3, ux1 }, oy[3] = { uy2, uy3, uy1 };
    // Tile pairs: every tile of the clipped bbox (the raster kernel rejects
    // the blocks an edge leaves uncovered; testing tiles here measured
    // slower, the per-lane loops over long thin triangles' spans dominate).
    pairsOut = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    if (PHASE == 0) return true;
    if (rtSlot >= A.capRT) return false;
    // PRELOAD (the clip kernels, whose lanes hold different states): the
    // state's raster fields, flags .. numLights, in two 16-byte loads issued
    // here; read one by one where they are tested, they made a chain of
    // dependent vector loads.  With a state uniform over the wavefront they
    // are scalar loads and are read where used (fewer live registers).
    typedef const __attribute__((address_space(4))) Words4 CW4;
    Words4 sf0 = { 0u, 0u, 0u, 0u }, sf1 = { 0u, 0u, 0u, 0u };
    if (PRELOAD) {
        sf0 = reinterpret_cast<CW4*>(&st->flags)[0];                 // flags, depthFunc, blendMode, shadeModel
        sf1 = reinterpret_cast<CW4*>(&st->flags)[1];                 // lightMode, texture, numLights
    }
    uint32_t Cx[3];
    for (int k = 0; k < 3; k++) Cx[k] = 0u - (ox[k] * Ax[k] + oy[k] * By[k]);
    // No int32 wrap of any edge within 64 px of the bbox (then a block of a
    // tile of the span has its largest edge value at a corner).
    bool nowrap = true;
    for (int k = 0; k < 3; k++) {
        int64_t lo, hi;
        edge_range((int32_t)Ax[k], (int32_t)By[k], (int32_t)Cx[k], (int64_t)xMin - 64, (int64_t)xMax + 64,
                   (int64_t)yMin - 64, (int64_t)yMax + 64, lo, hi);
        nowrap = nowrap && lo >= INT32_MIN && hi <= INT32_MAX;
    }

    RTri r;
    uint32_t sum = 0;
    for (int k = 0; k < 3; k++) {
        r.A[k] = (int32_t)Ax[k];
        r.B[k] = (int32_t)By[k];
        r.C[k] = (int32_t)Cx[k];
        sum += Ax[k] * (uint32_t)xMin + By[k] * (uint32_t)yMin + Cx[k];
    }
    r.xMin = xMin; r.yMin = yMin; r.xMax = xMax; r.yMax = yMax;
    r.wInvSum = __fdiv_rn(1.0f, (float)(int32_t)sum);
    r.z[0] = a.hom[2]; r.z[1] = b.hom[2]; r.z[2] = c.hom[2];
    r.c[0] = a.col; r.c[1] = b.col; r.c[2] = c.col;
    uint32_t meta = (face ? META_BACK : 0u) | (is3D ? META_3D : 0u);
    bool small = true;
    for (int k = 0; k < 3; k++) {
        int32_t va = (int32_t)Ax[k], vb = (int32_t)By[k];
        small = small && va > -(1 << 23) && va < (1 << 23) && vb > -(1 << 23) && vb < (1 << 23);
    }
    if (small) meta |= META_SMALL;
    if (nowrap) meta |= META_NOWRAP | (is3D ? META_EXACT : 0u);
    const uint32_t flags = PRELOAD ? sf0.x : st->flags;
    const int32_t texIdx = PRELOAD ? (int32_t)sf1.y : st->texture;
    if ((PRELOAD ? sf0.w : st->shadeModel) == 1) meta |= META_SMOOTH;
    if (flags & PFF_DEPTH_TEST) meta |= META_DEPTH | (((PRELOAD ? sf0.y : st->depthFunc) & 7u) << META_DFUNC_SH);
    if (flags & PFF_BLEND) meta |= META_BLEND | (((PRELOAD ? sf0.z : st->blendMode) & 7u) << META_BMODE_SH);
    if ((flags & PFF_TEXTURE_2D) && texIdx >= 0) meta |= META_TEX;
    const bool phong = EXTRA && (flags & PFF_LIGHTING) && (PRELOAD ? sf1.x : st->lightMode) == 1 &&
                       (PRELOAD ? sf1.z : st->numLights) > 0;
    if (phong) meta |= META_PHONG;
    r.meta = meta;
    r.state = stateIdx;
    r.uv[0] = a.uv[0]; r.uv[1] = a.uv[1]; r.uv[2] = b.uv[0]; r.uv[3] = b.uv[1]; r.uv[4] = c.uv[0]; r.uv[5] = c.uv[1];
    r.tex[0] = r.tex[1] = r.tex[2] = 0;
    if (meta & META_TEX) {
        const DTex T = A.texs[texIdx];
        const uint64_t addr = (uint64_t)(uintptr_t)T.pixels;
        r.tex[0] = (uint32_t)addr;
        r.tex[1] = ((uint32_t)(addr >> 32) & 0xFFFFu) | ((T.format & 1u) << 16) | ((T.wrap & 3u) << 17) | ((T.filter & 1u) << 19);
        r.tex[2] = (T.w & 0xFFFFu) | (T.h << 16);
    }
    A.rts[rtSlot] = r;
    if (phong) {
        RTriL l;
        for (int i = 0; i < 3; i++) {
            l.n[i] = a.nrm[i]; l.n[3 + i] = b.nrm[i]; l.n[6 + i] = c.nrm[i];
            l.p[i] = a.pos[i]; l.p[3 + i] = b.pos[i]; l.p[6 + i] = c.pos[i];
        }
        l.pad[0] = l.pad[1] = 0.0f;
        A.rtl[rtSlot] = l;
    }
    if (EXTRA && A.rtd) {
        RTriD d;
        for (int j = 0; j < 3; j++) { d.M[j] = a.bw[j]; d.M[3 + j] = b.bw[j]; d.M[6 + j] = c.bw[j]; }
        d.src = src;
        d.face = (uint32_t)face;
        d.pad = 0;
        A.rtd[rtSlot] = d;
    }
    // Tile pairs of this raster triangle, in submission order: handed back
    // as a rectangle (clipped triangles: large, few) or written here (the
    // rest: mostly a few pairs, where the per-lane loop is cheaper).
    const uint32_t quads = pair_quads(ts, cx0, cx1, cy0, cy1);
    const uint32_t w = tx1 - tx0 + 1, h = ty1 - ty0 + 1;
    if (A.pairOff) {
        // Many pairs per raster triangle: k_expand_pairs writes them.
        A.pairOff[rtSlot] = pairSlot;
        A.pairRect[rtSlot] = make_uint4(F.tileBase + ty0 * F.tilesX + tx0, w | (h << 16), F.tilesX, quads);
        if (rect) *rect = PairRect{ 0u, 0u, 0u };
        return true;
    }
    if (rect) {
        *rect = PairRect{ F.tileBase + ty0 * F.tilesX + tx0, w | (h << 16), quads };
        return true;
    }
    uint32_t p = pairSlot;
    for (uint32_t ty = ty0; ty <= ty1; ty++)
        for (uint32_t tx = tx0; tx <= tx1; tx++) {
            if (p >= A.capPairs) return true;
            A.keys[p] = F.tileBase + ty * F.tilesX + tx;
            A.vals[p] = pair_value(rtSlot, tx - tx0, w, ty - ty0, h, quads);
            p++;
        }
    return true;
}

// Clip-space position (pfmVec4Transform, pfm.h:1996; no contraction).
__device__ __forceinline__ void transform_pos(CState* st, float4 p, float* h) {
    h[0] = st->mvp[0] * p.x + st->mvp[4] * p.y + st->mvp[8] * p.z + st->mvp[12] * p.w;
    h[1] = st->mvp[1] * p.x + st->mvp[5] * p.y + st->mvp[9] * p.z + st->mvp[13] * p.w;
    h[2] = st->mvp[2] * p.x + st->mvp[6] * p.y + st->mvp[10] * p.z + st->mvp[14] * p.w;
    h[3] = st->mvp[3] * p.x + st->mvp[7] * p.y + st->mvp[11] * p.z + st->mvp[15] * p.w;
}

// Corner k of recorded triangle t for one culling pass: position, clip
// position and (PHASE 1) colour, texture coordinate and normal, lit when
// lighting is on (Process_ProjectAndClipTriangle's per-vertex part,
// triangles.c:62-104).
template <int PHASE, bool LIT, bool EXTRA = true>
__device__ __forceinline__ Vtx corner_vertex(const SetupArgs& A, CState* st, uint32_t si, int face, uint32_t t, int k,
                                             float4 p) {
    Vtx v;
    const size_t vi = (size_t)t * 3 + k;
    v.pos[0] = p.x; v.pos[1] = p.y; v.pos[2] = p.z; v.pos[3] = p.w;
    transform_pos(st, p, v.hom);
    v.col = (PHASE == 1) ? A.col[vi] : 0u;
    if (PHASE == 1 && A.uv) { float2 q = A.uv[vi]; v.uv[0] = q.x; v.uv[1] = q.y; }
    else { v.uv[0] = v.uv[1] = 0.0f; }
    if (PHASE == 1 && (LIT || EXTRA) && A.nrm) { v.nrm[0] = A.nrm[3 * vi]; v.nrm[1] = A.nrm[3 * vi + 1]; v.nrm[2] = A.nrm[3 * vi + 2]; }
    else { v.nrm[0] = v.nrm[1] = v.nrm[2] = 0.0f; }
    v.scr[0] = v.scr[1] = 0.0f;
    v.bw[0] = k == 0 ? 1.0f : 0.0f;
    v.bw[1] = k == 1 ? 1.0f : 0.0f;
    v.bw[2] = k == 2 ? 1.0f : 0.0f;
    if (PHASE == 1 && LIT && (st->flags & PFF_LIGHTING)) {                // triangles.c:78-104
        const PFFstate& stg = A.states[si];
        const float m0 = st->matNormal[0], m1 = st->matNormal[1], m2 = st->matNormal[2];
        const float m4 = st->matNormal[4], m5 = st->matNormal[5], m6 = st->matNormal[6];
        const float m8 = st->matNormal[8], m9 = st->matNormal[9], m10 = st->matNormal[10];
        const float m12 = st->matNormal[12], m13 = st->matNormal[13], m14 = st->matNormal[14];
        float nx = m0 * v.nrm[0] + m4 * v.nrm[1] + m8 * v.nrm[2] + m12;
        float ny = m1 * v.nrm[0] + m5 * v.nrm[1] + m9 * v.nrm[2] + m13;
        float nz = m2 * v.nrm[0] + m6 * v.nrm[1] + m10 * v.nrm[2] + m14;
        v.nrm[0] = nx; v.nrm[1] = ny; v.nrm[2] = nz;
        float sq = nx * nx + ny * ny + nz * nz;
        if (sq != 0.0f) {
            float inv = __fdiv_rn(1.0f, sqrt_cr(sq));
            for (int i = 0; i < 3; i++) v.nrm[i] = v.nrm[i] * inv;
        }
        uint32_t dcol = st->material[face].diffuse, cc = 0;
        for (int i = 0; i < 4; i++) cc |= (uint32_t)((ch(v.col, i) * ch(dcol, i)) / 255) << (8 * i);
        v.col = cc;
        if (st->lightMode == 0) {
            float vr[3] = { st->viewRow2[0], st->viewRow2[1], st->viewRow2[2] };
            float ndv = dot3(v.nrm, vr);
            v.col = light_vertex(stg, stg.material[ndv < 0.0f ? 0 : 1], v.col, stg.viewPos, v.pos, v.nrm);
        }
    }
    return v;
}

// All culling passes of recorded triangle t under state st.  PHASE 0 counts
// raster triangles and tile pairs (geometry only); PHASE 1 also lights the
// vertices and writes records and pairs from the given slots.  With CLIP
// false a triangle that needs clipping is left alone and true is returned
// (a clip kernel sets it up); with CLIP true the lane clips it itself in
// the LDS buffers it is given (no per-lane private arrays, so no scratch).
template <int PHASE, bool LIT, bool CLIP, int VF = VF_FULL, bool EXTRA = true>
__device__ __forceinline__ bool setup_triangle(const SetupArgs& A, CState* st, uint32_t si, const DFrame& F,
                                               uint32_t t, uint32_t info, const float4* P, uint32_t rtBase,
                                               uint32_t pairBase, uint32_t& nRT, uint32_t& nPairs,
                                               CVtx<VF>* cpoly = nullptr, CVtx<VF>* ctmp = nullptr,
                                               PairRect* rects = nullptr, LaneRects* lr = nullptr) {
    // Process_ProjectAndClipTriangle (triangles.c:246-280): the 2D and clip
    // tests from the positions alone.  The clip planes are taken in the
    // clipper's order (w >= eps, x <= w, -x <= w, y.., z..); a plane that has
    // all three vertices outside while every earlier plane had all of them
    // inside (so the polygon reaching it is still the triangle) clips it to
    // nothing, exactly as the clipper would.
    bool is2D, inside, reject = false;
    {
        float h[3][4];
        float wsum = 0.0f;
        for (int k = 0; k < 3; k++) {
            transform_pos(st, P[k], h[k]);
            wsum += h[k][3];
        }
        is2D = fabsf(wsum - 3.0f) < 1e-5f;
        bool unchanged = true;
        for (int pl = 0; pl < 7 && unchanged; pl++) {
            int nout = 0;
            for (int k = 0; k < 3; k++) {
                bool out;
                if (pl == 0) out = h[k][3] < 1e-5f;
                else {
                    const int ax = (pl - 1) >> 1;
                    const float a = ((pl - 1) & 1) ? -h[k][ax] : h[k][ax];
                    out = !(a <= h[k][3]);
                }
                nout += out ? 1 : 0;
            }
            if (nout == 3) reject = true;
            if (nout > 0) unchanged = false;
        }
        inside = unchanged;
    }
    if (!is2D && reject) return false;
    if (!CLIP && !is2D && !inside) return true;
    const float stl[4] = { (float)st->vpPos[0], (float)st->vpPos[1], (float)st->vpDim[0], (float)st->vpDim[1] };
    for (int face = 0; face < 2; face++) {
        if (!(info & (face ? PFF_PASS_BACK : PFF_PASS_FRONT))) continue;
        // Vertex attributes are (re)loaded per pass rather than kept live
        // across both; the positions come in P.
        Vtx v[3];
        for (int k = 0; k < 3; k++) v[k] = corner_vertex<PHASE, LIT, EXTRA>(A, st, si, face, t, k, P[k]);
        uint32_t np = 0;
        // With lr (PHASE 1) the pairs are handed back as rectangles for the
        // caller to write.
        PairRect rr{ 0u, 0u };
        if (is2D) {                                                       // "2D" path
            for (int k = 0; k < 3; k++) to_screen(stl, v[k]);
            if (emit_raster<PHASE, EXTRA>(A, st, F, si, t, face, false, v[0], v[1], v[2], rtBase + nRT, pairBase + nPairs, np,
                                          (PHASE == 1 && lr) ? &rr : nullptr)) {
                if (PHASE == 1 && lr) lr->put(nRT == 0, rr);
                nRT++;
                nPairs += np;
            }
            continue;
        }
        if (inside) {
            for (int k = 0; k < 3; k++) {
                v[k].hom[2] = __fdiv_rn(1.0f, v[k].hom[2]);
                v[k].uv[0] = v[k].uv[0] * v[k].hom[2];
                v[k].uv[1] = v[k].uv[1] * v[k].hom[2];
                float invW = __fdiv_rn(1.0f, v[k].hom[3]);
                v[k].hom[0] *= invW;
                v[k].hom[1] *= invW;
                to_screen(stl, v[k]);
            }
            if (emit_raster<PHASE, EXTRA>(A, st, F, si, t, face, true, v[0], v[1], v[2], rtBase + nRT, pairBase + nPairs, np,
                                          (PHASE == 1 && lr) ? &rr : nullptr)) {
                if (PHASE == 1 && lr) lr->put(nRT == 0, rr);
                nRT++;
                nPairs += np;
            }
            continue;
        }
        if (CLIP) {
            for (int k = 0; k < 3; k++) cv_store(cpoly[k], v[k]);
            CVtx<VF>* poly = cpoly;              // the buffer holding the result (either one)
            int n = clip_polygon<VF>(cpoly, ctmp, 3, &poly);
            if (n < 3) continue;
            for (int k = 0; k < n; k++) {
                Vtx q = cv_load(poly[k]);
                q.hom[2] = __fdiv_rn(1.0f, q.hom[2]);
                q.uv[0] = q.uv[0] * q.hom[2];
                q.uv[1] = q.uv[1] * q.hom[2];
                float invW = __fdiv_rn(1.0f, q.hom[3]);
                q.hom[0] *= invW;
                q.hom[1] *= invW;
                cv_store(poly[k], q);
            }
            for (int i = 0; i < n - 2; i++) {                             // triangles.c:113-115
                Vtx a = cv_load(poly[0]), b = cv_load(poly[i + 1]), c = cv_load(poly[i + 2]);
                to_screen(stl, a);
                to_screen(stl, b);
                to_screen(stl, c);
                if (emit_raster<PHASE>(A, st, F, si, t, face, true, a, b, c, rtBase + nRT, pairBase + nPairs, np,
                                       (rects && nRT < MAX_FAN_RECTS) ? &rects[nRT] : nullptr)) {
                    nRT++;
                    nPairs += np;
                }
            }
        }
    }
    return false;
}

// Inclusive prefix sum over the wavefront, by DPP in the VALU: row shifts
// 1, 2, 4, 8 scan each row of 16 lanes (row_shr:n = 0x110 + n; lanes
// shifted in from outside the row add 0), then row_bcast:15 (0x142) adds
// the last lane of rows 0 and 2 to rows 1 and 3 (row mask 0xA) and
// row_bcast:31 (0x143) lane 31 to rows 2 and 3 (row mask 0xC); rows a mask
// leaves out add 0.  Every lane of the wavefront must take part.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

// Exclusive block scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint32_t s = lane < nw ? lds[lane] : 0u;
        uint32_t si = wave_incl_scan(s);
        if (lane < nw) lds[lane] = si - s;
        if (lane == nw - 1) lds[32] = si;
    }
    __syncthreads();
    uint32_t r = inc - v + lds[wave];
    total = lds[32];
    __syncthreads();
    return r;
}

// Exclusive block scans of N values per thread at once (one pass of
// barriers for all of them); total[i] is the block total of value i.
// Blocks of at most 16 wavefronts; lds holds at least 16 N + N words.
template <int N>
__device__ __forceinline__ void block_excl_scan_n(const uint32_t (&v)[N], uint32_t* lds, uint32_t (&r)[N],
                                                  uint32_t (&total)[N]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc[N];
#pragma unroll
    for (int i = 0; i < N; i++) inc[i] = wave_incl_scan(v[i]);
    if (lane == 63)
#pragma unroll
        for (int i = 0; i < N; i++) lds[16 * i + wave] = inc[i];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const uint32_t w = lane < nw ? lds[16 * i + lane] : 0u;
            const uint32_t wi = wave_incl_scan(w);
            if (lane < nw) lds[16 * i + lane] = wi - w;
            if (lane == nw - 1) lds[16 * N + i] = wi;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; i++) {
        r[i] = inc[i] - v[i] + lds[16 * i + wave];
        total[i] = lds[16 * N + i];
    }
    __syncthreads();
}

// Recorded triangle, state and frame of pass-triangle gid: the last
// segment whose passTriBase <= gid, found forward from its block's first.
__device__ __forceinline__ void pass_triangle(const SetupArgs& A, uint32_t gid, uint32_t& t, uint32_t& info,
                                              uint32_t& si, DFrame& F, uint32_t& frame) {
    uint32_t lo = A.blockSeg[gid / 256];
    while (lo + 1 < A.numSegs && A.segs[lo + 1].passTriBase <= gid) lo++;
    const DSeg sg = A.segs[lo];
    t = sg.triFirst + (gid - sg.passTriBase);
    frame = sg.frame;
    F = A.frames[sg.frame];
    info = A.info[t];
    si = info & PFF_STATE_MASK;
}

// The wavefront writes one pair rectangle per lane (w x h tiles from tile0,
// rows tilesX apart, pairs from slot with raster triangle rt) together, 64
// pairs per step: the clipped triangles of a wavefront are few but large,
// and a lane looping over its own tiles would hold the others.  Every lane
// of the wavefront must call it; lanes without a rectangle pass w = 0.
__device__ __forceinline__ void write_pair_rects(const SetupArgs& A, uint32_t slot, uint32_t tile0, uint32_t w,
                                                 uint32_t h, uint32_t tilesX, uint32_t rt, uint32_t quads) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = w * h;
    const uint32_t incl = wave_incl_scan(n);
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    const uint32_t excl = incl - n;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t i = base + (uint32_t)lane;
        // Owner: the last lane whose first pair is <= i (lanes without
        // pairs share the next lane's start and are passed over).
        int o = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const uint32_t e = (uint32_t)__shfl((int)excl, o + step, 64);
            if (e <= i) o += step;
        }
        const uint32_t q = i - (uint32_t)__shfl((int)excl, o, 64);
        const uint32_t oSlot = (uint32_t)__shfl((int)slot, o, 64);
        const uint32_t oTile = (uint32_t)__shfl((int)tile0, o, 64);
        const uint32_t oW = (uint32_t)__shfl((int)w, o, 64);
        const uint32_t oTX = (uint32_t)__shfl((int)tilesX, o, 64);
        const uint32_t oRt = (uint32_t)__shfl((int)rt, o, 64);
        const uint32_t oH = (uint32_t)__shfl((int)h, o, 64);
        const uint32_t oQ = (uint32_t)__shfl((int)quads, o, 64);
        if (i < total) {
            const uint32_t qy = q / oW, qx = q - qy * oW;
            const uint32_t p = oSlot + q;
            if (p < A.capPairs) {
                A.keys[p] = oTile + qy * oTX + qx;
                A.vals[p] = pair_value(oRt, qx, oW, qy, oH, oQ);
            }
        }
    }
}

// The listed triangles that need clipping, one per lane (k_clip_lane), each lane
// clipping in its own LDS area (so no per-lane private arrays: no
// scratch), in the smallest vertex form the pass needs.  One wave per
// workgroup; the list is walked with a grid stride.  PHASE 0 stores the
// counts and adds them to the block sums; PHASE 1 writes from the slots
// k_setup<1> handed over.
template <int VF>
constexpr int clip_words() { return 2 * 12 * (int)(sizeof(CVtx<VF>) / 4) + 1; }   // odd: lanes in distinct banks

template <int PHASE, bool LIT, int VF>
__global__ __launch_bounds__(64) void k_clip_lane(SetupArgs A) {
    constexpr int CLIP_WORDS = clip_words<VF>();
    __shared__ uint32_t mem[64 * CLIP_WORDS];
    __shared__ PairRect rectMem[PHASE == 1 ? 64 * MAX_FAN_RECTS : 1];
    const int lane = threadIdx.x;
    CVtx<VF>* poly = reinterpret_cast<CVtx<VF>*>(&mem[lane * CLIP_WORDS]);
    CVtx<VF>* tmp = poly + 12;
    PairRect* rects = PHASE == 1 ? &rectMem[lane * MAX_FAN_RECTS] : nullptr;
    const uint32_t n = min(A.clipKnown != ~0u ? A.clipKnown : *A.clipCount, A.clipCap);
    // Wavefront-uniform loop: the pair writing below needs every lane.
    for (uint32_t i0 = blockIdx.x * 64; i0 < n; i0 += gridDim.x * 64) {
        const uint32_t i = i0 + (uint32_t)lane;
        const bool valid = i < n;
        uint32_t nRT = 0, nPairs = 0, rtBase = 0, pairBase = 0, tilesX = 0;
        if (valid) {
            // The listed triangle with its recorded index, info and frame, so
            // that its inputs are fetched without the segment lookup's chain
            // of dependent loads.
            const uint4 cr = A.clipList[i];
            const uint32_t gid = cr.x, t = cr.y, info = cr.z, si = info & PFF_STATE_MASK;
            const DFrame F = A.frames[cr.w];
            tilesX = F.tilesX;
            float4 P[3];
            for (int k = 0; k < 3; k++) P[k] = A.pos[(size_t)t * 3 + k];
            if (PHASE == 1) { rtBase = A.cntRT[gid]; pairBase = A.cntPairs[gid]; }
            setup_triangle<PHASE, LIT, true, VF>(A, as_const(A.states + si), si, F, t, info, P, rtBase, pairBase,
                                                 nRT, nPairs, poly, tmp, rects);
            if (PHASE == 0) {
                A.cntRT[gid] = nRT;
                A.cntPairs[gid] = nPairs;
                if (nRT) atomicAdd(&A.blockRT[gid / 256], nRT);
                if (nPairs) atomicAdd(&A.blockPairs[gid / 256], nPairs);
            }
        }
        if (PHASE == 1) {
            // The lanes' k-th raster triangles' pairs, k = 0, 1, ...
            uint32_t most = min(nRT, (uint32_t)MAX_FAN_RECTS);
            for (int d = 1; d < 64; d <<= 1) most = max(most, (uint32_t)__shfl_xor((int)most, d, 64));
            uint32_t slot = pairBase;
            for (uint32_t k = 0; k < most; k++) {
                const bool has = k < min(nRT, (uint32_t)MAX_FAN_RECTS);
                const PairRect r = has ? rects[k] : PairRect{ 0u, 0u };
                const uint32_t w = r.wh & 0xFFFFu, h = r.wh >> 16;
                write_pair_rects(A, slot, r.tile0, w, h, tilesX, rtBase + k, r.q);
                slot += w * h;
            }
        }
    }
}

// ---------------------------------------------------------------------------

// The pairs of one pair rectangle written by its lane alone, in the same
// order (rows of w tiles from tile0, tilesX apart).
__device__ __forceinline__ void write_pairs_lane(const SetupArgs& A, uint32_t slot, const PairRect& r, uint32_t tilesX,
                                                 uint32_t rt) {
    const uint32_t w = r.wh & 0xFFFFu, h = r.wh >> 16;
    uint32_t p = slot;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++, p++) {
            if (p >= A.capPairs) return;
            A.keys[p] = r.tile0 + y * tilesX + x;
            A.vals[p] = pair_value(rt, x, w, y, h, r.q);
        }
}

// LDS writes of a wavefront made visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane groups are the DPP unit's rows (lanes 16g .. 16g+15 of the
// wavefront): the scans below move values within a row by DPP, in the VALU,
// instead of through LDS.  Every lane of the group must take part.
// DPP controls: row_shr:n = 0x110 + n (lane i takes lane i-n of its row;
// lanes i < n take `old`, here 0), row_ror:n = 0x120 + n (rotation within
// the row).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_row(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

// Inclusive prefix sum over the lane group.
__device__ __forceinline__ uint32_t group_incl_scan(uint32_t v) {
    v += dpp_row<0x111>(v);
    v += dpp_row<0x112>(v);
    v += dpp_row<0x114>(v);
    v += dpp_row<0x118>(v);
    return v;
}

// Sum over the lane group, in every lane of it.
__device__ __forceinline__ uint32_t group_sum(uint32_t v) {
    v += dpp_row<0x128>(v);
    v += dpp_row<0x124>(v);
    v += dpp_row<0x122>(v);
    v += dpp_row<0x121>(v);
    return v;
}

// The listed triangles that need clipping, 16 lanes per triangle (four per
// wavefront; k_clip_group): lower latency per triangle than k_clip_lane,
// lower throughput, so it takes passes that clip few triangles.  Per culling pass the group's lanes 0-2 build (and light) the
// corners; the polygon is clipped plane by plane against w >= eps, then
// +-x, +-y, +-z <= w (Sutherland-Hodgman, triangles.c:157-244) with lane i
// taking vertex i and the edge from its predecessor: where an edge crosses
// the plane the lane emits the intersection (lerp of the two, ictx.c:67-90)
// and then, when inside, its vertex, at places given by a prefix sum of the
// emissions over the group, which is the serial clipper's output order
// (and its cap of 12 vertices).  A plane with every vertex inside leaves
// the polygon as it is and is skipped.  The fan triangles (0, i+1, i+2)
// are then set up one per lane, their record and pair slots from prefix
// sums of the survivors.  PHASE 0 stores the counts and adds them to the
// block sums; PHASE 1 writes from the slots k_scan_local handed over, the
// pairs of the whole wavefront's fan triangles together.
// One step of a wavefront (every lane must call it): listed items i0 ..
// i0+3 below n, one per lane group, in the wavefront's LDS buffers mem.
template <int PHASE, bool LIT, int VF>
__device__ __forceinline__ void clip_group_items(const SetupArgs& A, CVtx<VF> (*mem)[2][12], uint32_t i0, uint32_t n) {
    const int lane = threadIdx.x & 63, grp = lane >> 4, gl = lane & 15;
    {
        const uint32_t item = i0 + (uint32_t)grp;
        const bool valid = item < n;
        // The listed triangle with its recorded index, info and frame.  An
        // idle group reads the wavefront's first item (i0 < n) and does
        // nothing with it: loads outside a branch join no values, so the
        // loads that depend on them are issued without waiting.
        const Words4 cr = reinterpret_cast<const Words4*>(A.clipList)[valid ? item : i0];
        const uint32_t gid = cr.x, t = cr.y, info = cr.z, si = info & PFF_STATE_MASK;
        const DFrame F = A.frames[cr.w];
        const uint32_t tilesX = valid ? F.tilesX : 0u;
        CState* st = as_const(A.states + si);
        const uint32_t rtBase = (PHASE == 1 && valid) ? A.cntRT[gid] : 0u;
        const uint32_t pairBase = (PHASE == 1 && valid) ? A.cntPairs[gid] : 0u;
        uint32_t nRT = 0, nPairs = 0;
        // Read before the clipping, whose LDS fences would keep the loads
        // behind them.
        const float stl[4] = { (float)st->vpPos[0], (float)st->vpPos[1], (float)st->vpDim[0], (float)st->vpDim[1] };
        // Unlit corners are the same in both culling passes: built once, with
        // their loads issued beside the state's.
        Vtx unlit;
        if (!LIT && valid && gl < 3) unlit = corner_vertex<PHASE, false>(A, st, si, 0, t, gl, A.pos[(size_t)t * 3 + gl]);
        for (int face = 0; face < 2; face++) {
            const bool act = valid && (info & (face ? PFF_PASS_BACK : PFF_PASS_FRONT));
            CVtx<VF>* in = mem[grp][0];
            CVtx<VF>* out = mem[grp][1];
            if (act && gl < 3)
                cv_store(in[gl], LIT ? corner_vertex<PHASE, LIT>(A, st, si, face, t, gl, A.pos[(size_t)t * 3 + gl]) : unlit);
            wave_sync();
            uint32_t cnt = act ? 3u : 0u;                      // uniform over the group
            for (int pl = 0; pl < 7; pl++) {
                const int ax = (pl - 1) >> 1, side = (pl - 1) & 1;
                const bool has = (uint32_t)gl < cnt;
                const int prev = gl == 0 ? (int)cnt - 1 : gl - 1;
                float cw = 0.0f, ca = 0.0f, pw = 0.0f, pa = 0.0f;
                if (has) {
                    const float* hc = cv_hom(in[gl]);
                    const float* hp = cv_hom(in[prev]);
                    cw = hc[3]; pw = hp[3];
                    if (pl > 0) { ca = side ? -hc[ax] : hc[ax]; pa = side ? -hp[ax] : hp[ax]; }
                }
                const bool insC = pl == 0 ? !(cw < 1e-5f) : (ca <= cw);
                const bool insP = pl == 0 ? !(pw < 1e-5f) : (pa <= pw);
                const uint32_t outside = (uint32_t)(__ballot(has && !insC) >> (16 * grp)) & 0xFFFFu;
                if (outside == 0u) continue;                   // uniform over the group
                const int cd = insC ? 1 : -1, pd = insP ? 1 : -1;
                const bool cross = has && (pl == 0 ? pd * cd < 0 : pd * cd <= 0);
                const bool keep = has && insC;
                const uint32_t e = (cross ? 1u : 0u) + (keep ? 1u : 0u);
                const uint32_t incl = group_incl_scan(e);
                const uint32_t at = incl - e;
                const uint32_t total = group_sum(e);
                if (cross && at < 12u) {
                    const float* hc = cv_hom(in[gl]);
                    const float* hp = cv_hom(in[prev]);
                    float tt;
                    if (pl == 0) tt = __fdiv_rn(1e-5f - hp[3], hc[3] - hp[3]);
                    else {
                        float num, den;
                        if (!side) { num = hp[3] - hp[ax]; den = num - (hc[3] - hc[ax]); }
                        else       { num = hp[3] + hp[ax]; den = num - (hc[3] + hc[ax]); }
                        tt = __fdiv_rn(num, den);
                    }
                    cv_store(out[at], lerp_vtx(cv_load(in[prev]), cv_load(in[gl]), tt));
                }
                const uint32_t atK = at + (cross ? 1u : 0u);
                if (keep && atK < 12u) out[atK] = in[gl];
                wave_sync();
                cnt = min(total, 12u);
                CVtx<VF>* sw = in; in = out; out = sw;
            }
            // Perspective divide (triangles.c:113-115 through
            // Process_ProjectAndClipTriangle's tail), one vertex per lane.
            if (cnt >= 3u && (uint32_t)gl < cnt) {
                Vtx q = cv_load(in[gl]);
                q.hom[2] = __fdiv_rn(1.0f, q.hom[2]);
                q.uv[0] = q.uv[0] * q.hom[2];
                q.uv[1] = q.uv[1] * q.hom[2];
                const float invW = __fdiv_rn(1.0f, q.hom[3]);
                q.hom[0] *= invW;
                q.hom[1] *= invW;
                cv_store(in[gl], q);
            }
            wave_sync();
            // Fan triangle gl: (0, gl+1, gl+2), in screen space.
            const bool fan = cnt >= 3u && (uint32_t)gl < cnt - 2u;
            auto corner = [&](int k) {
                Vtx q = cv_load(in[k]);
                to_screen(stl, q);
                return q;
            };
            uint32_t np = 0;
            bool ok = false;
            if (fan) {
                const Vtx a = corner(0), b = corner(gl + 1), c = corner(gl + 2);
                ok = emit_raster<0>(A, st, F, si, t, face, true, a, b, c, 0u, 0u, np);
            }
            if (!ok) np = 0;
            const uint32_t rInc = group_incl_scan(ok ? 1u : 0u);
            const uint32_t pInc = group_incl_scan(np);
            const uint32_t rAt = nRT + rInc - (ok ? 1u : 0u), pAt = nPairs + pInc - np;
            nRT += group_sum(ok ? 1u : 0u);
            nPairs += group_sum(np);
            if (PHASE == 1) {
                PairRect r{ 0u, 0u };
                if (ok) {
                    const Vtx a = corner(0), b = corner(gl + 1), c = corner(gl + 2);
                    uint32_t np2 = 0;
                    if (!emit_raster<1, true, true>(A, st, F, si, t, face, true, a, b, c, rtBase + rAt, pairBase + pAt, np2, &r))
                        r = PairRect{ 0u, 0u };
                }
                write_pair_rects(A, pairBase + pAt, r.tile0, r.wh & 0xFFFFu, r.wh >> 16, tilesX, rtBase + rAt, r.q);
            }
            wave_sync();                                   // the next pass rebuilds the corners
        }
        if (PHASE == 0 && valid && gl == 0) {
            A.cntRT[gid] = nRT;
            A.cntPairs[gid] = nPairs;
            if (nRT) atomicAdd(&A.blockRT[gid / 256], nRT);
            if (nPairs) atomicAdd(&A.blockPairs[gid / 256], nPairs);
        }
    }
}

template <int PHASE, bool LIT, int VF>
__global__ __launch_bounds__(64) void k_clip_group(SetupArgs A) {
    __shared__ CVtx<VF> mem[4][2][12];
    const uint32_t n = min(A.clipKnown != ~0u ? A.clipKnown : *A.clipCount, A.clipCap);
    // Wavefront-uniform loop: the pair writing needs every lane.
    for (uint32_t i0 = blockIdx.x * 4; i0 < n; i0 += gridDim.x * 4) clip_group_items<PHASE, LIT, VF>(A, mem, i0, n);
}

// One thread per recorded triangle of the pass.  PHASE 0 counts and sums
// per block; PHASE 1 turns the block sums (scanned in between) and a
// block scan of the counts into output slots and writes.  Triangles that
// need clipping are only listed here (PHASE 0) and handed their slots
// (PHASE 1, through cntRT/cntPairs); a clip kernel, or with CLIPW this
// kernel's leading workgroups, does their work.
// CLIPW (PHASE 1 of unlit passes without float records): the first
// A.clipBlocks workgroups clip the listed triangles instead, 16 per
// workgroup and step (clip_group_items), beside the others' setup: one
// launch, and the clip waves, dispatched first, overlap it.
template <int PHASE, bool LIT, bool EXTRA = true, bool CLIPW = false>
__global__ __launch_bounds__(256) void k_setup(SetupArgs A) {
    __shared__ uint32_t lds[64];
    if constexpr (CLIPW) {
        if (blockIdx.x < A.clipBlocks) {
            __shared__ CVtx<VF_SLIM> cmem[4][4][2][12];
            const uint32_t n = min(A.clipKnown, A.clipCap);
            const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            for (uint32_t i0 = (blockIdx.x * 4 + wave) * 4; i0 < n; i0 += A.clipBlocks * 16)
                clip_group_items<1, false, VF_SLIM>(A, cmem[wave], i0, n);
            return;
        }
    }
    const uint32_t gid = (blockIdx.x - (CLIPW ? A.clipBlocks : 0u)) * blockDim.x + threadIdx.x;
    const bool valid = gid < A.numPassTris;
    if (PHASE == 0) {
        // Reset the tile ranges of this pass (the raster reads [begin, end)).
        for (uint32_t i = gid; i < A.numTiles; i += gridDim.x * blockDim.x) { A.tileBegin[i] = 0; A.tileEnd[i] = 0; }
        // pfClear copies pixel 0 into the pixels past the last multiple of 8
        // before any later draw (context.c:707-710).  The raster kernel of
        // this pass may already be writing pixel 0 while other tiles clear
        // their tail, so the tail takes pixel 0 as it was when the pass
        // began: every earlier pass has finished by the time this runs.
        if (blockIdx.x == 0)
            for (uint32_t s = threadIdx.x; s < A.numSegs; s += blockDim.x) {
                const DSeg sg = A.segs[s];
                if (!sg.clear) continue;
                const DFrame F = A.frames[sg.frame];
                A.clearSnap[8 * sg.frame] = F.kind < PFF_TARGET_FIRST_WIDE ? tgt_load(F.color, F.kind, 0) : 0u;
                A.clearSnap[8 * sg.frame + 1] = __float_as_uint(F.depth[0]);
                uint8_t* snap = (uint8_t*)&A.clearSnap[8 * sg.frame + 2];
                if (F.kind >= PFF_TARGET_FIRST_WIDE)
                    for (uint32_t b = 0; b < F.bpp; b++) snap[b] = F.color[b];
            }
    }
    // The triangle, its frame and positions are requested first, so their
    // latency overlaps the block scans of PHASE 1.
    uint32_t t = 0, info = 0, si = 0, frame = 0;
    DFrame F;
    float4 P[3] = { make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f) };
    if (valid) {
        pass_triangle(A, gid, t, info, si, F, frame);
        for (int k = 0; k < 3; k++) P[k] = A.pos[(size_t)t * 3 + k];
    }
    // PHASE 1: the slots of this triangle's records and pairs (k_scan_local
    // turned the counts into offsets).
    const uint32_t rtBase = (PHASE == 1 && valid) ? A.cntRT[gid] : 0u;
    const uint32_t pairBase = (PHASE == 1 && valid) ? A.cntPairs[gid] : 0u;
    uint32_t nRT = 0, nPairs = 0;
    // One iteration per distinct state in the wavefront: the state is then
    // uniform and read with scalar loads.
    uint64_t todo = __builtin_amdgcn_ballot_w64(valid);
    bool deferred = false;
    LaneRects lr{ { 0u, 0u }, { 0u, 0u } };
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const uint32_t s0 = __builtin_amdgcn_readlane(si, leader);
        const bool mine = valid && si == s0;
        todo &= ~__builtin_amdgcn_ballot_w64(mine);
        if (mine)
            deferred = setup_triangle<PHASE, LIT, false, VF_FULL, EXTRA>(A, as_const(A.states + s0), s0, F, t, info,
                                                                         P, rtBase, pairBase, nRT, nPairs, nullptr,
                                                                         nullptr, nullptr, &lr);
    }
    if (PHASE == 1) {
        // The tile pairs: each lane its own when all are few, else the
        // wavefront together, 64 per step (a lane looping over a large
        // triangle's tiles would hold the others: 800x600 frame of 10k
        // large triangles, 510 pairs each).
        const uint32_t na = (lr.a.wh & 0xFFFFu) * (lr.a.wh >> 16), nb = (lr.b.wh & 0xFFFFu) * (lr.b.wh >> 16);
        if (__builtin_amdgcn_ballot_w64(na + nb > 16u)) {
            write_pair_rects(A, pairBase, lr.a.tile0, lr.a.wh & 0xFFFFu, lr.a.wh >> 16, F.tilesX, rtBase, lr.a.q);
            write_pair_rects(A, pairBase + na, lr.b.tile0, lr.b.wh & 0xFFFFu, lr.b.wh >> 16, F.tilesX, rtBase + 1u,
                             lr.b.q);
        } else {
            write_pairs_lane(A, pairBase, lr.a, F.tilesX, rtBase);
            write_pairs_lane(A, pairBase + na, lr.b, F.tilesX, rtBase + 1u);
        }
    }
    if (PHASE == 0) {
        if (valid) { A.cntRT[gid] = nRT; A.cntPairs[gid] = nPairs; }
        // Block sums, and the triangles to clip listed in order within the
        // block: one counter update per block (one per wavefront contended
        // on the counter).
        const uint32_t sv[3] = { nRT, nPairs, deferred ? 1u : 0u };
        uint32_t ex[3], tot[3];
        block_excl_scan_n<3>(sv, lds, ex, tot);
        const uint32_t totD = tot[2], dAt = ex[2];
        if (threadIdx.x == 0) {
            A.blockRT[blockIdx.x] = tot[0];
            A.blockPairs[blockIdx.x] = tot[1];
            if (totD) lds[63] = atomicAdd(A.clipCount, totD);
        }
        if (totD) {                                            // uniform over the block
            __syncthreads();
            // The list holds every triangle of the pass once the counter
            // starts at zero; the bound keeps a counter left over by a failed
            // run from writing past it.
            const uint32_t at = lds[63] + dAt;
            if (deferred && at < A.clipCap) A.clipList[at] = make_uint4(gid, t, info, frame);
        }
    }
}

// Per triangle of the pass, the first raster-triangle and pair slot: the
// block's scanned sums plus a scan within the block, in place over the
// counts.  Both k_setup<1> and the clip kernel then find their slots there, so
// the two run side by side.
// FUSED: the block sums are still raw (k_scan_top did not run); each block
// adds up those of the blocks before it itself (few blocks).
template <bool FUSED>
__global__ __launch_bounds__(256) void k_scan_local(SetupArgs A) {
    __shared__ uint32_t lds[64];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = gid < A.numPassTris;
    uint32_t baseR, baseP, ex[2], tot[2];
    if (FUSED) {
        uint32_t sb[2] = { 0u, 0u };
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256) { sb[0] += A.blockRT[b]; sb[1] += A.blockPairs[b]; }
        block_excl_scan_n<2>(sb, lds, ex, tot);
        baseR = tot[0];
        baseP = tot[1];
    } else {
        baseR = A.blockRT[blockIdx.x];
        baseP = A.blockPairs[blockIdx.x];
    }
    const uint32_t sc[2] = { valid ? A.cntRT[gid] : 0u, valid ? A.cntPairs[gid] : 0u };
    block_excl_scan_n<2>(sc, lds, ex, tot);
    if (valid) { A.cntRT[gid] = ex[0] + baseR; A.cntPairs[gid] = ex[1] + baseP; }
}


// The tile pairs of passes whose raster triangles average many pairs, from
// the rectangles the setup recorded: pair i belongs to the last raster
// triangle whose first pair is at or before i (first pairs ascend with the
// raster triangles, each of which has at least one pair), and is the
// (i - first)-th tile of its rectangle, rows of w from tile0.  Four
// consecutive pairs per thread share one search.
__global__ __launch_bounds__(256) void k_expand_pairs(const uint32_t* __restrict__ pairOff,
                                                      const uint4* __restrict__ pairRect, uint32_t numRT,
                                                      uint32_t numPairs, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
    const uint32_t i0 = (blockIdx.x * 256u + threadIdx.x) * 4u;
    if (i0 >= numPairs) return;
    uint32_t lo = 0, hi = numRT;                      // pairOff[lo] <= i0 < pairOff[hi] (hi: numPairs)
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pairOff[mid] <= i0) lo = mid; else hi = mid;
    }
    uint32_t rt = lo;
    uint32_t first = pairOff[rt], next = rt + 1u < numRT ? pairOff[rt + 1u] : numPairs;
    uint4 r = pairRect[rt];
    for (uint32_t i = i0; i < min(i0 + 4u, numPairs); i++) {
        while (i >= next) {                           // into the next raster triangle
            rt++;
            first = next;
            next = rt + 1u < numRT ? pairOff[rt + 1u] : numPairs;
            r = pairRect[rt];
        }
        const uint32_t q = i - first, w = r.y & 0xFFFFu;
        const uint32_t qy = q / w, qx = q - qy * w;
        keys[i] = r.x + qy * r.z + qx;
        vals[i] = pair_value(rt, qx, w, qy, r.y >> 16, r.w);
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-block sums of the setup (one workgroup)
// ---------------------------------------------------------------------------

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// One block scans an array in place (exclusive) and writes its total;
// with a second array (b != nullptr) it scans that one too.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint32_t nb,
                                                           uint32_t* __restrict__ totalOut) {
    __shared__ uint32_t lds[64];
    for (int which = 0; which < 2; which++) {
        uint32_t* arr = which ? b : a;
        if (!arr) break;
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < nb; c0 += SCAN_TILE) {
            uint32_t base = c0 + threadIdx.x * SCAN_ITEMS;
            uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
            for (int i = 0; i < SCAN_ITEMS; i++) { v[i] = (base + i < nb) ? arr[base + i] : 0u; s += v[i]; }
            uint32_t total;
            uint32_t ex = block_excl_scan(s, lds, total) + carry;
#pragma unroll
            for (int i = 0; i < SCAN_ITEMS; i++) { if (base + i < nb) arr[base + i] = ex; ex += v[i]; }
            carry += total;
        }
        if (threadIdx.x == 0) totalOut[which] = carry;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Tile pairs and the stable radix sort
// ---------------------------------------------------------------------------


constexpr int RADIX_THREADS = 256;

// Digits of DB bits: 2^DB buckets (DB = 8 is used).  A block sorts ROUNDS
// x RADIX_THREADS consecutive items: more, smaller blocks fill the chip
// when the pairs are few; fewer, larger ones keep the per-block work of
// the column scan and the scatter's set-up small when they are many.
template <int DB, int ROUNDS>
__global__ __launch_bounds__(RADIX_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                              uint32_t nb, uint32_t* __restrict__ hist) {
    constexpr uint32_t ND = 1u << DB;
    __shared__ uint32_t h[ND];
    for (uint32_t d = threadIdx.x; d < ND; d += RADIX_THREADS) h[d] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * (RADIX_THREADS * ROUNDS);
    for (int r = 0; r < ROUNDS; r++) {
        size_t i = base + (size_t)r * RADIX_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (ND - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < ND; d += RADIX_THREADS) hist[(size_t)d * nb + blockIdx.x] = h[d];
}

// One workgroup per digit d: the exclusive prefix over radix blocks of
// hist[d * nb + b], in place, and the digit's total.  With the totals the
// scatter forms each block's bases itself, so a pass needs no device-wide
// scan.
__global__ __launch_bounds__(RADIX_THREADS) void k_radix_colscan(uint32_t* __restrict__ hist, uint32_t nb,
                                                                 uint32_t* __restrict__ digitTotal) {
    __shared__ uint32_t lds[64];
    constexpr int ITEMS = 8;
    uint32_t* col = hist + (size_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += RADIX_THREADS * ITEMS) {
        const uint32_t base = c0 + threadIdx.x * ITEMS;
        uint32_t v[ITEMS], sum = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; i++) { v[i] = (base + i < nb) ? col[base + i] : 0u; sum += v[i]; }
        uint32_t total;
        uint32_t ex = block_excl_scan(sum, lds, total) + carry;
#pragma unroll
        for (int i = 0; i < ITEMS; i++) { if (base + i < nb) col[base + i] = ex; ex += v[i]; }
        carry += total;
    }
    if (threadIdx.x == 0) digitTotal[blockIdx.x] = carry;
}

// Stable scatter: items of a block are taken in index order, 256 per round;
// an item's rank among equal digits is its rank within its wave (ballot
// match on the DB digit bits) plus the counts of lower waves and of earlier
// rounds.
template <int DB, int ROUNDS>
__global__ __launch_bounds__(RADIX_THREADS) void k_radix_scatter(const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                                                 uint32_t n, uint32_t shift, uint32_t nb,
                                                                 const uint32_t* __restrict__ colPrefix,
                                                                 const uint32_t* __restrict__ digitTotal,
                                                                 uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    constexpr uint32_t ND = 1u << DB;
    constexpr int PER = (int)ND / RADIX_THREADS;          // digits per thread
    __shared__ uint32_t base[ND];
    __shared__ uint32_t cnt[RADIX_THREADS / 64][ND];
    __shared__ uint32_t lds[64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Items of digit d start at (items of lower digits) + (items of digit d
    // in earlier blocks); thread t owns digits t*PER .. t*PER + PER-1.
    {
        uint32_t tot[PER], s = 0;
#pragma unroll
        for (int q = 0; q < PER; q++) { tot[q] = digitTotal[threadIdx.x * PER + q]; s += tot[q]; }
        uint32_t all;
        uint32_t lower = block_excl_scan(s, lds, all);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint32_t d = threadIdx.x * PER + q;
            base[d] = lower + colPrefix[(size_t)d * nb + blockIdx.x];
            lower += tot[q];
        }
    }
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    size_t blockBase = (size_t)blockIdx.x * (RADIX_THREADS * ROUNDS);
    for (int r = 0; r < ROUNDS; r++) {
        for (int w = 0; w < RADIX_THREADS / 64; w++)
            for (uint32_t d = threadIdx.x; d < ND; d += RADIX_THREADS) cnt[w][d] = 0;
        __syncthreads();
        size_t i = blockBase + (size_t)r * RADIX_THREADS + threadIdx.x;
        bool valid = i < n;
        uint32_t key = valid ? keysIn[i] : 0u;
        uint32_t val = valid ? valsIn[i] : 0u;
        uint32_t d = (key >> shift) & (ND - 1u);
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t wrank = __popcll(match & ltMask);
        if (valid && wrank == 0) cnt[wave][d] = __popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + wrank;
            for (int w = 0; w < wave; w++) pos += cnt[w][d];
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
        for (uint32_t dd = threadIdx.x; dd < ND; dd += RADIX_THREADS) {
            uint32_t add = 0;
            for (int w = 0; w < RADIX_THREADS / 64; w++) add += cnt[w][dd];
            base[dd] += add;
        }
        __syncthreads();
    }
}

// Also clears the clip-list counter for the next pass (both clip kernels of
// this pass have read it by now).  Every key is a tile of the pass; one that
// is not (none, while the counts are right) is skipped rather than written
// past the ranges.
__global__ void k_tile_bounds(const uint32_t* __restrict__ keys, uint32_t n, uint32_t numTiles,
                              uint32_t* __restrict__ tileBegin, uint32_t* __restrict__ tileEnd,
                              uint32_t* __restrict__ clipCount) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *clipCount = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t k = keys[i];
        if (k >= numTiles) continue;
        if (i == 0 || keys[i - 1] != k) tileBegin[k] = (uint32_t)i;
        if (i == n - 1 || keys[i + 1] != k) tileEnd[k] = (uint32_t)i + 1;
    }
}

// ---------------------------------------------------------------------------
// Tile raster
// ---------------------------------------------------------------------------

struct RasterArgs {
    const RTri*     rts;
    const RTriL*    rtl;
    const PFFstate* states;
    const uint32_t* tileFrame;      // per tile: its frame
    const DTileFrame* tileDesc;     // per frame: frame and clear of this pass
    const uint32_t* tileBegin;
    const uint32_t* tileEnd;
    const uint32_t* vals;           // sorted pair values (raster triangle ids)
    uint32_t        numTiles;
    uint32_t*       tileFrags;      // [numTiles][4]: covered samples per tile and wave (plain
                                    // stores: one counter shared by all waves costs more than
                                    // the raster itself)
    uint32_t*       winner;         // per pixel of frame 0: 1 + raster triangle of the last
                                    // fragment written, 0 for none; null unless differentiable
    const uint32_t* clearSnap;      // per frame, 8 words: pixel 0 (colour word, depth bits, bytes) when the pass began
};

// ---------------------------------------------------------------------------
// Tile raster
//
// Each wave owns one 8x8 block of a tile (one pixel per lane; colour and
// depth stay in VGPRs from the fused clear / load to the single write-back)
// and walks the tile's pair list in batches of 32: the 32 lanes fetch 32
// list entries and their raster-triangle records in one round trip, keep
// the ones whose bbox meets the block and stage them, compacted and still in
// list order, in the wave's LDS.  Each staged triangle is then tested
// against the block (edge functions per lane, ballot).  Two ways to shade:
//   packed   (mode 0): covered (triangle, pixel) fragments are queued in LDS
//            with a per-pixel chain in submission order; when 64 are queued
//            lane i shades fragment i, then every lane walks its own pixel's
//            chain doing the depth test and blend.  Shading does not read
//            the framebuffer, so shading before testing gives the same
//            pixels as the reference's order.
//   in place (mode 1): the lanes a triangle covers shade and update their
//            pixel immediately (best when triangles cover whole blocks).
// ---------------------------------------------------------------------------

constexpr int BATCH = 32;

// Words between staged records.  36 (16-byte aligned, off the 32-bank
// period) spreads lanes reading one field of different records over the
// banks, but the larger queue lowers the workgroups per CU from 8 to 7 and
// measured slower on the batched workload; 32 is kept.
constexpr int REC_STRIDE = 32;

// Staged RTriL records of per-fragment lit triangles; no LDS otherwise
// (the queue of an unlit kernel then fits eight workgroups per CU).
template <bool PHONG> struct PhongStage { uint32_t recL[BATCH][20]; };
template <> struct PhongStage<false> {};

template <bool PHONG, typename Q>
__device__ __forceinline__ const float* staged_light(const Q& q, int j) {
    if constexpr (PHONG) return reinterpret_cast<const float*>(&q.recL[j][0]);
    else return nullptr;
}

template <bool PHONG>
struct WaveQueueT : PhongStage<PHONG> {
    uint32_t rec[BATCH][REC_STRIDE];        // staged RTri records
    unsigned long long rmask[BATCH];        // 8x8 blocks: bbox pixels of each staged record (bit = lane)
    uint32_t recId[BATCH];                  // raster-triangle index of each staged record
    uint32_t fcol[64];
    float    fz[64];
    uint8_t  ftri[64];
    uint8_t  fpix[64];
};

// Staged records, read and written as 16-byte chunks.  With ROT the chunks
// of record j would sit rotated by j (chunk c at (c + j) mod 8), which
// spreads one field of different records over the LDS banks; it measured
// slower with chunked reads and is off.
template <bool ROT, typename QT>
__device__ __forceinline__ void stage_rec(QT& Q, int slot, const uint4* q) {
    uint4* dst = reinterpret_cast<uint4*>(&Q.rec[slot][0]);
#pragma unroll
    for (int c = 0; c < 8; c++) dst[ROT ? ((c + slot) & 7) : c] = q[c];
}
// Chunks 0..NC-1 of staged record j (its leading fields); the rest zero.
template <bool ROT, int NC, typename QT>
__device__ __forceinline__ RTri staged_rec(const QT& Q, int j) {
    const uint4* src = reinterpret_cast<const uint4*>(&Q.rec[j][0]);
    RTri r;
    uint4* d = reinterpret_cast<uint4*>(&r);
#pragma unroll
    for (int c = 0; c < 8; c++) d[c] = c < NC ? src[ROT ? ((c + j) & 7) : c] : make_uint4(0u, 0u, 0u, 0u);
    return r;
}
template <bool ROT, typename QT>
__device__ __forceinline__ uint32_t staged_meta(const QT& Q, int j) {
    static_assert(offsetof(RTri, meta) == 80, "meta opens chunk 5");
    return Q.rec[j][(ROT ? ((5 + j) & 7) : 5) * 4];
}


// Edge values at (px, py); small says every coefficient fits 24 bits
// (META_SMALL), which allows 24-bit multiplies.
template <typename TRI>
__device__ __forceinline__ void edge_values_as(const TRI& t, bool small, int px, int py, int32_t& w1, int32_t& w2,
                                               int32_t& w3) {
    if (small) {
        // 24x24-bit multiplies keep the low 32 bits of the product: the same
        // value mod 2^32 as the full multiply when |A|, |B| < 2^23.
        w1 = __mul24(t.A[0], px) + __mul24(t.B[0], py) + t.C[0];
        w2 = __mul24(t.A[1], px) + __mul24(t.B[1], py) + t.C[1];
        w3 = __mul24(t.A[2], px) + __mul24(t.B[2], py) + t.C[2];
        return;
    }
    const uint32_t ux = (uint32_t)px, uy = (uint32_t)py;
    w1 = (int32_t)((uint32_t)t.A[0] * ux + (uint32_t)t.B[0] * uy + (uint32_t)t.C[0]);
    w2 = (int32_t)((uint32_t)t.A[1] * ux + (uint32_t)t.B[1] * uy + (uint32_t)t.C[1]);
    w3 = (int32_t)((uint32_t)t.A[2] * ux + (uint32_t)t.B[2] * uy + (uint32_t)t.C[2]);
}

template <typename TRI>
__device__ __forceinline__ void edge_values(const TRI& t, int px, int py, int32_t& w1, int32_t& w2, int32_t& w3) {
    edge_values_as(t, (t.meta & META_SMALL) != 0, px, py, w1, w2, w3);
}

// Colour and depth of one fragment of raster triangle t at pixel (px, py)
// (triangles.c:422-432, 500-521).
// The depth test of a fragment at depth z against the pixel's depth
// (depth.h:80-124: NOTEQUAL behaves as EQUAL); true when there is none.
// Branch-free: the outcome of the comparison as one bit (1 less, 2 equal,
// 4 greater; none when either value is NaN, where every comparison is
// false) tested against the outcomes each function passes, three bits per
// function: 0 and 1 equal, 2 less, 3 less or equal, 4 greater, 5-7
// greater or equal, and entry 8 (no depth test) all.
__device__ __forceinline__ bool depth_ok(uint32_t meta, float z, float dep) {
    constexpr uint32_t PASS = (2u << 0) | (2u << 3) | (1u << 6) | (3u << 9) | (4u << 12) | (6u << 15) | (6u << 18) |
                              (6u << 21) | (7u << 24);
    const uint32_t f = (meta & META_DEPTH) ? ((meta >> META_DFUNC_SH) & 7u) : 8u;
    const uint32_t o = (z < dep ? 1u : 0u) | (z == dep ? 2u : 0u) | (z > dep ? 4u : 0u);
    return ((PASS >> (3u * f)) & o) != 0u || f == 8u;
}

// EARLY: the depth test against `dep` first; a fragment that fails it is
// discarded by the ROP, so its colour is not computed and false is
// returned (the pixel's result is the same).  W: the edge values at the
// pixel are given (w1, w2, w3) rather than computed.
template <bool PHONG, bool EARLY = false, bool W = false>
__device__ __forceinline__ bool shade_fragment(const RasterArgs& A, const RTri& t, const float* lrec,
                                               int px, int py, uint32_t& outCol, float& outZ, float dep = 0.0f,
                                               int32_t w1 = 0, int32_t w2 = 0, int32_t w3 = 0) {
    if (!W) edge_values(t, px, py, w1, w2, w3);
    const float inv = t.wInvSum;
    float b1 = (float)w1 * inv, b2 = (float)w2 * inv, b3 = (float)w3 * inv;
    float zq = rcp_x86((t.z[0] * b1 + t.z[1] * b2) + t.z[2] * b3);
    const uint32_t meta = t.meta;
    outZ = zq;
    if (EARLY && !depth_ok(meta, zq, dep)) return false;
    uint32_t frag = (meta & META_SMOOTH) ? bary_smooth(t.c[0], t.c[1], t.c[2], b1, b2, b3)
                                         : bary_flat(t.c[0], t.c[1], t.c[2], b1, b2, b3);
    if (meta & META_TEX) {
        DTex tex;
        tex.pixels = (const void*)(uintptr_t)(((uint64_t)(t.tex[1] & 0xFFFFu) << 32) | t.tex[0]);
        tex.format = (t.tex[1] >> 16) & 1u;
        tex.wrap = (t.tex[1] >> 17) & 3u;
        tex.filter = (t.tex[1] >> 19) & 1u;
        tex.w = t.tex[2] & 0xFFFFu;
        tex.h = t.tex[2] >> 16;
        tex.pad = 0;
        float u = (t.uv[0] * b1 + t.uv[2] * b2) + t.uv[4] * b3;
        float v = (t.uv[1] * b1 + t.uv[3] * b2) + t.uv[5] * b3;
        if (meta & META_3D) { u = u * zq; v = v * zq; }
        frag = modulate(sample_tex(tex, u, v), frag);
    }
    if (PHONG && (meta & META_PHONG)) {
        const PFFstate* st = &A.states[t.state];
        float N[3], P[3];
        for (int i = 0; i < 3; i++) {
            N[i] = (lrec[i] * b1 + lrec[3 + i] * b2) + lrec[6 + i] * b3;
            P[i] = (lrec[9 + i] * b1 + lrec[12 + i] * b2) + lrec[15 + i] * b3;
        }
        frag = light_fragment(st, st->material[(meta & META_BACK) ? 1 : 0], frag, P, N);
    }
    outCol = frag;
    return true;
}

// Wide targets' SIMD accessors (defined with the texel decoders below).
// A wide pixel's `win` is 0 while the frame has not written it, WIN_SNAP
// when a clear's tail gave it pixel 0's bytes (kept in clearSnap), else
// the writer's raster triangle + 1.
constexpr uint32_t WIN_SNAP = 0xFFFFFFFEu;
__device__ __forceinline__ uint32_t wide_get_simd(const uint8_t* p, uint32_t kind, size_t i);
__device__ __forceinline__ void wide_set_simd(uint8_t* p, uint32_t kind, size_t i, uint32_t c);
__device__ __forceinline__ uint32_t wide_rt_simd(uint32_t kind, uint32_t c);

// Depth test and blend of one fragment against the lane's pixel
// (depth.h:80-124, blend.h:137-274, triangles.c:434-438, 523-529).
// TK, the targets a kernel instance meets: 0 RGBA8 only, 1 any kind of at
// most 32 bits, 2 wide ones too (each instance compiles only the paths it
// needs: the others cost registers and instructions).
template <int TK>
__device__ __forceinline__ void rop(uint32_t meta, uint32_t c, float z, uint32_t id, uint32_t kind, uint32_t& col,
                                    float& dep, uint32_t& win) {
    if (!depth_ok(meta, z, dep)) return;
    if (TK == 0 || kind == PFF_TARGET_RGBA8) {
        if (meta & META_BLEND) c = blend_op((meta >> META_BMODE_SH) & 7u, c, col);
        col = c;
    } else if (TK == 2 && kind >= PFF_TARGET_FIRST_WIDE) {
        // col is the RGBA8 value of the pixel: the getter's reading of its
        // bytes while unwritten, else the colour last written, which the
        // target holds through the SIMD setter.
        if (meta & META_BLEND)
            c = blend_op((meta >> META_BMODE_SH) & 7u, c, (win != 0u && win != WIN_SNAP) ? wide_rt_simd(kind, col) : col);
        col = c;
    } else {
        // col holds the target's stored pixel: blending reads it through the
        // SIMD getter and the result goes back through the SIMD setter
        // (triangles.c:523-529).
        if (meta & META_BLEND) c = blend_op((meta >> META_BMODE_SH) & 7u, c, tgt_get_simd(kind, col));
        col = tgt_set_simd(kind, c);
    }
    dep = z;
    win = id + 1u;
}



// Shades the queued fragments and applies them in order (packed mode).
// Queued pixel p of a TW-wide block is (bx0 + p % TW, by0 + p / TW); lane
// l owns pixels l + 64k, k < K, whose fragments sit in the queue slots set
// in pm[k] (slots grow in submission order).
template <bool PHONG, int K, bool ROT, int TK>
__device__ __forceinline__ void flush_queue(const RasterArgs& A, WaveQueueT<PHONG>& Q, int bx0, int by0, int lane,
                                            int nf, uint64_t* pm, uint32_t kind, uint32_t* col, float* dep,
                                            uint32_t* win) {
    constexpr int TW = (K == 1) ? 8 : 16;
    wave_sync();
    if (lane < nf) {
        const int j = Q.ftri[lane];
        const int p = Q.fpix[lane];
        const RTri t = staged_rec<ROT, 8>(Q, j);
        uint32_t c;
        float z;
        shade_fragment<PHONG>(A, t, staged_light<PHONG>(Q, j), bx0 + (p % TW),
                              by0 + (p / TW), c, z);
        Q.fcol[lane] = c;
        Q.fz[lane] = z;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < K; k++) {
        uint64_t m = pm[k];
        pm[k] = 0ull;
        while (m) {
            const int s = __builtin_ctzll(m);
            m &= m - 1ull;
            const int j = Q.ftri[s];
            rop<TK>(staged_meta<ROT>(Q, j), Q.fcol[s], Q.fz[s], Q.recId[j], kind, col[k], dep[k], win[k]);
        }
    }
    wave_sync();
}

// Set bits of the wavefront mask m below the calling lane (v_mbcnt; the
// same as popcount(m & lanes-below-mask)).
__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Pixels x0..x1 of rows y0..y1 of an 8x8 block as a 64-bit mask, bit
// 8 y + x (the lane of pixel (x, y) when a wave holds one pixel per lane).
__device__ __forceinline__ uint64_t rect_mask8(int x0, int x1, int y0, int y1) {
    const uint64_t row = (uint64_t)(((2u << x1) - 1u) & ~((1u << x0) - 1u));
    return (row * 0x0101010101010101ull) & (~0ull >> (8 * (7 - y1))) & (~0ull << (8 * y0));
}

// Renders the block at (bx0, by0) of a tile whose pair list is [e0, e1):
// 8x8 with one pixel per lane (K = 1) or 16x16 with four pixels per lane
// (K = 4: lane l owns column l % 16 of rows l / 16 + 4k), so that one walk
// of the list serves 256 pixels.  Returns the covered samples it saw.
// QUAD: the block is quadrant qbit of a 16x16 tile, and only the list
// entries holding that quadrant's bit are staged.
template <int MODE, bool PHONG, int K, int TK, bool QUAD>
__device__ __forceinline__ uint32_t raster_block(const RasterArgs& A, const DFrame& F, const DTileFrame& sg,
                                                 uint32_t frame, int bx0, int by0, uint32_t e0, uint32_t e1,
                                                 WaveQueueT<PHONG>& Q, int qbit) {
    constexpr bool ROT = false;                   // staged records are not rotated (slower, measured)
    constexpr int PM = MODE;                      // 0 packed, 1 in place
    constexpr int TW = (K == 1) ? 8 : 16;
    constexpr int RS = 64 / TW;                    // rows between a lane's pixels
    const int lane = threadIdx.x & 63;
    const int px = bx0 + (lane % TW);
    int py[K];
    bool in[K];
    size_t pi[K];
    uint32_t col[K], win[K];
    float dep[K];
    const size_t npix = (size_t)F.W * F.H;
    const size_t aligned = npix - (npix % 8);
#pragma unroll
    for (int k = 0; k < K; k++) {
        py[k] = by0 + lane / TW + RS * k;
        in[k] = px < (int)F.W && py[k] < (int)F.H;
        pi[k] = in[k] ? (size_t)py[k] * F.W + px : 0;
        col[k] = 0u;
        win[k] = 0u;
        dep[k] = 0.0f;
        if (in[k]) {
            const bool wide = TK == 2 && F.kind >= PFF_TARGET_FIRST_WIDE;
            if (sg.clear && pi[k] >= 8 && pi[k] < aligned) {       // context.c:693-713
                col[k] = sg.clearColor;                              // as the SIMD setter stores it
                dep[k] = sg.clearDepth;
                if (wide) win[k] = 1u;                               // written by the clear
            } else if (sg.clear && pi[k] >= aligned) {              // the tail copies pixel 0
                col[k] = wide ? wide_get_simd((const uint8_t*)&A.clearSnap[8 * frame + 2], F.kind, 0)
                              : A.clearSnap[8 * frame];
                dep[k] = __uint_as_float(A.clearSnap[8 * frame + 1]);
                if (wide) win[k] = WIN_SNAP;
            } else {
                col[k] = wide ? wide_get_simd(F.color, F.kind, pi[k])
                              : (TK == 0 ? reinterpret_cast<const uint32_t*>(F.color)[pi[k]] : tgt_load(F.color, F.kind, pi[k]));
                dep[k] = F.depth[pi[k]];
            }
        }
    }
    // Last column and row of the block inside the target.
    const int xLim = min(bx0 + TW - 1, (int)F.W - 1), yLim = min(by0 + TW - 1, (int)F.H - 1);
    uint32_t frags = 0;
    for (uint32_t e = e0; e < e1; e += QUAD ? 0u : (uint32_t)BATCH) {
        // Up to 32 list entries: the next 32, or the next 32 holding this
        // block's quadrant bit (read 64 at a time, compacted in list order
        // through the queue's record-id slots).
        uint32_t nb, rIn = 0;
        if constexpr (QUAD) {
            nb = 0;
            while (nb < (uint32_t)BATCH && e < e1) {
                const bool has = e + (uint32_t)lane < e1;
                const uint32_t v = has ? A.vals[e + lane] : 0u;
                const bool mine = has && ((v >> (PAIR_QUAD_SH + (uint32_t)qbit)) & 1u);
                const uint64_t M = __builtin_amdgcn_ballot_w64(mine);
                const uint32_t room = (uint32_t)BATCH - nb, found = (uint32_t)__popcll(M);
                const uint32_t rank = (uint32_t)lane_rank(M);
                if (mine && rank < room) Q.recId[nb + rank] = v & PAIR_RT_MASK;
                if (found > room) {
                    // The batch fills here: continue after its last entry.
                    e += (uint32_t)__builtin_ctzll(__builtin_amdgcn_ballot_w64(mine && rank == room - 1u)) + 1u;
                    nb = (uint32_t)BATCH;
                } else {
                    e += min(64u, e1 - e);
                    nb += found;
                }
            }
            wave_sync();
            if ((uint32_t)lane < nb) rIn = Q.recId[lane];
            wave_sync();
        } else {
            nb = min((uint32_t)BATCH, e1 - e);
        }
        // Their records; keep the ones that meet this block, compacted in
        // list order.
        bool ov = false;
        const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
        uint4 q0 = z4, q1 = z4, q2 = z4, q3 = z4, q4 = z4, q5 = z4, q6 = z4, q7 = z4;
        uint32_t r = 0;
        uint64_t rm = 0;
        if ((uint32_t)lane < nb) {
            if constexpr (QUAD) r = rIn;
            else r = A.vals[e + lane] & PAIR_RT_MASK;
            const uint4* src = reinterpret_cast<const uint4*>(&A.rts[r]);
            q0 = src[0]; q1 = src[1]; q2 = src[2]; q3 = src[3];
            q4 = src[4]; q5 = src[5]; q6 = src[6]; q7 = src[7];
            const int32_t xMin = (int32_t)q2.y, yMin = (int32_t)q2.z;
            const int32_t xMax = (int32_t)q2.w, yMax = (int32_t)q3.x;
            ov = !(xMax <= bx0 || xMin > bx0 + TW - 1 || yMax < by0 || yMin > by0 + TW - 1);
            // An edge negative over the whole block leaves it uncovered.  With
            // NOWRAP the block (inside a tile that meets the bbox, so within 64
            // px of it) sees no int32 wrap: its largest w is at a corner.
            if (ov && (q5.x & META_NOWRAP)) {
                const int32_t xs[2] = { bx0, bx0 + TW - 1 }, ys[2] = { by0, by0 + TW - 1 };
                const int32_t eA[3] = { (int32_t)q0.x, (int32_t)q0.y, (int32_t)q0.z };
                const int32_t eB[3] = { (int32_t)q0.w, (int32_t)q1.x, (int32_t)q1.y };
                const int32_t eC[3] = { (int32_t)q1.z, (int32_t)q1.w, (int32_t)q2.x };
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int32_t hi = eA[k] * xs[eA[k] > 0] + eB[k] * ys[eB[k] > 0] + eC[k];
                    ov = ov && hi >= 0;
                }
            }
            if (ov && K == 1) {
                // The coverage rule's bbox part (x in [xMin, xMax), y in
                // [yMin, yMax]) and the target, as a pixel mask: for exact
                // triangles the edges imply it, for the others it is their
                // bbox test.  An empty mask drops the record.
                const int32_t rx0 = max(bx0, xMin), rx1 = min(xLim, xMax - 1);
                const int32_t ry0 = max(by0, yMin), ry1 = min(yLim, yMax);
                if (rx0 > rx1 || ry0 > ry1) ov = false;
                else rm = rect_mask8(rx0 - bx0, rx1 - bx0, ry0 - by0, ry1 - by0);
            }
        }
        const uint64_t OV = __builtin_amdgcn_ballot_w64(ov);
        const int nt = __popcll(OV);
        // 24-bit edge multiplies for the whole batch when every staged
        // record allows them (decided before the loop: no per-record branch).
        const bool allSmall = __builtin_amdgcn_ballot_w64(ov && !(q5.x & META_SMALL)) == 0ull;
        if (ov) {
            const int slot = lane_rank(OV);
            const uint4 q[8] = { q0, q1, q2, q3, q4, q5, q6, q7 };
            stage_rec<ROT>(Q, slot, q);
            Q.recId[slot] = r;
            if (K == 1) Q.rmask[slot] = rm;
            if constexpr (PHONG) {
                if (q5.x & META_PHONG) {
                    const uint4* ls = reinterpret_cast<const uint4*>(&A.rtl[r]);
                    uint4* ld = reinterpret_cast<uint4*>(&Q.recL[slot][0]);
#pragma unroll
                    for (int k = 0; k < 5; k++) ld[k] = ls[k];
                }
            }
        }
        wave_sync();
        int nf = 0;
        uint64_t pm[K];                    // per owned pixel: its queue slots
#pragma unroll
        for (int k = 0; k < K; k++) pm[k] = 0ull;
        for (int j = 0; j < nt; j++) {
            // Edges (chunks 0..2), for 16x16 blocks also bbox and meta; the
            // whole record when this loop shades (in place).
            RTri t;
            if (PM == 1) t = staged_rec<ROT, 8>(Q, j);
            else if (K == 1) t = staged_rec<ROT, 3>(Q, j);
            else { t = staged_rec<ROT, 4>(Q, j); t.meta = staged_meta<ROT>(Q, j); }
            // A lane's pixels are RS rows apart: each edge value steps by
            // RS * B, and stepping mod 2^32 gives the direct value exactly.
            int32_t w1, w2, w3;
            edge_values_as(t, allSmall, px, py[0], w1, w2, w3);
            const uint32_t d1 = (uint32_t)t.B[0] * RS, d2 = (uint32_t)t.B[1] * RS, d3 = (uint32_t)t.B[2] * RS;
            // Branch-free bbox test (K = 4; K = 1 uses the staged mask):
            // every field is read up front.
            const uint64_t rmj = (K == 1) ? Q.rmask[j] : 0ull;
            const bool exact = (t.meta & META_EXACT) != 0;
            const int32_t xMin = t.xMin, yMin = t.yMin, xMax = t.xMax, yMax = t.yMax;
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (k > 0) {
                    w1 = (int32_t)((uint32_t)w1 + d1);
                    w2 = (int32_t)((uint32_t)w2 + d2);
                    w3 = (int32_t)((uint32_t)w3 + d3);
                }
                const bool inBox = (K == 1) ? (((rmj >> lane) & 1ull) != 0)
                                            : (in[k] & (px < xMax) & (exact | ((px >= xMin) & (py[k] >= yMin) & (py[k] <= yMax))));
                const bool inEdges = (w1 | w2 | w3) > 0;
                const bool cov = inBox & inEdges;
                // The two ballots, each of a comparison, and their AND in
                // scalar registers: a ballot of the combined flag rebuilds
                // its mask through two vector instructions.
                const uint64_t M = __builtin_amdgcn_ballot_w64(inEdges) & __builtin_amdgcn_ballot_w64(inBox);
                if (!M) continue;
                const int c = __popcll(M);
                frags += (uint32_t)c;
                if (PM == 1) {
                    if (cov) {
                        uint32_t fc;
                        float fz;
                        if (shade_fragment<PHONG, true, true>(A, t, staged_light<PHONG>(Q, j), px, py[k], fc, fz, dep[k],
                                                              w1, w2, w3))
                            rop<TK>(t.meta, fc, fz, Q.recId[j], F.kind, col[k], dep[k], win[k]);
                    }
                    continue;
                }
                // Queued after the fragments before them; a full queue is
                // flushed first.  That rare branch takes the lane's coverage
                // from the ballot, so `cov` need not be kept across the flush
                // (which cost instructions on every record).
                const auto enqueue = [&](int base) {
                    const int slot = base + lane_rank(M);
                    Q.ftri[slot] = (uint8_t)j;
                    Q.fpix[slot] = (uint8_t)(lane + 64 * k);
                    pm[k] |= 1ull << slot;
                };
                if (nf + c > 64) {
                    flush_queue<PHONG, K, ROT, TK>(A, Q, bx0, by0, lane, nf, pm, F.kind, col, dep, win);
                    nf = 0;
                    if ((M >> lane) & 1ull) enqueue(0);
                } else if (cov) {
                    enqueue(nf);
                }
                nf += c;
            }
        }
        if (PM == 0 && nf > 0) flush_queue<PHONG, K, ROT, TK>(A, Q, bx0, by0, lane, nf, pm, F.kind, col, dep, win);
        wave_sync();
    }
#pragma unroll
    for (int k = 0; k < K; k++)
        if (in[k]) {
            if (TK == 0) {
                reinterpret_cast<uint32_t*>(F.color)[pi[k]] = col[k];
            } else if (TK == 1 || F.kind < PFF_TARGET_FIRST_WIDE) {
                tgt_store(F.color, F.kind, pi[k], col[k]);
            } else if (win[k] == WIN_SNAP) {                         // pixel 0's bytes at the clear
                const uint32_t bpp = F.bpp;
                const uint8_t* snap = (const uint8_t*)&A.clearSnap[8 * frame + 2];
                for (uint32_t b = 0; b < bpp; b++) F.color[pi[k] * bpp + b] = snap[b];
            } else if (win[k] != 0u) {
                wide_set_simd(F.color, F.kind, pi[k], col[k]);
            }
            F.depth[pi[k]] = dep[k];
            if (A.winner) A.winner[pi[k]] = win[k];
        }
    return frags;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i mod 8), and each XCD has its own L2.  This maps the workgroups of each
// XCD to chunks of XCD_CHUNK consecutive work items, the chunks themselves
// dealt round-robin: neighbouring tiles -- which share raster-triangle
// records -- mostly run on one XCD and read them through one L2, while
// every XCD still gets work from all over the frame (one contiguous band
// per XCD left some XCDs with most of the dense tiles).  A bijection of
// [0, n): the last n mod (8 XCD_CHUNK) items keep their index.
constexpr uint32_t XCD_CHUNK = 16;

__device__ __forceinline__ uint32_t xcd_chunked(uint32_t i, uint32_t n) {
    const uint32_t span = 8 * XCD_CHUNK, full = n - n % span;
    if (i >= full) return i;
    const uint32_t x = i % 8, j = i / 8;
    return ((j / XCD_CHUNK) * 8 + x) * XCD_CHUNK + j % XCD_CHUNK;
}

// Four waves per workgroup.  A wave renders blocks of BW x BW pixels
// (BW = 8 for K = 1, 16 for K = 4).  When a tile is one block, each wave
// takes its own tile (four tiles per workgroup); otherwise one tile per
// workgroup, its blocks dealt round-robin to the four waves.
template <int TS, int MODE, bool PHONG, int K, int TK>
__device__ __forceinline__ void raster_tiles(const RasterArgs& A) {
    typedef WaveQueueT<PHONG> Queue;
    __shared__ Queue Qs[4];
    constexpr int BW = (K == 1) ? 8 : 16;
    static_assert(TS >= BW, "tile smaller than a block");
    constexpr int BPR = TS / BW;
    constexpr int NB = BPR * BPR;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Queue& Q = Qs[wave];
    const uint32_t wg = xcd_chunked(blockIdx.x, gridDim.x);
    const uint32_t tile = (NB == 1) ? wg * 4 + wave : wg;
    if (tile >= A.numTiles) return;
    uint32_t* fragOut = A.tileFrags + (size_t)tile * 4 + (NB == 1 ? 0 : wave);
    const bool lead = (threadIdx.x & 63) == 0;
    const uint32_t frame = A.tileFrame[tile];
    const DTileFrame sg = A.tileDesc[frame];
    if (!sg.active) { if (lead) *fragOut = 0u; return; }
    const DFrame& F = sg.F;
    const uint32_t e0 = A.tileBegin[tile], e1 = A.tileEnd[tile];
    if (e0 == e1 && !sg.clear) { if (lead) *fragOut = 0u; return; }
    const uint32_t local = tile - F.tileBase;
    const int tx = (int)(local % F.tilesX), ty = (int)(local / F.tilesX);
    uint32_t frags = 0;
    for (int b = (NB == 1 ? 0 : wave); b < NB; b += (NB == 1 ? 1 : 4)) {
        const int bx0 = tx * TS + (b % BPR) * BW, by0 = ty * TS + (b / BPR) * BW;
        if (bx0 >= (int)F.W || by0 >= (int)F.H) continue;
        // Packed mode at 16x16 tiles stages each quadrant's own entries
        // (1080p frame: raster 100 -> 95 us; 4K frame 548 -> 484 us); in
        // place, whose large triangles meet every quadrant, it only costs.
        constexpr bool QUAD = TS == 16 && K == 1 && MODE == 0;
        frags += raster_block<MODE, PHONG, K, TK, QUAD>(A, F, sg, frame, bx0, by0, e0, e1, Q, b);
    }
    if (lead) *fragOut = frags;
}

template <int TS, int MODE, bool PHONG, int K, int TK>
__global__ __launch_bounds__(256) void k_raster(RasterArgs A) {
    raster_tiles<TS, MODE, PHONG, K, TK>(A);
}

// The same for RGBA8 jobs without per-fragment lighting, held to 64 VGPRs
// so that a SIMD holds 8 waves (as many as the LDS allows): 1080p frame
// -2 %, 256-context batch -3 %.  Only the instances that fit without
// spilling use it (launch_raster).
template <int TS, int MODE, int K>
__global__ __launch_bounds__(256, 8) void k_raster_o8(RasterArgs A) {
    raster_tiles<TS, MODE, false, K, 0>(A);
}

// ---------------------------------------------------------------------------
// Differentiable shading: the float-domain forward and its adjoint
//
// Coverage and the depth-test winner of every pixel come from the integer
// pipeline above (the raster kernel records 1 + the raster triangle of the
// last fragment written).  Each pixel's winning fragment is then shaded
// again without quantisation:
//   b_k      = w_k * wInvSum (the raster's own barycentrics)
//   c_k      = sum_j M[k][j] * C_j  (recorded-triangle corner colours C_j:
//              parameters where the corner has a colour-array index, the
//              recorded RGBA8 / 255 otherwise), times material diffuse / 255
//              when lighting is on (triangles.c:88-91)
//   frag     = sum_k b_k c_k
//   texture  = sum of texel taps of the reference index maps with weights
//              (1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy; frag *= texel
//   Phong    = lighting.c:148-258 with exact sqrt, division and pow:
//              out.rgb = clamp(sum_l amb_l + kd_l ld_l D + ks_l ls_l S, 0, 1),
//              out.a = 0; amb/kd/ks hold attenuation and spot factors
// Background pixels take the target's RGBA8 / 255.  The adjoint recomputes
// the same quantities and scatters dL/dimage into the texel, vertex-colour
// and light parameters (coverage fixed: edges stay hard).
// ---------------------------------------------------------------------------

constexpr int DIFF_LIGHT_VALS = 2 * PFF_MAX_LIGHTS * 3;   // diffuse then specular, by light number

struct DiffArgs {
    const RTri*     rts;
    const RTriL*    rtl;
    const RTriD*    rtd;
    const uint32_t* winner;
    const uint32_t* color;          // RGBA8 target after the integer pass
    const uint32_t* srcCol;         // recorded corner colours
    const uint32_t* vid;            // corner colour-array indices, or null
    const PFFstate* states;
    const DTex*     texs;
    int32_t         paramTex;       // texture whose texels are parameters, -1 for none
    const float*    texParams;      // [h][w][4]
    const float*    colParams;      // [numVertices][4], or null
    uint32_t        numVertices;
    const float*    lightParams;    // [DIFF_LIGHT_VALS], or null: recorded colours / 255
    uint32_t        W, H;
    float*          image;          // forward output [H][W][4]
    const float*    dImage;         // adjoint input [H][W][4]
    float*          gTex;           // adjoint outputs, accumulated; each may be null
    float*          gCol;
    float*          gLightPart;     // [blocks][DIFF_LIGHT_VALS]
};

struct Taps {
    int32_t off[4];
    float   w[4];
    int     n;
};

__device__ __forceinline__ int32_t clamp_off(const DTex& t, int32_t off) {
    const int64_t n = (int64_t)t.w * t.h;
    int64_t o = off;
    return (int32_t)(o < 0 ? 0 : (o >= n ? n - 1 : o));
}

// Texel offsets and weights of one lookup, with the index maps of
// sample_tex (sampler.h:202-334) and the same out-of-texture clamp.
__device__ __forceinline__ void tex_taps(const DTex& t, float u, float v, Taps& tp) {
    const uint32_t w = t.w;
    if (t.filter == 0) {
        int32_t x = wrap_coord(t.wrap, u, t.w), y = wrap_coord(t.wrap, v, t.h);
        tp.n = 1;
        tp.off[0] = clamp_off(t, (int32_t)((uint32_t)y * w + (uint32_t)x));
        tp.w[0] = 1.0f;
        return;
    }
    float tx = rcp_int_rn(t.w), ty = rcp_int_rn(t.h);   // texture sides are at most 65535
    int32_t x0 = wrap_coord(t.wrap, u, t.w), y0 = wrap_coord(t.wrap, v, t.h);
    int32_t x1 = wrap_coord(t.wrap, u + tx, t.w), y1 = wrap_coord(t.wrap, v + ty, t.h);
    float fx = clamp_ps(u * (float)t.w - (float)x0, 0.0f, 1.0f);
    float fy = clamp_ps(v * (float)t.h - (float)y0, 0.0f, 1.0f);
    tp.n = 4;
    tp.off[0] = clamp_off(t, (int32_t)((uint32_t)y0 * w + (uint32_t)x0));
    tp.off[1] = clamp_off(t, (int32_t)((uint32_t)y0 * w + (uint32_t)x1));
    tp.off[2] = clamp_off(t, (int32_t)((uint32_t)y1 * w + (uint32_t)x0));
    tp.off[3] = clamp_off(t, (int32_t)((uint32_t)y1 * w + (uint32_t)x1));
    tp.w[0] = (1.0f - fx) * (1.0f - fy);
    tp.w[1] = fx * (1.0f - fy);
    tp.w[2] = (1.0f - fx) * fy;
    tp.w[3] = fx * fy;
}

__device__ __forceinline__ float4 texel_f(const DiffArgs& A, const DTex& t, bool param, int32_t off) {
    if (param) return reinterpret_cast<const float4*>(A.texParams)[off];
    if (t.format == PFF_TEX_RGBA8) {
        const uint32_t c = ((const uint32_t*)t.pixels)[off];
        const float k = 1.0f / 255.0f;
        return make_float4((float)ch(c, 0) * k, (float)ch(c, 1) * k, (float)ch(c, 2) * k, (float)ch(c, 3) * k);
    }
    return ((const float4*)t.pixels)[off];
}

__device__ __forceinline__ float light_param(const DiffArgs& A, const PFFlight& Lt, int which, int c) {
    if (A.lightParams) return A.lightParams[which * PFF_MAX_LIGHTS * 3 + (Lt.id & 7u) * 3 + c];
    return (float)ch(which ? Lt.specular : Lt.diffuse, c) * (1.0f / 255.0f);
}

// Per-light factors at one fragment: amb[c] = att * la/255 * ambient/255 (to
// multiply D), kd = att * I * max(N.L, 0), ks = att * I * max(N.H, 0)^shininess.
struct LightTerms {
    float amb[3];
    float kd, ks;
};

__device__ LightTerms light_terms(const PFFlight& Lt, const PFFmaterial& mat, const float* P, const float* N,
                                  const float* V) {
    LightTerms r;
    float L[3], d2 = 0.0f;
    for (int i = 0; i < 3; i++) { L[i] = Lt.position[i] - P[i]; d2 += L[i] * L[i]; }
    const float il = 1.0f / sqrtf(fmaxf(d2, 1e-5f));
    for (int i = 0; i < 3; i++) L[i] *= il;
    float nl = fmaxf(N[0] * L[0] + N[1] * L[1] + N[2] * L[2], 0.0f);
    float H[3] = { L[0] + V[0], L[1] + V[1], L[2] + V[2] };
    float hs = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
    float ih = hs > 0.0f ? 1.0f / sqrtf(hs) : 0.0f;
    float nh = fmaxf((N[0] * H[0] + N[1] * H[1] + N[2] * H[2]) * ih, 0.0f);
    float spec = powf(nh, mat.shininess);
    float I = 1.0f, att = 1.0f;
    if ((double)Lt.innerCutOff < PF_PI_D) {
        float theta = -(L[0] * Lt.direction[0] + L[1] * Lt.direction[1] + L[2] * Lt.direction[2]);
        I = fminf(fmaxf((theta - Lt.outerCutOff) / (Lt.innerCutOff - Lt.outerCutOff), 0.0f), 1.0f);
    }
    if (Lt.attLinear != 0.0f || Lt.attQuadratic != 0.0f) {
        float d0 = Lt.position[0] - P[0], d1 = Lt.position[1] - P[1];
        float dsq = d0 * d0 + 2.0f * d1 * d1;                 // simd.h:2427-2429 reads y twice
        att = 1.0f / (Lt.attConstant + Lt.attLinear * sqrtf(dsq) + Lt.attQuadratic * dsq);
    }
    for (int c = 0; c < 3; c++)
        r.amb[c] = att * ((float)ch(Lt.ambient, c) / 255.0f) * ((float)ch(mat.ambient, c) / 255.0f);
    r.kd = att * I * nl;
    r.ks = att * I * spec;
    return r;
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// One thread per pixel, 16x16 pixels per block.  BWD = false writes the
// image; BWD = true accumulates the parameter gradients.
template <bool BWD>
__global__ __launch_bounds__(256) void k_diff(DiffArgs A) {
    __shared__ float lg[DIFF_LIGHT_VALS];
    const int tid = threadIdx.x, lane = tid & 63;
    if (BWD) {
        for (int i = tid; i < DIFF_LIGHT_VALS; i += 256) lg[i] = 0.0f;
        __syncthreads();
    }
    const uint32_t px = blockIdx.x * 16 + (tid & 15), py = blockIdx.y * 16 + (tid >> 4);
    const bool in = px < A.W && py < A.H;
    const size_t pix = in ? (size_t)py * A.W + px : 0;
    const uint32_t win = in ? A.winner[pix] : 0u;
    const bool cov = win != 0u;

    float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (BWD && cov) {
        const float4 d = reinterpret_cast<const float4*>(A.dImage)[pix];
        g[0] = d.x; g[1] = d.y; g[2] = d.z; g[3] = d.w;
    }
    // Fragment inputs
    const PFFstate* st = nullptr;
    const PFFmaterial* mat = nullptr;
    float beta[3] = {0.0f, 0.0f, 0.0f}, m[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float cf[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tex[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float N[3] = {0.0f, 0.0f, 0.0f}, P[3] = {0.0f, 0.0f, 0.0f}, V[3] = {0.0f, 0.0f, 0.0f};
    Taps tp;
    tp.n = 0;
    bool texOn = false, texParam = false, phong = false;
    uint32_t src = 0, nl = 0;
    DTex T;
    if (cov) {
        const uint32_t r = win - 1u;
        const RTri& t = A.rts[r];
        const RTriD& d = A.rtd[r];
        int32_t w1, w2, w3;
        edge_values(t, (int)px, (int)py, w1, w2, w3);
        const float b[3] = { (float)w1 * t.wInvSum, (float)w2 * t.wInvSum, (float)w3 * t.wInvSum };
        st = &A.states[t.state];
        mat = &st->material[d.face & 1u];
        src = d.src;
        for (int j = 0; j < 3; j++) beta[j] = b[0] * d.M[j] + b[1] * d.M[3 + j] + b[2] * d.M[6 + j];
        if (st->flags & PFF_LIGHTING)
            for (int c = 0; c < 4; c++) m[c] = (float)ch(mat->diffuse, c) / 255.0f;
        for (int j = 0; j < 3; j++) {
            const size_t ci = (size_t)src * 3 + j;
            const uint32_t vi = A.vid ? A.vid[ci] : PFF_NO_VERTEX;
            float C[4];
            if (A.colParams && vi < A.numVertices) {
                const float4 q = reinterpret_cast<const float4*>(A.colParams)[vi];
                C[0] = q.x; C[1] = q.y; C[2] = q.z; C[3] = q.w;
            } else {
                const uint32_t cc = A.srcCol[ci];
                for (int c = 0; c < 4; c++) C[c] = (float)ch(cc, c) / 255.0f;
            }
            for (int c = 0; c < 4; c++) cf[c] += beta[j] * (C[c] * m[c]);
        }
        if (t.meta & META_TEX) {
            texOn = true;
            T = A.texs[st->texture];
            texParam = st->texture == A.paramTex;
            float u = (t.uv[0] * b[0] + t.uv[2] * b[1]) + t.uv[4] * b[2];
            float v = (t.uv[1] * b[0] + t.uv[3] * b[1]) + t.uv[5] * b[2];
            if (t.meta & META_3D) {
                const float zq = __fdiv_rn(1.0f, (t.z[0] * b[0] + t.z[1] * b[1]) + t.z[2] * b[2]);
                u = u * zq;
                v = v * zq;
            }
            tex_taps(T, u, v, tp);
            for (int c = 0; c < 4; c++) tex[c] = 0.0f;
            for (int k = 0; k < tp.n; k++) {
                const float4 q = texel_f(A, T, texParam, tp.off[k]);
                tex[0] += tp.w[k] * q.x; tex[1] += tp.w[k] * q.y; tex[2] += tp.w[k] * q.z; tex[3] += tp.w[k] * q.w;
            }
        }
        if (t.meta & META_PHONG) {
            phong = true;
            nl = st->numLights;
            const RTriL& l = A.rtl[r];
            for (int i = 0; i < 3; i++) {
                N[i] = l.n[i] * b[0] + l.n[3 + i] * b[1] + l.n[6 + i] * b[2];
                P[i] = l.p[i] * b[0] + l.p[3 + i] * b[1] + l.p[6 + i] * b[2];
                V[i] = st->viewPos[i] - P[i];
            }
            const float iv = 1.0f / sqrtf(fmaxf(V[0] * V[0] + V[1] * V[1] + V[2] * V[2], 1e-5f));
            for (int i = 0; i < 3; i++) V[i] *= iv;
        }
    }
    float frag[4];
    for (int c = 0; c < 4; c++) frag[c] = cf[c] * tex[c];
    float dfrag[4] = { g[0], g[1], g[2], g[3] };
    float gcs[3] = {0.0f, 0.0f, 0.0f}, S[3] = {0.0f, 0.0f, 0.0f};
    if (phong) {
        float D[3], acc[3] = {0.0f, 0.0f, 0.0f};
        for (int c = 0; c < 3; c++) { D[c] = frag[c]; S[c] = (float)ch(mat->specular, c) / 255.0f; }
        for (uint32_t li = 0; li < nl; li++) {
            const PFFlight& Lt = st->lights[li];
            const LightTerms lt = light_terms(Lt, *mat, P, N, V);
            for (int c = 0; c < 3; c++)
                acc[c] += lt.amb[c] * D[c] + lt.kd * light_param(A, Lt, 0, c) * D[c] + lt.ks * light_param(A, Lt, 1, c) * S[c];
        }
        for (int c = 0; c < 3; c++) out[c] = fminf(fmaxf(acc[c], 0.0f), 1.0f);
        out[3] = 0.0f;
        if (BWD) {
            float dD[3] = {0.0f, 0.0f, 0.0f};
            for (int c = 0; c < 3; c++) gcs[c] = (acc[c] > 0.0f && acc[c] < 1.0f) ? g[c] : 0.0f;
            for (uint32_t li = 0; li < nl; li++) {
                const PFFlight& Lt = st->lights[li];
                const LightTerms lt = light_terms(Lt, *mat, P, N, V);
                for (int c = 0; c < 3; c++) dD[c] += gcs[c] * (lt.amb[c] + lt.kd * light_param(A, Lt, 0, c));
            }
            for (int c = 0; c < 3; c++) dfrag[c] = dD[c];
            dfrag[3] = 0.0f;
        }
    } else {
        for (int c = 0; c < 4; c++) out[c] = frag[c];
    }
    if (!BWD) {
        if (in) {
            if (!cov) {
                const uint32_t cc = A.color[pix];
                for (int c = 0; c < 4; c++) out[c] = (float)ch(cc, c) / 255.0f;
            }
            reinterpret_cast<float4*>(A.image)[pix] = make_float4(out[0], out[1], out[2], out[3]);
        }
        return;
    }
    // Light parameters: per light slot, a wave sum per light number, then
    // one LDS add per wave and value.
    const bool lit = phong && A.gLightPart;
    uint32_t maxL = lit ? nl : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) maxL = max(maxL, (uint32_t)__shfl_xor((int)maxL, d, 64));
    maxL = __builtin_amdgcn_readfirstlane(maxL);
    for (uint32_t li = 0; li < maxL; li++) {
        const bool has = lit && li < nl;
        uint32_t id = 0;
        float v[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (has) {
            const PFFlight& Lt = st->lights[li];
            id = Lt.id & 7u;
            const LightTerms lt = light_terms(Lt, *mat, P, N, V);
            for (int c = 0; c < 3; c++) {
                v[c] = gcs[c] * lt.kd * frag[c];
                v[3 + c] = gcs[c] * lt.ks * S[c];
            }
        }
        uint64_t pend = __builtin_amdgcn_ballot_w64(has);
        while (pend) {
            const uint32_t lead = __builtin_amdgcn_readlane(id, __builtin_ctzll(pend));
            const bool mine = has && id == lead;
            for (int c = 0; c < 6; c++) {
                const float x = wave_sum(mine ? v[c] : 0.0f);
                if (lane == 0) atomicAdd(&lg[(c < 3 ? 0 : PFF_MAX_LIGHTS * 3) + lead * 3 + (c % 3)], x);
            }
            pend &= ~__builtin_amdgcn_ballot_w64(mine);
        }
    }
    // Texels
    if (cov) {
        if (texOn && texParam && A.gTex) {
            for (int k = 0; k < tp.n; k++) {
                float* gp = A.gTex + (size_t)tp.off[k] * 4;
                for (int c = 0; c < 4; c++) {
                    const float x = tp.w[k] * dfrag[c] * cf[c];
                    if (x != 0.0f) unsafeAtomicAdd(gp + c, x);
                }
            }
        }
    }
    // Vertex colours: neighbouring pixels share triangles and vertices, so
    // each wave first sums the contributions per vertex (one atomic per
    // vertex and channel instead of one per pixel).
    if (A.colParams && A.gCol && A.vid) {
        float dcf[4];
        for (int c = 0; c < 4; c++) dcf[c] = dfrag[c] * tex[c];
        for (int j = 0; j < 3; j++) {
            const uint32_t vi = cov ? A.vid[(size_t)src * 3 + j] : PFF_NO_VERTEX;
            const bool has = vi < A.numVertices;
            float x[4];
            for (int c = 0; c < 4; c++) x[c] = has ? beta[j] * m[c] * dcf[c] : 0.0f;
            uint64_t pend = __builtin_amdgcn_ballot_w64(has);
            while (pend) {
                const uint32_t lead = __builtin_amdgcn_readlane(vi, __builtin_ctzll(pend));
                const bool mine = has && vi == lead;
                pend &= ~__builtin_amdgcn_ballot_w64(mine);
                float sum[4];
                for (int c = 0; c < 4; c++) sum[c] = wave_sum(mine ? x[c] : 0.0f);
                if (lane == 0)
                    for (int c = 0; c < 4; c++)
                        if (sum[c] != 0.0f) unsafeAtomicAdd(A.gCol + (size_t)lead * 4 + c, sum[c]);
            }
        }
    }
    __syncthreads();
    if (A.gLightPart) {
        const uint32_t blk = blockIdx.y * gridDim.x + blockIdx.x;
        for (int i = tid; i < DIFF_LIGHT_VALS; i += 256) A.gLightPart[(size_t)blk * DIFF_LIGHT_VALS + i] = lg[i];
    }
}

// Sums the per-block light partials in a fixed order.
__global__ __launch_bounds__(256) void k_diff_light_sum(const float* __restrict__ part, uint32_t blocks,
                                                        float* __restrict__ out) {
    __shared__ float red[256];
    const uint32_t i = blockIdx.x;          // one block per value
    float acc = 0.0f;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256) acc += part[(size_t)b * DIFF_LIGHT_VALS + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
}

// ---------------------------------------------------------------------------
// Texels of the other source formats, decoded once to RGBA8 as the
// reference's AVX2 texel getters read them (pixel.h:2250-3040).  Each getter
// is a function of the texture bytes alone, so decoding at upload and
// sampling RGBA8 gives the texels the reference's sampler sees.  The getters
// gather 32-bit words at byte offset (texel index x a stride); the
// single-byte formats keep the top byte of that word (the texel three
// ahead), LUMINANCE_ALPHA bytes the third and fourth (the next texel), ALPHA
// bytes leaves red, green and blue at 0, and BGR half / float pack their
// first component into red.  Half and float components are converted by
// cvt_rne(v * 255) with no clamp, ORed into place.  Bytes past the last
// texel read as 0 (the source copy carries 4 zero bytes).
// ---------------------------------------------------------------------------

// pfmHalfToFloatI (pfm.h:128-145): subnormals flush to signed zero;
// infinities and NaNs keep their payload.
__device__ __forceinline__ float half_to_float_ref(uint32_t h) {
    const uint32_t s = (h & 0x8000u) << 16;
    const int32_t em = (int32_t)(h & 0x7FFFu);
    int32_t r = (em + (112 << 10)) << 13;
    r = em < (1 << 10) ? 0 : r;
    r += em >= (31 << 10) ? (112 << 23) : 0;
    return __uint_as_float(s | (uint32_t)r);
}

// Little-endian 32-bit word at byte offset o (any alignment).
__device__ __forceinline__ uint32_t src_word(const uint8_t* __restrict__ p, size_t o) {
    return (uint32_t)p[o] | ((uint32_t)p[o + 1] << 8) | ((uint32_t)p[o + 2] << 16) | ((uint32_t)p[o + 3] << 24);
}
__device__ __forceinline__ uint32_t c8_half(const uint8_t* __restrict__ p, size_t o) {
    return (uint32_t)cvt_rne_i32(half_to_float_ref(src_word(p, o) & 0xFFFFu) * 255.0f);
}
__device__ __forceinline__ uint32_t c8_float(const uint8_t* __restrict__ p, size_t o) {
    return (uint32_t)cvt_rne_i32(__uint_as_float(src_word(p, o)) * 255.0f);
}

__device__ uint32_t decode_source_texel(const uint8_t* __restrict__ p, size_t i, uint32_t source) {
    const uint32_t f = source & 0xFFu, t = (source >> 8) & 0xFFu;
    const uint32_t A = 0xFF000000u;
    if (t == 0u) {                                                    // UNSIGNED_BYTE
        switch (f) {
        case 0u: return (src_word(p, i) >> 24) | A;                   // pixel.h:2389-2397
        case 1u: return ((src_word(p, i) >> 24) << 8) | A;            // :2400-2407
        case 2u: return ((src_word(p, i) >> 24) << 16) | A;           // :2410-2417
        case 3u: return src_word(p, i) << 24;                         // :2420-2424
        case 4u: { const uint32_t g = src_word(p, i) >> 24;           // :2253-2263
                   return A | (g << 16) | (g << 8) | g; }
        case 5u: { const uint32_t x = src_word(p, 2 * i), g = (x >> 16) & 0xFFu;   // :2309-2325
                   return (x & 0xFF000000u) | (g << 16) | (g << 8) | g; }
        case 6u: return (src_word(p, 3 * i) & 0x00FFFFFFu) | A;       // :2604-2632
        case 8u: { const uint32_t x = src_word(p, 3 * i);             // :2635-2663
                   return A | ((x & 0xFFu) << 16) | (x & 0xFF00u) | ((x >> 16) & 0xFFu); }
        case 9u: { const uint32_t x = src_word(p, 4 * i);             // :2916-2920
                   return (x & 0xFF00FF00u) | ((x & 0xFFu) << 16) | ((x >> 16) & 0xFFu); }
        default: return 0u;
        }
    }
    if (t == 9u || t == 10u) {                                        // HALF_FLOAT, FLOAT
        const bool h = t == 9u;
        const size_t cs = h ? 2 : 4;                                  // component stride
        auto c = [&](size_t o) { return h ? c8_half(p, o) : c8_float(p, o); };
        switch (f) {
        case 0u: return c(cs * i) | A;                                // :2427-2435, :2480-2490
        case 1u: return (c(cs * i) << 8) | A;                         // :2440-2451, :2493-2504
        case 2u: return (c(cs * i) << 16) | A;                        // :2454-2465, :2507-2518
        case 3u: return c(cs * i) << 24;                              // :2468-2477, :2521-2530
        case 4u: { const uint32_t l = c(cs * i);                      // :2266-2283, :2286-2303
                   return A | (l << 16) | (l << 8) | l; }
        case 5u: { const uint32_t l = c(2 * cs * i), a = c(2 * cs * i + cs);       // :2328-2354, :2357-2383
                   return (a << 24) | (l << 16) | (l << 8) | l; }
        case 6u: case 8u: {                                           // :2666-2694, :2697-2725, :2728-2790
                   const size_t o = 3 * cs * i;
                   return A | (c(o + 2 * cs) << 16) | (c(o + cs) << 8) | c(o); }
        case 7u: { const size_t o = 4 * cs * i;                       // :2923-2950 (RGBA float is PFF_TEX_RGBAF32)
                   return (c(o + 3 * cs) << 24) | (c(o + 2 * cs) << 16) | (c(o + cs) << 8) | c(o); }
        case 9u: { const size_t o = 4 * cs * i;                       // :2953-2980, :3013-3040
                   return (c(o + 3 * cs) << 24) | (c(o) << 16) | (c(o + cs) << 8) | c(o + 2 * cs); }
        default: return 0u;
        }
    }
    const uint32_t x = src_word(p, 2 * i);                            // packed 16-bit texels
    if (t == 2u) {                                                    // 5_6_5, :2536-2601
        const uint32_t hi = ((x & 0xF800u) >> 11) << 3, g = ((x & 0x07E0u) >> 5) << 2, lo = (x & 0x1Fu) << 3;
        return f == 6u ? (A | (lo << 16) | (g << 8) | hi) : (A | (hi << 16) | (g << 8) | lo);
    }
    if (t == 3u) {                                                    // 5_5_5_1, :2794-2849
        const uint32_t c0 = ((x >> 11) & 0x1Fu) * 8u, c1 = ((x >> 6) & 0x1Fu) * 8u, c2 = ((x >> 1) & 0x1Fu) * 8u;
        const uint32_t a = (x & 1u) * 255u;
        return f == 7u ? ((a << 24) | (c2 << 16) | (c1 << 8) | c0) : ((a << 24) | (c0 << 16) | (c1 << 8) | c2);
    }
    if (t == 4u) {                                                    // 4_4_4_4, :2852-2907
        const uint32_t c0 = ((x >> 12) & 0xFu) * 17u, c1 = ((x >> 8) & 0xFu) * 17u, c2 = ((x >> 4) & 0xFu) * 17u;
        const uint32_t a = (x & 0xFu) * 17u;
        return f == 7u ? ((a << 24) | (c2 << 16) | (c1 << 8) | c0) : ((a << 24) | (c0 << 16) | (c1 << 8) | c2);
    }
    return 0u;
}

// One thread per texel; src holds n texels of the source pair plus 4 bytes.
__global__ __launch_bounds__(256) void k_decode_texels(const uint8_t* __restrict__ src, size_t n, uint32_t source,
                                                       uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = decode_source_texel(src, i, source);
}

// ---------------------------------------------------------------------------
// pfDrawPixels (context.c:1987-2080): the image is decoded to RGBA8 with the
// scalar getters (pixel.h:405-760) when the job is created; after the raster
// kernel of the pass that holds the call, one thread per destination pixel
// runs the scalar depth test (depth.h:29-63, NOTEQUAL a real !=), writes
// the depth and the scalar blend (blend.h:29-118) of the source texel.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t pack4(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return (r & 255u) | ((g & 255u) << 8) | ((b & 255u) << 16) | ((a & 255u) << 24);
}

__device__ uint32_t decode_scalar_texel(const uint8_t* __restrict__ p, size_t i, uint32_t source) {
    const uint32_t f = source & 0xFFu, t = (source >> 8) & 0xFFu;
    if (t == 0u) {                                                    // UNSIGNED_BYTE
        switch (f) {
        case 0u: return pack4(p[i], 0, 0, 255);                      // pixel.h:462-487
        case 1u: return pack4(0, p[i], 0, 255);
        case 2u: return pack4(0, 0, p[i], 255);
        case 3u: return pack4(255, 255, 255, p[i]);
        case 4u: return pack4(p[i], p[i], p[i], 255);                 // :410-414
        case 5u: return pack4(p[2 * i], p[2 * i], p[2 * i], p[2 * i + 1]);          // :434-438
        case 6u: return pack4(p[3 * i], p[3 * i + 1], p[3 * i + 2], 255);           // :575-587
        case 8u: return pack4(p[3 * i + 2], p[3 * i + 1], p[3 * i], 255);
        case 7u: return pack4(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);  // :696-707
        case 9u: return pack4(p[4 * i + 2], p[4 * i + 1], p[4 * i], p[4 * i + 3]);
        default: return 0u;
        }
    }
    if (t == 9u || t == 10u) {                                        // HALF_FLOAT, FLOAT: (PFubyte)(v * 255)
        const uint32_t comps = f <= 4u ? 1u : f == 5u ? 2u : (f == 6u || f == 8u) ? 3u : 4u;
        uint32_t c[4] = { 0u, 0u, 0u, 0u };
        for (uint32_t k = 0; k < comps; k++) {
            const size_t o = (t == 9u ? 2 : 4) * ((size_t)comps * i + k);
            const float v = t == 9u ? half_to_float_ref((uint32_t)p[o] | ((uint32_t)p[o + 1] << 8)) : __uint_as_float(src_word(p, o));
            c[k] = f2u8(v * 255.0f);
        }
        switch (f) {
        case 0u: return pack4(c[0], 0, 0, 255);                       // :490-532
        case 1u: return pack4(0, c[0], 0, 255);
        case 2u: return pack4(0, 0, c[0], 255);
        case 3u: return pack4(255, 255, 255, c[0]);
        case 4u: return pack4(c[0], c[0], c[0], 255);                 // :417-431
        case 5u: return pack4(c[0], c[0], c[0], c[1]);                // :441-457
        case 6u: return pack4(c[0], c[1], c[2], 255);                 // :589-639
        case 8u: return pack4(c[2], c[1], c[0], 255);
        case 7u: return pack4(c[0], c[1], c[2], c[3]);                // :709-759
        case 9u: return pack4(c[2], c[1], c[0], c[3]);
        default: return 0u;
        }
    }
    const uint32_t x = (uint32_t)p[2 * i] | ((uint32_t)p[2 * i + 1] << 8);
    if (t == 2u) {                                                    // 5_6_5, :549-573
        const uint32_t hi = f2u8((float)((x & 0xF800u) >> 11) * (255.0f / 31));
        const uint32_t g = f2u8((float)((x & 0x07E0u) >> 5) * (255.0f / 63));
        const uint32_t lo = f2u8((float)(x & 0x001Fu) * (255.0f / 31));
        return f == 6u ? pack4(hi, g, lo, 255) : pack4(lo, g, hi, 255);
    }
    if (t == 3u) {                                                    // 5_5_5_1, :644-668
        const uint32_t c11 = f2u8((float)((x & 0xF800u) >> 11) * (255.0f / 31));
        const uint32_t c6 = f2u8((float)((x & 0x07C0u) >> 6) * (255.0f / 31));
        const uint32_t c1 = f2u8((float)((x & 0x003Eu) >> 1) * (255.0f / 31));
        const uint32_t a = (x & 1u) * 255u;
        return f == 7u ? pack4(c11, c6, c1, a) : pack4(c1, c6, c11, a);
    }
    if (t == 4u) {                                                    // 4_4_4_4, :670-694
        const uint32_t c12 = f2u8((float)((x & 0xF000u) >> 12) * (255.0f / 15));
        const uint32_t c8 = f2u8((float)((x & 0x0F00u) >> 8) * (255.0f / 15));
        const uint32_t c4 = f2u8((float)((x & 0x00F0u) >> 4) * (255.0f / 15));
        const uint32_t a = f2u8((float)(x & 0x000Fu) * (255.0f / 15));
        return f == 7u ? pack4(c12, c8, c4, a) : pack4(c4, c8, c12, a);
    }
    return 0u;
}

// ---------------------------------------------------------------------------
// Wide render targets (PFF_TARGET_LA_HALF and after: luminance-alpha, RGB,
// BGR, RGBA and BGRA in half or float components).  A kernel keeps such a
// pixel as an RGBA8 value and whether the frame has written it: an
// unwritten pixel keeps its bytes and its value is what the target's getter
// reads from them; a written pixel will hold the setter's encoding of its
// value, which blending reads back through the getter (the round trips
// below).  The raster uses the SIMD accessors, the other kernels the scalar
// ones, as the reference does.
// ---------------------------------------------------------------------------

// pfmFloatToHalfI (pfm.h:107-126): mantissa rounded half up, magnitudes
// below 2^-14 flushed to signed zero, overflow to infinity, NaN to 0x7E00.
__device__ __forceinline__ uint32_t float_to_half_ref(float x) {
    const uint32_t ui = __float_as_uint(x);
    const int32_t s = (int32_t)((ui >> 16) & 0x8000u);
    const int32_t em = (int32_t)(ui & 0x7FFFFFFFu);
    int32_t h = (em - (112 << 23) + (1 << 12)) >> 13;
    if (em < (113 << 23)) h = 0;
    if (em >= (143 << 23)) h = 0x7C00;
    if (em > (255 << 23)) h = 0x7E00;
    return (uint32_t)(s | h);
}

__device__ __forceinline__ bool wide_half(uint32_t kind) {
    return kind == PFF_TARGET_LA_HALF || kind == PFF_TARGET_RGB_HALF || kind == PFF_TARGET_BGR_HALF
        || kind == PFF_TARGET_RGBA_HALF || kind == PFF_TARGET_BGRA_HALF;
}
__device__ __forceinline__ uint32_t wide_comps(uint32_t kind) {
    return kind <= PFF_TARGET_LA_FLOAT ? 2u : kind <= PFF_TARGET_BGR_FLOAT ? 3u : 4u;
}
__device__ __forceinline__ uint32_t wide_source(uint32_t kind) {
    switch (kind) {
    case PFF_TARGET_LA_HALF:    return PFF_SOURCE(5, 9);
    case PFF_TARGET_LA_FLOAT:   return PFF_SOURCE(5, 10);
    case PFF_TARGET_RGB_HALF:   return PFF_SOURCE(6, 9);
    case PFF_TARGET_BGR_HALF:   return PFF_SOURCE(8, 9);
    case PFF_TARGET_RGB_FLOAT:  return PFF_SOURCE(6, 10);
    case PFF_TARGET_BGR_FLOAT:  return PFF_SOURCE(8, 10);
    case PFF_TARGET_RGBA_HALF:  return PFF_SOURCE(7, 9);
    case PFF_TARGET_BGRA_HALF:  return PFF_SOURCE(9, 9);
    case PFF_TARGET_RGBA_FLOAT: return PFF_SOURCE(7, 10);
    default:                    return PFF_SOURCE(9, 10);
    }
}

// Stores component k of pixel i: a half or a float.
__device__ __forceinline__ void wide_put(uint8_t* p, uint32_t kind, size_t i, uint32_t k, float v) {
    const size_t n = wide_comps(kind);
    if (wide_half(kind)) reinterpret_cast<uint16_t*>(p)[n * i + k] = (uint16_t)float_to_half_ref(v);
    else reinterpret_cast<float*>(p)[n * i + k] = v;
}

// The components the SIMD setter stores for RGBA8 colour c, in memory
// order: channels times 1.0f/255.0f (simd.h:88); LA takes the integer grey
// (b*29 + g*150 + r*77) >> 8 (color.h:205-240); BGR stores red first
// (pixel.h:1731-1848), BGRA blue first (pixel.h:2123-2250).
__device__ __forceinline__ void wide_simd_comps(uint32_t kind, uint32_t c, float* v) {
    const float inv = 1.0f / 255.0f;
    const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u, a = c >> 24;
    if (kind <= PFF_TARGET_LA_FLOAT) {                                        // pixel.h:1013-1090
        const uint32_t grey = ((b * 29u + g * 150u) + r * 77u) >> 8;
        v[0] = (float)(grey & 255u) * inv;
        v[1] = (float)a * inv;
        return;
    }
    const bool bgra = kind == PFF_TARGET_BGRA_HALF || kind == PFF_TARGET_BGRA_FLOAT;
    v[0] = (float)(bgra ? b : r) * inv;
    v[1] = (float)g * inv;
    v[2] = (float)(bgra ? r : b) * inv;
    v[3] = (float)a * inv;
}

__device__ __forceinline__ void wide_set_simd(uint8_t* p, uint32_t kind, size_t i, uint32_t c) {
    float v[4];
    wide_simd_comps(kind, c, v);
    const uint32_t n = wide_comps(kind);
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
        if (k < n) wide_put(p, kind, i, k, v[k]);
}

// The SIMD getter (pixel.h:2328-2383, 2666-2790, 2923-3040).
__device__ __forceinline__ uint32_t wide_get_simd(const uint8_t* p, uint32_t kind, size_t i) {
    return decode_source_texel(p, i, wide_source(kind));
}

// The SIMD getter's reading of the SIMD setter's encoding of c: each
// stored component v read back as cvt_rne(v * 255), packed as the getter
// packs it (BGR's first component and BGRA's third land in red).
__device__ __forceinline__ uint32_t wide_rt_simd(uint32_t kind, uint32_t c) {
    float v[4];
    wide_simd_comps(kind, c, v);
    const bool h = wide_half(kind);
    uint32_t q[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float s = h ? half_to_float_ref(float_to_half_ref(v[k])) : v[k];
        q[k] = (uint32_t)cvt_rne_i32(s * 255.0f);
    }
    if (kind <= PFF_TARGET_LA_FLOAT) return ((q[1] << 24) | (q[0] << 16)) | ((q[0] << 8) | q[0]);
    if (kind <= PFF_TARGET_BGR_FLOAT) return (0xFF000000u | (q[2] << 16)) | ((q[1] << 8) | q[0]);
    const bool bgra = kind == PFF_TARGET_BGRA_HALF || kind == PFF_TARGET_BGRA_FLOAT;
    return ((q[3] << 24) | ((bgra ? q[0] : q[2]) << 16)) | ((q[1] << 8) | (bgra ? q[2] : q[0]));
}

// The components the scalar setter stores (pixel.h:102-127, 220-290,
// 347-400): channels times (float)PF_INV_255; LA the grey of
// PF_COLOR_GARYSCALE in float (pixel.h:69-74); BGR swaps red and blue, the
// BGRA normaliser does not (pixel.h:59-67).
__device__ __forceinline__ void wide_scalar_comps(uint32_t kind, uint32_t c, float* v) {
    const float inv = (float)(1.0 / 255.0);
    const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u, a = c >> 24;
    if (kind <= PFF_TARGET_LA_FLOAT) {
        v[0] = (float)r * inv * 0.299f + (float)g * inv * 0.587f + (float)b * inv * 0.114f;
        v[1] = (float)a * inv;
        return;
    }
    const bool bgr = kind == PFF_TARGET_BGR_HALF || kind == PFF_TARGET_BGR_FLOAT;
    v[0] = (float)(bgr ? b : r) * inv;
    v[1] = (float)g * inv;
    v[2] = (float)(bgr ? r : b) * inv;
    v[3] = (float)a * inv;
}

__device__ __forceinline__ void wide_set_scalar(uint8_t* p, uint32_t kind, size_t i, uint32_t c) {
    float v[4];
    wide_scalar_comps(kind, c, v);
    const uint32_t n = wide_comps(kind);
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
        if (k < n) wide_put(p, kind, i, k, v[k]);
}

// The scalar getter (pixel.h:441-457, 589-639, 709-759).
__device__ __forceinline__ uint32_t wide_get_scalar(const uint8_t* p, uint32_t kind, size_t i) {
    return decode_scalar_texel(p, i, wide_source(kind));
}

// The scalar getter's reading of the scalar setter's encoding of c: each
// component as (PFubyte)(v * 255), packed as the getter packs it.
__device__ __forceinline__ uint32_t wide_rt_scalar(uint32_t kind, uint32_t c) {
    float v[4];
    wide_scalar_comps(kind, c, v);
    const bool h = wide_half(kind);
    uint32_t q[4];
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = f2u8((h ? half_to_float_ref(float_to_half_ref(v[k])) : v[k]) * 255.0f);
    if (kind <= PFF_TARGET_LA_FLOAT) return pack4(q[0], q[0], q[0], q[1]);
    const bool bgr = kind == PFF_TARGET_BGR_HALF || kind == PFF_TARGET_BGR_FLOAT;
    const bool bgra = kind == PFF_TARGET_BGRA_HALF || kind == PFF_TARGET_BGRA_FLOAT;
    if (kind <= PFF_TARGET_BGR_FLOAT) return bgr ? pack4(q[2], q[1], q[0], 255) : pack4(q[0], q[1], q[2], 255);
    return bgra ? pack4(q[2], q[1], q[0], q[3]) : pack4(q[0], q[1], q[2], q[3]);
}

__global__ __launch_bounds__(256) void k_decode_image(const uint8_t* __restrict__ src, size_t n, uint32_t source,
                                                      uint32_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = decode_scalar_texel(src, i, source);
}

// The scalar getter of a target (pixel.h:549-707) on stored pixel v:
// 16-bit fields scaled by 255.0f/31, /63 or /15 in float and truncated.
__device__ __forceinline__ uint32_t tgt_get_scalar(uint32_t kind, uint32_t v) {
    switch (kind) {
    case PFF_TARGET_RGB8: return v | 0xFF000000u;
    case PFF_TARGET_BGR8: return ((v >> 16) & 255u) | (v & 0xFF00u) | ((v & 255u) << 16) | 0xFF000000u;
    case PFF_TARGET_RGB565: case PFF_TARGET_BGR565: {
        const uint32_t hi = f2u8((float)((v & 0xF800u) >> 11) * (255.0f / 31));
        const uint32_t g = f2u8((float)((v & 0x07E0u) >> 5) * (255.0f / 63));
        const uint32_t lo = f2u8((float)(v & 0x001Fu) * (255.0f / 31));
        return kind == PFF_TARGET_RGB565 ? pack4(hi, g, lo, 255) : pack4(lo, g, hi, 255);
    }
    case PFF_TARGET_RGBA5551: case PFF_TARGET_BGRA5551: {
        const uint32_t c11 = f2u8((float)((v & 0xF800u) >> 11) * (255.0f / 31));
        const uint32_t c6 = f2u8((float)((v & 0x07C0u) >> 6) * (255.0f / 31));
        const uint32_t c1 = f2u8((float)((v & 0x003Eu) >> 1) * (255.0f / 31));
        const uint32_t a = (v & 1u) * 255u;
        return kind == PFF_TARGET_RGBA5551 ? pack4(c11, c6, c1, a) : pack4(c1, c6, c11, a);
    }
    case PFF_TARGET_RGBA4444: case PFF_TARGET_BGRA4444: {
        const uint32_t c12 = f2u8((float)((v & 0xF000u) >> 12) * (255.0f / 15));
        const uint32_t c8 = f2u8((float)((v & 0x0F00u) >> 8) * (255.0f / 15));
        const uint32_t c4 = f2u8((float)((v & 0x00F0u) >> 4) * (255.0f / 15));
        const uint32_t a = f2u8((float)(v & 0x000Fu) * (255.0f / 15));
        return kind == PFF_TARGET_RGBA4444 ? pack4(c12, c8, c4, a) : pack4(c4, c8, c12, a);
    }
    default: return v;
    }
}

// The scalar setter of a target (pixel.h:206-400) for RGBA8 colour c.
// Channels are normalised by (float)(1/255) and rounded (roundf) to the
// field; the 5_5_5_1 alpha bit is n_a > 50/255 (config.h:45-47, compared
// in double).  The BGRA 5_5_5_1 and 4_4_4_4 setters normalise without the
// red/blue swap (pixel.h:59-67), so they store RGBA order.
__device__ __forceinline__ uint32_t tgt_set_scalar(uint32_t kind, uint32_t c) {
    const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u, a = c >> 24;
    const float inv = (float)(1.0 / 255.0);
    switch (kind) {
    case PFF_TARGET_RGB8: return r | (g << 8) | (b << 16);
    case PFF_TARGET_BGR8: return b | (g << 8) | (r << 16);
    case PFF_TARGET_RGB565: case PFF_TARGET_BGR565: {
        const uint32_t c0 = kind == PFF_TARGET_RGB565 ? r : b, c2 = kind == PFF_TARGET_RGB565 ? b : r;
        return (((uint32_t)(uint8_t)roundf((float)c0 * inv * 31.0f)) << 11)
             | (((uint32_t)(uint8_t)roundf((float)g * inv * 63.0f)) << 5)
             | (uint32_t)(uint8_t)roundf((float)c2 * inv * 31.0f);
    }
    case PFF_TARGET_RGBA5551: case PFF_TARGET_BGRA5551: {
        const float na = (float)a * inv;
        return (((uint32_t)(uint8_t)roundf((float)r * inv * 31.0f)) << 11)
             | (((uint32_t)(uint8_t)roundf((float)g * inv * 31.0f)) << 6)
             | (((uint32_t)(uint8_t)roundf((float)b * inv * 31.0f)) << 1)
             | ((double)na > 50.0 * (1.0 / 255.0) ? 1u : 0u);
    }
    case PFF_TARGET_RGBA4444: case PFF_TARGET_BGRA4444:
        return (((uint32_t)(uint8_t)roundf((float)r * inv * 15.0f)) << 12)
             | (((uint32_t)(uint8_t)roundf((float)g * inv * 15.0f)) << 8)
             | (((uint32_t)(uint8_t)roundf((float)b * inv * 15.0f)) << 4)
             | (uint32_t)(uint8_t)roundf((float)a * inv * 15.0f);
    default: return c;
    }
}

// One pixel of a target through the scalar getter / setter (points,
// lines, pfDrawPixels, rects, fog).
__device__ __forceinline__ uint32_t px_get_scalar(const uint8_t* p, uint32_t kind, size_t i) {
    return kind >= PFF_TARGET_FIRST_WIDE ? wide_get_scalar(p, kind, i) : tgt_get_scalar(kind, tgt_load(p, kind, i));
}
__device__ __forceinline__ void px_set_scalar(uint8_t* p, uint32_t kind, size_t i, uint32_t c) {
    if (kind >= PFF_TARGET_FIRST_WIDE) wide_set_scalar(p, kind, i, c);
    else tgt_store(p, kind, i, tgt_set_scalar(kind, c));
}

struct DPixOp {
    uint32_t  kind;               // 0 pfDrawPixels, 1 pfFogProcess
    const uint32_t* texels;       // w*h RGBA8
    uint8_t*  color;              // render target
    float*    depth;
    uint32_t  target;             // PFF_TARGET_* of color
    uint32_t  W, H;
    uint32_t  w, h;
    int32_t   xScreen, yScreen, xMin, yMin, xMax, yMax;
    float     zPos, invXLen, invYLen;
    uint32_t  flags, depthFunc, blendMode;
    uint32_t  fogMode, fogColor;
    float     fogDensity, fogStart, fogEnd;
};

// (PFsizei)(float) of gcc on x86-64: cvttss2si into a 64-bit register, low
// 32 bits kept; NaN and values outside int64 give 0.
__device__ __forceinline__ uint32_t f2u32_x86(float f) {
    if (!(f >= -9223372036854775808.0f && f < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}

__device__ __forceinline__ bool depth_scalar(uint32_t func, float s, float d) {
    switch (func) {
    case 0u: return s == d;
    case 1u: return s != d;
    case 2u: return s < d;
    case 3u: return s <= d;
    case 4u: return s > d;
    default: return s >= d;
    }
}

__device__ __forceinline__ uint32_t blend_scalar(uint32_t mode, uint32_t s, uint32_t d) {
    uint32_t o = 0u;
    const int32_t al = (int32_t)(s >> 24) + 1, inv = 256 - al;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int32_t a = (int32_t)((s >> (8 * i)) & 255u), b = (int32_t)((d >> (8 * i)) & 255u);
        int32_t v;
        switch (mode) {
        case 0u: v = (a + b) >> 1; break;
        case 1u: v = i < 3 ? (al * a + inv * b) >> 8 : (al * 255 + inv * b) >> 8; break;
        case 2u: v = min(a + b, 255); break;
        case 3u: v = max(b - a, 0); break;
        case 4u: v = (a * b) / 255; break;
        case 5u: v = min(((b * (255 - a)) >> 8) + a, 255); break;
        case 6u: v = max(a, b); break;
        default: v = min(a, b); break;
        }
        o |= ((uint32_t)v & 255u) << (8 * i);
    }
    return o;
}

// pfFogProcess (context.c:2275-2343) on one pixel: the fog colour blended
// (scalar ALPHA) from the pixel's depth; expf / exp2f as the float rounding
// of the double results.
__device__ __forceinline__ uint32_t fog_pixel(const DPixOp& o, float depth, uint32_t dst) {
    const uint32_t fc = o.fogColor, alpha = fc >> 24;
    if (depth >= o.fogEnd) return alpha == 255u ? fc : blend_scalar(1u, fc, dst);
    float t = 0.0f;
    const float d = -o.fogDensity * (depth - o.fogStart);
    if (o.fogMode == 0u) t = (depth - o.fogStart) * __fdiv_rn(1.0f, o.fogEnd - o.fogStart);
    else if (o.fogMode == 1u) t = 1.0f - (float)exp((double)d);
    else if (o.fogMode == 2u) t = 1.0f - (float)exp2((double)d);
    const uint32_t a = f2u8(t * (float)alpha);
    return blend_scalar(1u, (fc & 0x00FFFFFFu) | (a << 24), dst);
}

// grid.y = the pass's calls; each thread strides over its call's rectangle
// (the whole target for fog).
__global__ __launch_bounds__(256) void k_draw_pixels(const DPixOp* __restrict__ ops) {
    const DPixOp o = ops[blockIdx.y];
    if (o.kind == 1u) {
        const size_t n = (size_t)o.W * o.H;
        for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
            const float z = o.depth[k];
            if (!(z >= o.fogEnd) && !(z > o.fogStart)) continue;      // untouched: no getter or setter call
            px_set_scalar(o.color, o.target, k, fog_pixel(o, z, px_get_scalar(o.color, o.target, k)));
        }
        return;
    }
    if (o.xMax < o.xMin || o.yMax < o.yMin) return;
    const uint32_t rw = (uint32_t)(o.xMax - o.xMin) + 1u, rh = (uint32_t)(o.yMax - o.yMin) + 1u;
    const uint64_t area = (uint64_t)rw * rh;
    const uint32_t n = o.w * o.h;
    for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < area; k += (uint64_t)gridDim.x * 256) {
        const int32_t x = o.xMin + (int32_t)(k % rw), y = o.yMin + (int32_t)(k / rw);
        if (x < 0 || y < 0 || (uint32_t)x >= o.W || (uint32_t)y >= o.H) continue;
        const size_t dst = (size_t)y * o.W + (size_t)x;
        if ((o.flags & PFF_DEPTH_TEST) && !depth_scalar(o.depthFunc, o.zPos, o.depth[dst])) continue;
        const float v = (float)(y - o.yScreen) * o.invYLen;
        const float u = (float)(x - o.xScreen) * o.invXLen;
        uint32_t src = f2u32_x86(v * (float)(o.h - 1u)) * o.w + f2u32_x86(u * (float)(o.w - 1u));
        if (src >= n) src = n - 1u;
        if (!(o.flags & PFF_PIXELS_RECT)) o.depth[dst] = o.zPos;      // pfRect writes no depth
        uint32_t c = o.texels[src];
        if (o.flags & PFF_BLEND) c = blend_scalar(o.blendMode, c, px_get_scalar(o.color, o.target, dst));
        px_set_scalar(o.color, o.target, dst, c);
    }
}

// ---------------------------------------------------------------------------
// Points (points.c:35-183).  k_point_setup projects every point of the job
// once: MVP, the clip test and the division of x and y when w != 1, the
// screen mapping (ictx.c:51-65) and the viewport test.  After the raster
// kernel of the pass a run of points occupies, k_points gives each 16x16
// tile of the target to a workgroup whose threads own one pixel each and
// walk the run's points in order: one pixel at size <= 1, else the disc
// x^2 + y^2 <= (size/2)^2; scalar depth test and blending; the depth is
// written also without the depth test.
// ---------------------------------------------------------------------------

// Ordered compaction of one chunk of up to 256 primitives: `hit` per thread
// (primitive base + threadIdx.x); the indices of the hits land in idx[] in
// primitive order and the count is returned.  Every thread of the
// workgroup must call it (it synchronises the workgroup).
__device__ __forceinline__ uint32_t compact_hits(bool hit, uint32_t i, uint32_t* idx, uint32_t* waveCount) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(hit);
    if (lane == 0) waveCount[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += waveCount[w];
    const uint32_t total = waveCount[0] + waveCount[1] + waveCount[2] + waveCount[3];
    if (hit) idx[before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    __syncthreads();
    return total;
}

struct DPointScr {
    int32_t  cx, cy;              // screen pixel
    float    z;                   // clip z (not divided)
    uint32_t color;
    int32_t  ir;                  // (int)(size / 2), 0 at size <= 1
    float    rSq;                 // (size / 2)^2
    uint32_t flags;               // PFF_DEPTH_TEST | PFF_BLEND; 0x80000000: rejected
    uint16_t depthFunc, blendMode;
};

__global__ __launch_bounds__(256) void k_point_setup(const PFFpoint* __restrict__ pts, uint32_t n,
                                                     const PFFstate* __restrict__ states,
                                                     DPointScr* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const PFFpoint pt = pts[i];
    const PFFstate& st = states[pt.state];
    DPointScr o;
    o.flags = 0x80000000u;
    o.cx = o.cy = 0; o.z = 0.0f; o.color = pt.color; o.ir = 0; o.rSq = 0.0f; o.depthFunc = 0; o.blendMode = 0;
    float h[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        h[k] = st.mvp[k] * pt.position[0] + st.mvp[4 + k] * pt.position[1] + st.mvp[8 + k] * pt.position[2]
             + st.mvp[12 + k] * pt.position[3];
    bool ok = true;
    if (h[3] != 1.0f) {
        for (int k = 0; k < 3; k++) ok = ok && !(h[k] < -h[3] || h[k] > h[3]);
        const float invW = __fdiv_rn(1.0f, h[3]);
        h[0] *= invW;
        h[1] *= invW;
    }
    const float sx = ((float)st.vpPos[0] + ((h[0] + 1.0f) * 0.5f) * (float)st.vpDim[0]) + 0.5f;
    const float sy = ((float)st.vpPos[1] + ((1.0f - h[1]) * 0.5f) * (float)st.vpDim[1]) + 0.5f;
    ok = ok && sx >= (float)st.vpMin[0] && sy >= (float)st.vpMin[1] && sx <= (float)st.vpMax[0] && sy <= (float)st.vpMax[1];
    if (ok) {
        o.cx = cvtt_i32(sx);
        o.cy = cvtt_i32(sy);
        o.z = h[2];
        if (pt.size > 1.0f) {
            const float r = pt.size * 0.5f;
            o.rSq = r * r;
            o.ir = cvtt_i32(r);
        }
        o.flags = st.flags & (PFF_DEPTH_TEST | PFF_BLEND);
        o.depthFunc = (uint16_t)st.depthFunc;
        o.blendMode = (uint16_t)st.blendMode;
    }
    out[i] = o;
}

struct DPtOp {
    uint8_t*  color;
    float*    depth;
    uint32_t  target;             // PFF_TARGET_* of color
    uint32_t  W, H, tilesX, tilesY;
    uint32_t  first, last;        // job point indices
};

__global__ __launch_bounds__(256) void k_points(const DPtOp* __restrict__ ops, const DPointScr* __restrict__ scr) {
    const DPtOp o = ops[blockIdx.y];
    if (blockIdx.x >= o.tilesX * o.tilesY) return;
    const int32_t x0 = (int32_t)(blockIdx.x % o.tilesX) * 16, y0 = (int32_t)(blockIdx.x / o.tilesX) * 16;
    const int32_t px = x0 + (int32_t)(threadIdx.x & 15u), py = y0 + (int32_t)(threadIdx.x >> 4);
    const bool inside = (uint32_t)px < o.W && (uint32_t)py < o.H;
    const size_t off = inside ? (size_t)py * o.W + (size_t)px : 0;
    // Narrow targets: c is the stored word.  Wide targets: c is the RGBA8
    // value (the scalar getter's reading while unwritten, then the colour
    // last written; dirty says written).
    const bool wide = o.target >= PFF_TARGET_FIRST_WIDE;
    uint32_t c = inside ? (wide ? wide_get_scalar(o.color, o.target, off) : tgt_load(o.color, o.target, off)) : 0u;
    float d = inside ? o.depth[off] : 0.0f;
    bool dirty = false;
    __shared__ uint32_t idx[256];
    __shared__ uint32_t waveCount[4];
    for (uint32_t base = o.first; base < o.last; base += 256) {
      // The chunk's points that touch this tile, in order.
      const uint32_t mine = base + threadIdx.x;
      bool hit = false;
      if (mine < o.last) {
          const DPointScr t = scr[mine];
          hit = !(t.flags & 0x80000000u)
             && !(t.cx + t.ir < x0 || t.cx - t.ir > x0 + 15 || t.cy + t.ir < y0 || t.cy - t.ir > y0 + 15);
      }
      const uint32_t nHit = compact_hits(hit, mine, idx, waveCount);
      for (uint32_t h = 0; h < nHit; h++) {
        const DPointScr s = scr[idx[h]];
        const int32_t dx = px - s.cx, dy = py - s.cy;
        const bool cov = s.ir == 0 && s.rSq == 0.0f
                       ? (dx == 0 && dy == 0)
                       : (dx >= -s.ir && dx <= s.ir && dy >= -s.ir && dy <= s.ir && (float)(dy * dy + dx * dx) <= s.rSq);
        if (!inside || !cov) continue;
        if ((s.flags & PFF_DEPTH_TEST) && !depth_scalar(s.depthFunc, s.z, d)) continue;
        if (wide)
            c = (s.flags & PFF_BLEND) ? blend_scalar(s.blendMode, s.color, dirty ? wide_rt_scalar(o.target, c) : c) : s.color;
        else
            c = tgt_set_scalar(o.target, (s.flags & PFF_BLEND) ? blend_scalar(s.blendMode, s.color, tgt_get_scalar(o.target, c))
                                                                : s.color);
        d = s.z;
        dirty = true;
      }
      __syncthreads();                                          // idx[] is rewritten by the next chunk
    }
    if (dirty) {
        if (wide) wide_set_scalar(o.color, o.target, off, c);
        else tgt_store(o.color, o.target, off, c);
        o.depth[off] = d;
    }
}

// ---------------------------------------------------------------------------
// Lines (lines.c:56-532).  k_line_setup runs one thread per group (a
// PF_LINES pair or the edges of one polygon): projection, the 2D clipper
// (Cohen-Sutherland on screen positions, BOTTOM mapped to the top edge and
// TOP to the bottom one as in the reference, at most 64 steps) or the 3D
// one (Liang-Barsky on clip coordinates), the division of x and y, the
// screen mapping; a rejected edge ends its group.  k_lines gives each 16x16
// tile to a workgroup whose threads walk the run's lines in order.  A line
// is thin, or above width 1.5 a depth-tested centre line plus copies
// shifted by 1..w across the shorter axis.  Each thin line is the
// reference's fixed-point DDA: step k of the longer axis lands on the
// shorter-axis coordinate start + ((k * decInc) >> 16), so a pixel knows in
// O(1) which step, if any, covers it.  Pixels are addressed y * W + x as in
// the reference (an x past a side wraps to the neighbouring row), so each
// thread also tests the two wrapped positions of its offset and applies the
// steps it owns in step order.
// ---------------------------------------------------------------------------

struct DLineScr {
    float    s1[2], s2[2];        // screen end points after clipping
    float    z1, z2;
    uint32_t c1, c2;
    uint32_t flags;               // PFF_DEPTH_TEST | PFF_BLEND; 0x80000000 inactive; 0x40000000 thick
    uint16_t depthFunc, blendMode;
    int32_t  w, axis;             // thick copies: 1..w each side, along axis (0 x, 1 y)
    int32_t  bx0, by0, bx1, by1;  // pixel bounds of every copy
};

__device__ __forceinline__ void line_screen(const PFFstate& st, const float* h, float* sc) {
    sc[0] = ((float)st.vpPos[0] + ((h[0] + 1.0f) * 0.5f) * (float)st.vpDim[0]) + 0.5f;
    sc[1] = ((float)st.vpPos[1] + ((1.0f - h[1]) * 0.5f) * (float)st.vpDim[1]) + 0.5f;
}

__device__ __forceinline__ uint32_t line_code(const PFFstate& st, const float* sc) {
    uint32_t c = 0u;
    if (sc[0] < (float)st.vpMin[0]) c |= 1u;
    if (sc[0] > (float)st.vpMax[0]) c |= 2u;
    if (sc[1] < (float)st.vpMin[1]) c |= 8u;
    if (sc[1] > (float)st.vpMax[1]) c |= 4u;
    return c;
}

__device__ __forceinline__ bool clip_coord3d(float q, float p, float& t1, float& t2) {
    if (fabsf(p) < 1e-5f) return !(q < -1e-5f);
    const float r = __fdiv_rn(q, p);
    if (p < 0.0f) {
        if (r > t2) return false;
        if (r > t1) t1 = r;
    } else {
        if (r < t1) return false;
        if (r < t2) t2 = r;
    }
    return true;
}

// Process_ProjectAndClipLine (lines.c:140-258) of one line.
__device__ bool project_line(const PFFstate& st, const PFFline& l, float* h1, float* h2, float* s1, float* s2,
                             uint32_t& c1, uint32_t& c2) {
    float h[2][4];
    for (int v = 0; v < 2; v++)
        for (int k = 0; k < 4; k++)
            h[v][k] = st.mvp[k] * l.position[v][0] + st.mvp[4 + k] * l.position[v][1] + st.mvp[8 + k] * l.position[v][2]
                    + st.mvp[12 + k] * l.position[v][3];
    c1 = l.color[0];
    c2 = l.color[1];
    if (h[0][3] == 1.0f && h[1][3] == 1.0f) {
        float a[2], b[2];
        line_screen(st, h[0], a);
        line_screen(st, h[1], b);
        float m = 0.0f;
        if (a[0] != b[0]) m = __fdiv_rn(b[1] - a[1], b[0] - a[0]);
        bool accept = false;
        for (int step = 0; step < 64; step++) {
            uint32_t k0 = line_code(st, a), k1 = line_code(st, b);
            if ((k0 | k1) == 0u) { accept = true; break; }
            if (k0 & k1) break;
            if (k0 == 0u) {
                for (int q = 0; q < 2; q++) { float t = a[q]; a[q] = b[q]; b[q] = t; }
                for (int q = 0; q < 4; q++) { float t = h[0][q]; h[0][q] = h[1][q]; h[1][q] = t; }
                uint32_t t = c1; c1 = c2; c2 = t;
                k0 = k1;
            }
            if (k0 & 1u) { a[1] += ((float)st.vpMin[0] - a[0]) * m; a[0] = (float)st.vpMin[0]; }
            else if (k0 & 2u) { a[1] += ((float)st.vpMax[0] - a[0]) * m; a[0] = (float)st.vpMax[0]; }
            else if (k0 & 4u) { if (m != 0.0f) a[0] += __fdiv_rn((float)st.vpMin[1] - a[1], m); a[1] = (float)st.vpMin[1]; }
            else if (k0 & 8u) { if (m != 0.0f) a[0] += __fdiv_rn((float)st.vpMax[1] - a[1], m); a[1] = (float)st.vpMax[1]; }
        }
        if (!accept) return false;
        for (int q = 0; q < 4; q++) { h1[q] = h[0][q]; h2[q] = h[1][q]; }
        s1[0] = a[0]; s1[1] = a[1]; s2[0] = b[0]; s2[1] = b[1];
        return true;
    }
    float t1 = 0.0f, t2 = 1.0f, d[4];
    for (int q = 0; q < 4; q++) d[q] = h[1][q] - h[0][q];
    const float* g = h[0];
    if (!clip_coord3d(g[3] - g[0], -d[3] + d[0], t1, t2)) return false;
    if (!clip_coord3d(g[3] + g[0], -d[3] - d[0], t1, t2)) return false;
    if (!clip_coord3d(g[3] - g[1], -d[3] + d[1], t1, t2)) return false;
    if (!clip_coord3d(g[3] + g[1], -d[3] - d[1], t1, t2)) return false;
    if (!clip_coord3d(g[3] - g[2], -d[3] + d[2], t1, t2)) return false;
    if (!clip_coord3d(g[3] + g[2], -d[3] - d[2], t1, t2)) return false;
    if (t2 < 1.0f) for (int q = 0; q < 4; q++) h[1][q] = h[0][q] + d[q] * t2;
    if (t1 > 0.0f) for (int q = 0; q < 4; q++) h[0][q] = h[0][q] + d[q] * t1;
    for (int v = 0; v < 2; v++) {
        const float invW = __fdiv_rn(1.0f, h[v][3]);
        h[v][0] *= invW;
        h[v][1] *= invW;
    }
    line_screen(st, h[0], s1);
    line_screen(st, h[1], s2);
    for (int q = 0; q < 4; q++) { h1[q] = h[0][q]; h2[q] = h[1][q]; }
    return true;
}

__global__ __launch_bounds__(256) void k_line_setup(const PFFline* __restrict__ lines, uint32_t n,
                                                    const PFFstate* __restrict__ states, DLineScr* __restrict__ out) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= n || !lines[g].groupStart) return;
    bool dead = false;
    for (uint32_t i = g; i < n && (i == g || !lines[i].groupStart); i++) {
        const PFFline l = lines[i];
        const PFFstate& st = states[l.state];
        DLineScr o;
        memset(&o, 0, sizeof o);
        o.flags = 0x80000000u;
        float h1[4], h2[4];
        if (!dead && project_line(st, l, h1, h2, o.s1, o.s2, o.c1, o.c2)) {
            o.z1 = h1[2];
            o.z2 = h2[2];
            o.flags = st.flags & (PFF_DEPTH_TEST | PFF_BLEND);
            o.depthFunc = (uint16_t)st.depthFunc;
            o.blendMode = (uint16_t)st.blendMode;
            const int32_t x1 = cvtt_i32(o.s1[0]), y1 = cvtt_i32(o.s1[1]), x2 = cvtt_i32(o.s2[0]), y2 = cvtt_i32(o.s2[1]);
            o.w = 0;
            o.axis = 0;
            if (l.width > 1.5f) {
                o.flags |= 0x40000000u;
                const int32_t dx = x2 - x1, dy = y2 - y1;
                if (dx != 0 && abs(dy / dx) < 1) {
                    o.w = cvtt_i32((l.width - 1.0f) * __fdiv_rn(1.0f, __fsqrt_rn((float)(dx * dx + dy * dy))) * (float)abs(dx)) >> 1;
                    o.axis = 1;
                } else if (dy != 0) {
                    o.w = cvtt_i32((l.width - 1.0f) * __fdiv_rn(1.0f, __fsqrt_rn((float)(dx * dx + dy * dy))) * (float)abs(dy)) >> 1;
                    o.axis = 0;
                }
                if (o.w < 0) o.w = 0;
                if (o.w > 16384) o.w = 16384;                   // copies past any target draw nothing
            }
            const int32_t ex = o.axis == 0 ? o.w + 1 : 1, ey = o.axis == 1 ? o.w + 1 : 1;
            o.bx0 = min(x1, x2) - ex; o.bx1 = max(x1, x2) + ex;
            o.by0 = min(y1, y2) - ey; o.by1 = max(y1, y2) + ey;
        } else {
            dead = true;                                        // lines.c:93
        }
        out[i] = o;
    }
}

struct DLnOp {
    uint8_t*  color;
    float*    depth;
    uint32_t  target;             // PFF_TARGET_* of color
    uint32_t  W, H, tilesX, tilesY;
    uint32_t  first, last;        // job line indices
};

// Step of a thin line (ints X1..Y2) covering line pixel (x, y), or -1.
__device__ __forceinline__ int32_t line_step(int32_t X1, int32_t Y1, int32_t X2, int32_t Y2, int32_t x, int32_t y,
                                             int32_t& iOut, float& invEnd) {
    int32_t shortLen = Y2 - Y1, longLen = X2 - X1;
    const bool yl = abs(shortLen) > abs(longLen);
    if (yl) { const int32_t t = shortLen; shortLen = longLen; longLen = t; }
    invEnd = __fdiv_rn(1.0f, (float)longLen);
    const int32_t endVal = longLen;
    int32_t sgn = 1;
    if (longLen < 0) { longLen = -longLen; sgn = -1; }
    const int32_t decInc = longLen == 0 ? 0 : (int32_t)((uint32_t)shortLen << 16) / longLen;
    const int32_t i = yl ? y - Y1 : x - X1;
    const int32_t k = i * sgn;
    if (k < 0 || k >= (endVal < 0 ? -endVal : endVal)) return -1;
    const int32_t minor = (yl ? X1 : Y1) + ((k * decInc) >> 16);
    if ((yl ? x : y) != minor) return -1;
    iOut = i;
    return k;
}

__global__ __launch_bounds__(256) void k_lines(const DLnOp* __restrict__ ops, const DLineScr* __restrict__ scr) {
    const DLnOp o = ops[blockIdx.y];
    if (blockIdx.x >= o.tilesX * o.tilesY) return;
    const int32_t W = (int32_t)o.W;
    const int32_t x0 = (int32_t)(blockIdx.x % o.tilesX) * 16, y0 = (int32_t)(blockIdx.x / o.tilesX) * 16;
    const int32_t px = x0 + (int32_t)(threadIdx.x & 15u), py = y0 + (int32_t)(threadIdx.x >> 4);
    const bool inside = (uint32_t)px < o.W && (uint32_t)py < o.H;
    const size_t off = inside ? (size_t)py * o.W + (size_t)px : 0;
    // Narrow targets: c is the stored word.  Wide targets: c is the RGBA8
    // value (the scalar getter's reading while unwritten, then the colour
    // last written; dirty says written).
    const bool wide = o.target >= PFF_TARGET_FIRST_WIDE;
    uint32_t c = inside ? (wide ? wide_get_scalar(o.color, o.target, off) : tgt_load(o.color, o.target, off)) : 0u;
    float d = inside ? o.depth[off] : 0.0f;
    bool dirty = false;
    // Line positions whose offset y * W + x is this pixel's.
    const int32_t cxs[3] = { px, px + W, px - W }, cys[3] = { py, py - 1, py + 1 };
    __shared__ uint32_t idx[256];
    __shared__ uint32_t waveCount[4];
    for (uint32_t base = o.first; base < o.last; base += 256) {
      // The chunk's lines that can touch this tile (directly or by a row
      // wrap), in order.
      const uint32_t mine = base + threadIdx.x;
      bool hit = false;
      if (mine < o.last) {
          const DLineScr t = scr[mine];
          const bool wraps = t.bx0 < 0 || t.bx1 >= W;
          hit = !(t.flags & 0x80000000u) && !(t.by1 < y0 - 1 || t.by0 > y0 + 16)
             && (wraps || !(t.bx1 < x0 || t.bx0 > x0 + 15));
      }
      const uint32_t nHit = compact_hits(hit, mine, idx, waveCount);
      for (uint32_t hI = 0; hI < nHit; hI++) {
        const DLineScr s = scr[idx[hI]];
        if (!inside) continue;
        const bool thick = (s.flags & 0x40000000u) != 0;
        const int32_t copies = thick ? 1 + 2 * s.w : 1;
        for (int32_t cp = 0; cp < copies; cp++) {
            // Copy 0 is the centre line; then -1, +1, -2, +2, ... along axis.
            const int32_t sh = cp == 0 ? 0 : ((cp + 1) >> 1) * ((cp & 1) ? -1 : 1);
            float a0 = s.s1[0], a1 = s.s1[1], b0 = s.s2[0], b1 = s.s2[1];
            if (s.axis == 0) { a0 += (float)sh; b0 += (float)sh; } else { a1 += (float)sh; b1 += (float)sh; }
            const int32_t X1 = cvtt_i32(a0), Y1 = cvtt_i32(a1), X2 = cvtt_i32(b0), Y2 = cvtt_i32(b1);
            const bool depthOn = (cp == 0 && thick) || (s.flags & PFF_DEPTH_TEST);
            int32_t ks[3], is[3];
            float invEnd = 0.0f;
            for (int q = 0; q < 3; q++) ks[q] = line_step(X1, Y1, X2, Y2, cxs[q], cys[q], is[q], invEnd);
            for (int pass = 0; pass < 3; pass++) {
                int best = -1;
                for (int q = 0; q < 3; q++)
                    if (ks[q] >= 0 && (best < 0 || ks[q] < ks[best])) best = q;
                if (best < 0) break;
                const float t = (float)is[best] * invEnd;
                ks[best] = -1;
                const float z = s.z1 + t * (s.z2 - s.z1);
                if (depthOn && !depth_scalar(s.depthFunc, z, d)) continue;
                uint32_t col = 0u;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int32_t ca = (int32_t)((s.c1 >> (8 * q)) & 255u), cb = (int32_t)((s.c2 >> (8 * q)) & 255u);
                    col |= (uint32_t)f2u8((float)ca + t * (float)(cb - ca)) << (8 * q);
                }
                if (wide)
                    c = (s.flags & PFF_BLEND) ? blend_scalar(s.blendMode, col, dirty ? wide_rt_scalar(o.target, c) : c) : col;
                else
                    c = tgt_set_scalar(o.target, (s.flags & PFF_BLEND) ? blend_scalar(s.blendMode, col, tgt_get_scalar(o.target, c))
                                                                        : col);
                d = z;
                dirty = true;
            }
        }
      }
      __syncthreads();                                          // idx[] is rewritten by the next chunk
    }
    if (dirty) {
        if (wide) wide_set_scalar(o.color, o.target, off, c);
        else tgt_store(o.color, o.target, off, c);
        o.depth[off] = d;
    }
}

// ---------------------------------------------------------------------------
// Array draws (pfDrawElements / pfDrawArrays recorded whole, PFFarrays):
// one thread per triangle forms it from the copied elements, in the order
// the reference's vertex loop emits triangles (context.c:1231-1575: the
// vertex buffer of the mode, ictx.c:94-244 per face, triangles.c:125-152
// for fans and strips), and writes it into the job's triangle streams.
// ---------------------------------------------------------------------------

struct DArrays {
    const uint8_t* pos;
    const uint8_t* nrm;        // null: zero normal
    const uint8_t* uv;         // null: zero texture coordinate
    const uint8_t* col;        // null: `color`, no colour-array index
    const uint8_t* idx;        // null: vertex v is element v
    uint32_t posSize, posType, nrmType, uvType, colSize, colType, idxType;
    uint32_t mode;
    uint32_t perBuf;           // triangles per full buffer and face group
    uint32_t stride;           // vertices a full buffer advances (buffer size minus the kept ones)
    uint32_t numFaces, facePass0, facePass1;
    uint32_t state;            // global state index
    uint32_t color, elementBase, numElements;
    uint32_t slot0;            // global slot of the draw's first triangle
    uint32_t cum;              // array triangles of the job before this draw
    uint32_t flags;            // PFF_ARRAYS_*
    float    uvm[6];           // texture matrix (PFF_ARRAYS_TEXMAT)
};

__device__ __forceinline__ float arr_f(const uint8_t* b, uint32_t type, size_t i) {
    switch (type) {                                // context.c:1274-1310 (float conversions)
    case PFF_SHORT:  return (float)((const int16_t*)b)[i];
    case PFF_INT:    return (float)((const int32_t*)b)[i];
    case PFF_DOUBLE: return (float)((const double*)b)[i];
    default:         return ((const float*)b)[i];
    }
}

// Colour component as the reference converts it (context.c:1356-1392); a
// float or double outside int32 converts to 0 (cvttss2si / cvttsd2si give
// INT_MIN, whose low byte is 0).
__device__ __forceinline__ uint32_t arr_u8(const uint8_t* b, uint32_t type, size_t i) {
    switch (type) {
    case PFF_UBYTE:  return b[i];
    case PFF_USHORT: return ((const uint16_t*)b)[i] >> 8;
    case PFF_UINT:   return ((const uint32_t*)b)[i] >> 24;
    case PFF_FLOAT: {
        const float f = ((const float*)b)[i] * 255.0f;
        return (f > -2147483904.0f && f < 2147483648.0f) ? ((uint32_t)(int32_t)f & 0xFFu) : 0u;
    }
    default: {
        const double d = ((const double*)b)[i] * 255.0;
        return (d > -2147483649.0 && d < 2147483648.0) ? ((uint32_t)(int32_t)d & 0xFFu) : 0u;
    }
    }
}

__global__ __launch_bounds__(256) void k_deindex(const DArrays* __restrict__ D, uint32_t numDraws, uint32_t total,
                                                 float4* __restrict__ pos, uint32_t* __restrict__ col,
                                                 float2* __restrict__ uv, float* __restrict__ nrm,
                                                 uint32_t* __restrict__ info, uint32_t* __restrict__ vid) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= total) return;
    uint32_t lo = 0, hi = numDraws;                // D[lo].cum <= g < D[hi].cum
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (D[mid].cum <= g) lo = mid; else hi = mid;
    }
    const DArrays& A = D[lo];
    const uint32_t w = g - A.cum;
    const uint32_t perFill = A.perBuf * A.numFaces;
    const uint32_t fill = w / perFill, r = w - fill * perFill;
    const uint32_t grp = r / A.perBuf, i = r - grp * A.perBuf;
    uint32_t sl[3];
    if (A.mode == 2u) {                            // PF_TRIANGLES
        sl[0] = 0; sl[1] = 1; sl[2] = 2;
    } else if (A.mode == 4u || A.mode == 7u) {     // strips: odd triangles reversed
        if (i & 1u) { sl[0] = i + 2; sl[1] = i + 1; sl[2] = i; }
        else        { sl[0] = i; sl[1] = i + 1; sl[2] = i + 2; }
    } else {                                       // fans and quads: around slot 0
        sl[0] = 0; sl[1] = i + 1; sl[2] = i + 2;
    }
    const size_t t = (size_t)A.slot0 + w;
    for (int k = 0; k < 3; k++) {
        const uint32_t v = fill * A.stride + sl[k];
        uint32_t j = v;
        if (A.idx) j = A.idxType == PFF_UBYTE ? A.idx[v] : A.idxType == PFF_USHORT ? ((const uint16_t*)A.idx)[v]
                                                                                  : ((const uint32_t*)A.idx)[v];
        if (j >= A.numElements) j = A.numElements - 1;        // not produced by the host library
        float p[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        for (uint32_t c = 0; c < A.posSize; c++) p[c] = arr_f(A.pos, A.posType, (size_t)j * A.posSize + c);
        pos[3 * t + k] = make_float4(p[0], p[1], p[2], p[3]);
        uint32_t cw = A.color, vi = PFF_NO_VERTEX;
        if (A.col) {
            uint32_t ch[4] = {255u, 255u, 255u, 255u};
            for (uint32_t c = 0; c < A.colSize; c++) ch[c] = arr_u8(A.col, A.colType, (size_t)j * A.colSize + c);
            cw = ch[0] | (ch[1] << 8) | (ch[2] << 16) | (ch[3] << 24);
            if (!(A.flags & PFF_ARRAYS_LIST)) vi = j + A.elementBase;
        }
        col[3 * t + k] = cw;
        if (vid) vid[3 * t + k] = vi;
        if (uv) {
            float2 q = make_float2(0.0f, 0.0f);
            if (A.uv) { q.x = arr_f(A.uv, A.uvType, (size_t)j * 2); q.y = arr_f(A.uv, A.uvType, (size_t)j * 2 + 1); }
            if (A.flags & PFF_ARRAYS_TEXMAT) {         // pfTexCoord2f (pfm.h:954-962)
                const float u = q.x, v = q.y;
                q.x = A.uvm[0] * u + A.uvm[2] * v + A.uvm[4];
                q.y = A.uvm[1] * u + A.uvm[3] * v + A.uvm[5];
            }
            uv[3 * t + k] = q;
        }
        if (nrm) {
            float n[3];
            for (int c = 0; c < 3; c++) n[c] = A.nrm ? arr_f(A.nrm, A.nrmType, (size_t)j * 3 + c) : 0.0f;
            if (A.flags & PFF_ARRAYS_NORMALIZE) {      // pfNormal3f under PF_NORMALIZE (pfm.h:1183)
                const float sq = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
                if (sq != 0.0f) {
                    const float inv = 1.0f / sqrtf(sq);
                    for (int c = 0; c < 3; c++) n[c] = n[c] * inv;
                }
            }
            for (int c = 0; c < 3; c++) nrm[9 * t + 3 * k + c] = n[c];
        }
    }
    info[t] = A.state | (grp ? A.facePass1 : A.facePass0);
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

thread_local char g_err[512] = "";

void set_error(const char* what, hipError_t e) {
    snprintf(g_err, sizeof g_err, "%s: %s", what, e == hipSuccess ? "error" : hipGetErrorString(e));
}

#define HIP_TRY(call, code)                                          \
    do {                                                             \
        hipError_t e_ = (call);                                      \
        if (e_ != hipSuccess) { set_error(#call, e_); return code; } \
    } while (0)

int g_initState = 0;   // 0 unknown, 1 ok, -1 failed

struct DevBuf {
    void*  p = nullptr;
    size_t bytes = 0;
    bool   borrowed = false;   // carved from an arena: never freed
};

// Device memory a job carves its buffers from instead of calling hipMalloc
// (the static scratch array of pfdExecuteFrameNoAlloc).  Bump allocation;
// reset when the job is reset.
struct Arena {
    uint8_t* base = nullptr;
    size_t   cap = 0, used = 0;
    const char* owner = "";
};

struct PassInfo {
    std::vector<DSeg> segs;
    std::vector<int32_t> frameSeg;
    uint32_t numPassTris = 0;
    uint32_t numRT = 0;        // known after the first run
    uint32_t numPairs = 0;
    uint32_t numClip = 0;      // triangles the pass clips (first run)
    bool     sized = false;
    std::vector<uint32_t> blockSeg;   // per 256-triangle block: segment of its first triangle
    std::vector<DPixOp> pix;   // pfDrawPixels calls run after this pass's raster
    uint32_t pixGrid = 0;
    std::vector<DPtOp> pts;    // runs of points, after the pixel calls
    uint32_t ptTiles = 0;
    std::vector<DLnOp> lns;    // runs of lines, after the points
    uint32_t lnTiles = 0;
    // Device copies of the tables above, inside the job's passTab buffer,
    // and of the pass's per-frame tile descriptors, inside tileDescs.
    const DSeg*       dSegs = nullptr;
    const uint32_t*   dBlockSeg = nullptr;
    const DPixOp*     dPix = nullptr;
    const DPtOp*      dPts = nullptr;
    const DLnOp*      dLns = nullptr;
    const DTileFrame* dTileDesc = nullptr;
};

}  // namespace

// Device memory of pfdExecuteFrameNoAlloc: every buffer of its frames
// (inputs, render target, raster records, tile pairs, sort and scan
// buffers, counters) is carved from this array, so that entry point never
// allocates.  Each GPU holds its own copy until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (no_alloc_setup); the array's name is not a device
// address on the host.
#define PFD_SCRATCH_BYTES ((size_t)PFD_SCRATCH_BYTES_MIB << 20)
static __device__ __align__(256) uint8_t scratch[PFD_SCRATCH_BYTES];

// Every device buffer of a job is a member of this struct (no vectors of
// buffers), so a job can be reset and planned again for the next frame
// while keeping its memory: buffers only grow.
struct PFDjob {
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;           // the clip kernel beside k_setup<1> (default priority); may be null
    hipEvent_t fork = nullptr, join = nullptr;
    Arena* arena = nullptr;               // non-null: buffers come from the arena, never from hipMalloc
    bool autoTiles = false;               // choose tile size and raster mode from the first counts
    uint64_t autoKey = 0;                 // one-shot job: shape of the last automatically configured
    uint32_t autoShift = 0, autoMode = 0; // frame and its choice (kept across job_reset)
    std::vector<PFFframe> src;            // host frame descriptors: only the target fields are
                                          // used after creation (the rest may be reused by the caller)
    std::vector<uint8_t>  leadClear;      // per frame: the first command is a clear
    std::vector<PFFtexture> srcTex;       // frame 0's textures and states (differentiable jobs)
    std::vector<PFFstate>   srcStates;
    std::vector<DFrame>   frames;
    DevBuf dFrames;
    DevBuf targetPool;                    // colour + depth of every frame
    DevBuf pos, col, uv, nrm, info, states, texTable;
    DevBuf texPool;                       // texels of every texture (RGBA8 or RGBA float)
    DevBuf pixPool;                       // decoded pfDrawPixels images
    DevBuf stage;                         // raw texels / pixel images before decoding
    DevBuf passTab, tileDescs;            // per-pass tables (PassInfo::d*)
    DevBuf   points, pointScr;     // PFFpoint of all frames, their projections
    DevBuf   lines, lineScr;       // PFFline of all frames, their clipped screen forms
    uint32_t numIn = 0;
    bool hasUV = false, hasNrm = false, anyPhong = false, anyLit = false;
    bool anyWide = false;                 // a frame's target is wide (PFF_TARGET_FIRST_WIDE on)
    bool anyNarrow = false;               // a frame's target is of at most 32 bits but not RGBA8
    std::vector<PassInfo> passes;
    PFDconfig cfg{};
    uint32_t tileShift = 3;
    uint32_t numTiles = 0;
    // scratch
    DevBuf cnt, off, blockRT, blockPairs, rts, rtl, keysA, valsA, keysB, valsB;
    DevBuf scanBlocks, scanTotal, hist, histScan, tileBegin, tileEnd, tileFrags, dTileFrame, clipList, clipCount, clearSnap;
    DevBuf pairOff, pairRect;             // per raster triangle, when the pass expands its pairs
    uint32_t sortedInB = 0;               // per-pass result buffer selection (last pass)
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    // stats
    uint64_t lastFragments = 0, totalRT = 0, totalPairs = 0;
    // differentiable jobs
    bool diff = false, diffRan = false;
    DevBuf rtd, winner, vid, texParams, colParams, lightParams, image, dImage, gTex, gCol, gLightPart, gLight;
    DevBuf arrPool, arrTable;             // array draws: element copies and their DArrays
    uint32_t numVertices = 0;
    int32_t paramTex = -1;
    // Page-locked staging for copies from and to pageable host memory
    // (staged_h2d / staged_d2h), made on first use.
    void* hstage[2] = { nullptr, nullptr };
    hipEvent_t hstageEv[2] = { nullptr, nullptr };
};

namespace {

template <typename F>
void for_each_buf(PFDjob* J, F f) {
    DevBuf* bufs[] = { &J->dFrames, &J->targetPool, &J->pos, &J->col, &J->uv, &J->nrm, &J->info, &J->states,
                       &J->texTable, &J->texPool, &J->pixPool, &J->stage, &J->passTab, &J->tileDescs, &J->points,
                       &J->pointScr, &J->lines, &J->lineScr, &J->cnt, &J->off, &J->blockRT, &J->blockPairs, &J->rts,
                       &J->rtl, &J->keysA, &J->valsA, &J->keysB, &J->valsB, &J->scanBlocks, &J->scanTotal, &J->hist,
                       &J->histScan, &J->tileBegin, &J->tileEnd, &J->tileFrags, &J->dTileFrame, &J->clipList,
                       &J->clipCount, &J->clearSnap, &J->rtd, &J->winner, &J->vid, &J->texParams, &J->colParams,
                       &J->lightParams, &J->image, &J->dImage, &J->gTex, &J->gCol, &J->gLightPart, &J->gLight,
                       &J->arrPool, &J->arrTable, &J->pairOff, &J->pairRect };
    for (DevBuf* b : bufs) f(*b);
}

constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Makes `b` hold at least `bytes`.  Without an arena the buffer grows by
// hipMalloc (after the device has finished with the old block); with one,
// a new block is carved from the arena, and running out of it is an error.
int ensure(PFDjob* J, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return 0;
    const size_t want = align256(bytes < 256 ? 256 : bytes);
    if (J->arena) {
        Arena& A = *J->arena;
        if (want > A.cap - A.used) {
            snprintf(g_err, sizeof g_err, "%s: the frame needs more than the %zu-byte static scratch "
                     "(%zu bytes in use, %zu more asked)", A.owner, A.cap, A.used, want);
            return PFD_E_NOMEM;
        }
        b.p = A.base + A.used;
        b.bytes = want;
        b.borrowed = true;
        A.used += want;
        return 0;
    }
    if (b.p && !b.borrowed) {
        // Kernels already queued may still read the old block.
        (void)hipDeviceSynchronize();
        (void)hipFree(b.p);
    }
    b = DevBuf();
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) { set_error("hipMalloc", e); b.p = nullptr; return PFD_E_NOMEM; }
    b.bytes = want;
    return 0;
}

void release(DevBuf& b) {
    if (b.p && !b.borrowed) (void)hipFree(b.p);
    b = DevBuf();
}

// ---------------------------------------------------------------------------
// Copies from and to pageable host memory (the application's colour buffer,
// its textures).  hipMemcpy stages pageable memory through the runtime's
// buffers with one thread (about 10 GB/s here); these copies go through two
// page-locked chunks of the job instead, the host side of each chunk copied
// by a few threads while the DMA of the neighbouring chunk runs.
// ---------------------------------------------------------------------------

constexpr size_t HSTAGE_CHUNK = (size_t)2 << 20;

// A few threads that split one memcpy between them and the caller.  Made
// on first use and kept for the life of the process (never destroyed, so
// no thread outlives the state it waits on).
class CopyPool {
  public:
    static CopyPool& get() {
        static CopyPool* pool = new CopyPool();
        return *pool;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        if (workers_.empty() || bytes < ((size_t)256 << 10)) { memcpy(dst, src, bytes); return; }
        std::lock_guard<std::mutex> one(callLock_);          // one parallel copy at a time
        const size_t parts = workers_.size() + 1;
        const size_t per = ((bytes + parts - 1) / parts + 63) & ~(size_t)63;
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = (uint8_t*)dst; src_ = (const uint8_t*)src; bytes_ = bytes; per_ = per;
            pending_ = workers_.size();
            gen_++;
        }
        cv_.notify_all();
        part(0);                                              // the caller's own share
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }
  private:
    CopyPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n = hw >= 16 ? 3u : hw >= 4 ? 1u : 0u;
        for (unsigned i = 0; i < n; i++) workers_.emplace_back([this, i] { run(i + 1); });
        for (auto& t : workers_) t.detach();
    }
    void part(size_t k) {
        const size_t b = k * per_;
        if (b < bytes_) memcpy(dst_ + b, src_ + b, std::min(per_, bytes_ - b));
    }
    void run(size_t k) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            part(k);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex callLock_, m_;
    std::condition_variable cv_, done_;
    uint8_t* dst_ = nullptr;
    const uint8_t* src_ = nullptr;
    size_t bytes_ = 0, per_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

// True when `p` is page-locked host memory (a direct DMA source or target).
bool host_pinned(const void* p) {
    hipPointerAttribute_t a;
    memset(&a, 0, sizeof a);
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeHost;
}

// The job's staging chunks; false when they cannot be made (or the job may
// not allocate), and the caller copies directly.
bool hstage_ready(PFDjob* J) {
    if (J->arena) return false;                              // pfdExecuteFrameNoAlloc: no allocation
    if (J->hstage[0]) return true;
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(&J->hstage[i], HSTAGE_CHUNK, hipHostMallocDefault) != hipSuccess
            || hipEventCreateWithFlags(&J->hstageEv[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (int k = 0; k < 2; k++) {
                if (J->hstage[k]) (void)hipHostFree(J->hstage[k]);
                if (J->hstageEv[k]) (void)hipEventDestroy(J->hstageEv[k]);
                J->hstage[k] = nullptr;
                J->hstageEv[k] = nullptr;
            }
            return false;
        }
    }
    return true;
}

// Host -> device on the job's stream; returns with the host memory read.
hipError_t staged_h2d(PFDjob* J, void* dev, const void* host, size_t bytes) {
    if (bytes < HSTAGE_CHUNK || host_pinned(host) || !hstage_ready(J)) {
        hipError_t e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, J->stream);
        return e == hipSuccess ? hipStreamSynchronize(J->stream) : e;
    }
    hipError_t e = hipSuccess;
    for (size_t off = 0, i = 0; e == hipSuccess && off < bytes; off += HSTAGE_CHUNK, i++) {
        const size_t n = std::min(HSTAGE_CHUNK, bytes - off), b = i & 1;
        if (i >= 2) e = hipEventSynchronize(J->hstageEv[b]);   // this chunk's last DMA is done
        if (e != hipSuccess) break;
        CopyPool::get().copy(J->hstage[b], (const uint8_t*)host + off, n);
        e = hipMemcpyAsync((uint8_t*)dev + off, J->hstage[b], n, hipMemcpyHostToDevice, J->stream);
        if (e == hipSuccess) e = hipEventRecord(J->hstageEv[b], J->stream);
    }
    return e == hipSuccess ? hipStreamSynchronize(J->stream) : e;
}

// Device -> host after the work queued on the job's stream; returns with
// the host memory written.
hipError_t staged_d2h(PFDjob* J, void* host, const void* dev, size_t bytes) {
    if (bytes < HSTAGE_CHUNK || host_pinned(host) || !hstage_ready(J)) {
        hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, J->stream);
        return e == hipSuccess ? hipStreamSynchronize(J->stream) : e;
    }
    const size_t chunks = (bytes + HSTAGE_CHUNK - 1) / HSTAGE_CHUNK;
    auto issue = [&](size_t i) -> hipError_t {
        const size_t off = i * HSTAGE_CHUNK, n = std::min(HSTAGE_CHUNK, bytes - off);
        hipError_t ee = hipMemcpyAsync(J->hstage[i & 1], (const uint8_t*)dev + off, n, hipMemcpyDeviceToHost, J->stream);
        return ee == hipSuccess ? hipEventRecord(J->hstageEv[i & 1], J->stream) : ee;
    };
    hipError_t e = issue(0);
    for (size_t i = 0; e == hipSuccess && i < chunks; i++) {
        if (i + 1 < chunks) e = issue(i + 1);                  // its chunk's host copy (i - 1) is done
        if (e == hipSuccess) e = hipEventSynchronize(J->hstageEv[i & 1]);
        if (e != hipSuccess) break;
        const size_t off = i * HSTAGE_CHUNK;
        CopyPool::get().copy((uint8_t*)host + off, J->hstage[i & 1], std::min(HSTAGE_CHUNK, bytes - off));
    }
    return e;
}

uint32_t read_total(PFDjob* J, int* rc) {
    uint32_t v = 0;
    hipError_t e = hipMemcpyAsync(&v, J->scanTotal.p, 4, hipMemcpyDeviceToHost, J->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
    if (e != hipSuccess) { set_error("read_total", e); *rc = PFD_E_LAUNCH; }
    return v;
}

uint32_t grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    return g == 0 ? 1u : (uint32_t)g;
}

// The raster kernel's TK for the job's targets.
int target_class(const PFDjob* J) { return J->anyWide ? 2 : (J->anyNarrow ? 1 : 0); }

// tk: the job's targets (the TK of the raster kernel); jobs with a target
// other than RGBA8 run the K = 1 instances that compile its path.
template <int TS, int MODE, int K>
void launch_raster(const RasterArgs& ra, uint32_t numTiles, bool phong, int tk, hipStream_t s) {
    constexpr int BW = (K == 1) ? 8 : 16;
    uint32_t grid = (TS == BW) ? (numTiles + 3) / 4 : numTiles;
    if (tk != 0) {
        const uint32_t g1 = (TS == 8) ? (numTiles + 3) / 4 : numTiles;
        if (tk == 2) {
            if (phong) hipLaunchKernelGGL((k_raster<TS, MODE, true, 1, 2>), dim3(g1), dim3(256), 0, s, ra);
            else       hipLaunchKernelGGL((k_raster<TS, MODE, false, 1, 2>), dim3(g1), dim3(256), 0, s, ra);
        } else {
            if (phong) hipLaunchKernelGGL((k_raster<TS, MODE, true, 1, 1>), dim3(g1), dim3(256), 0, s, ra);
            else       hipLaunchKernelGGL((k_raster<TS, MODE, false, 1, 1>), dim3(g1), dim3(256), 0, s, ra);
        }
        return;
    }
    // The instances that fit 64 VGPRs without spilling (checked with
    // -Rpass-analysis=kernel-resource-usage) take k_raster_o8.
    constexpr bool O8 = K == 1 && ((TS == 16 && MODE == 0) || (TS >= 16 && MODE == 1));
    if (phong) {
        hipLaunchKernelGGL((k_raster<TS, MODE, true, K, 0>), dim3(grid), dim3(256), 0, s, ra);
        return;
    }
    if constexpr (O8) hipLaunchKernelGGL((k_raster_o8<TS, MODE, K>), dim3(grid), dim3(256), 0, s, ra);
    else hipLaunchKernelGGL((k_raster<TS, MODE, false, K, 0>), dim3(grid), dim3(256), 0, s, ra);
}

template <int MODE, int K>
void launch_raster_ts(const RasterArgs& ra, uint32_t shift, uint32_t numTiles, bool phong, int tk, hipStream_t s) {
    switch (shift) {
    case 3: launch_raster<8, MODE, 1>(ra, numTiles, phong, tk, s); break;   // a 8x8 tile is one K = 1 block
    case 4: launch_raster<16, MODE, K>(ra, numTiles, phong, tk, s); break;
    case 5: launch_raster<32, MODE, K>(ra, numTiles, phong, tk, s); break;
    default: launch_raster<64, MODE, K>(ra, numTiles, phong, tk, s); break;
    }
}

void layout_tiles(PFDjob* J);
int upload_frames_table(PFDjob* J);

// Automatic configuration from the first pass's counts at 8x8 tiles: while
// raster triangles span more than 32 tiles on average, tiles double in
// side (each doubling divides the pairs of large triangles by about four),
// as long as at least 1024 tiles remain to spread over the chip's 256 CUs;
// triangles still spanning more than 16 tiles are shaded in place (raster
// mode 1), the rest packed (mode 0).  Returns the tile shift.
uint32_t auto_config(PFDjob* J, uint32_t numRT, uint32_t numPairs) {
    uint32_t s = J->tileShift;
    double per = (double)numPairs / (double)(numRT ? numRT : 1);
    auto tilesAt = [&](uint32_t shift) {
        uint64_t tiles = 0;
        for (const DFrame& F : J->frames)
            tiles += (uint64_t)((F.W + (1u << shift) - 1) >> shift) * ((F.H + (1u << shift) - 1) >> shift);
        return tiles;
    };
    while (s < 6 && per > 32.0) {
        if (tilesAt(s + 1) < 1024) break;
        s++;
        per *= 0.25;
    }
    // The packed raster at 16x16 tiles stages each 8x8 quadrant's own list
    // entries, which beats 8x8 tiles for small triangles too (1080p height
    // field 0.177 against 0.189 ms, 4K spheres 0.625 against 0.660 ms),
    // while the frames keep 1024 tiles.
    if (s == 3 && tilesAt(4) >= 1024) {
        s = 4;
        per *= 0.25;
    }
    J->cfg.rasterMode = per > 16.0 ? 1u : 0u;
    return s;
}

// The clip kernel of a pass for `clipped` listed triangles (an estimate
// before the pass is sized; either kernel walks the whole list): lane
// groups up to 16k triangles (1080p height field: 2.3k), one lane per
// triangle beyond (256-context batch: about 75k).
template <int PHASE, bool LIT, int VF>
void launch_clip(const SetupArgs& sa, uint32_t clipped, hipStream_t s) {
    if (clipped <= 16384u)
        hipLaunchKernelGGL((k_clip_group<PHASE, LIT, VF>), dim3(std::max(1u, (clipped + 3) / 4)), dim3(64), 0, s, sa);
    else
        hipLaunchKernelGGL((k_clip_lane<PHASE, LIT, VF>), dim3(std::min(4096u, (clipped + 63) / 64)), dim3(64), 0, s, sa);
}

// Radix blocks of ROUNDS x 256 pairs, by measurement (binning time of
// 0.36 M to 5.1 M pairs): blocks of 1,024 up to a million pairs, of 2,048
// up to four million, of 4,096 beyond.
uint32_t radix_rounds(uint32_t pairs) {
    if (pairs <= 1u << 20) return 4;
    if (pairs <= 4u << 20) return 8;
    return 16;
}

// One histogram + column scan + scatter pass of the pair sort, DB-bit
// digits from bit `shift`.
template <int DB, int ROUNDS>
void launch_radix_pass(PFDjob* J, const uint32_t* kIn, const uint32_t* vIn, uint32_t* kOut, uint32_t* vOut,
                       uint32_t n, uint32_t shift, uint32_t rb) {
    hipLaunchKernelGGL((k_radix_hist<DB, ROUNDS>), dim3(rb), dim3(RADIX_THREADS), 0, J->stream, kIn, n, shift, rb,
                       (uint32_t*)J->hist.p);
    hipLaunchKernelGGL(k_radix_colscan, dim3(1u << DB), dim3(RADIX_THREADS), 0, J->stream, (uint32_t*)J->hist.p, rb,
                       (uint32_t*)J->histScan.p);
    hipLaunchKernelGGL((k_radix_scatter<DB, ROUNDS>), dim3(rb), dim3(RADIX_THREADS), 0, J->stream, kIn, vIn, n, shift,
                       rb, (const uint32_t*)J->hist.p, (const uint32_t*)J->histScan.p, kOut, vOut);
}

template <int ROUNDS>
void launch_radix_pass_db(uint32_t db, PFDjob* J, const uint32_t* kIn, const uint32_t* vIn, uint32_t* kOut,
                          uint32_t* vOut, uint32_t n, uint32_t shift, uint32_t rb) {
    if (db == 9) launch_radix_pass<9, ROUNDS>(J, kIn, vIn, kOut, vOut, n, shift, rb);
    else launch_radix_pass<8, ROUNDS>(J, kIn, vIn, kOut, vOut, n, shift, rb);
}

// Digit width of the sort's passes for keys of `bits` bits: as few passes
// as digits of at most 9 bits allow, the bits split evenly, never below 8
// bits.  Measured: 4K frame at 8x8 tiles (17-bit keys, 1.4 M pairs) two
// 9-bit passes 67 us against three 8-bit ones 81 us; one 11-bit pass for
// 800x600's 11-bit keys (5.1 M pairs) 198 us against two 8-bit ones 173;
// 10 + 9 bits for the batch's 19-bit keys no faster than three passes.
uint32_t radix_digit_bits(uint32_t bits) {
    const uint32_t passes = (bits + 8) / 9;
    return std::max(8u, (bits + passes - 1) / passes);
}

// Issues one pass.  With discover set, the raster-triangle and pair counts
// are read back (synchronising) and scratch buffers are sized; otherwise
// the sizes of the previous run are used and nothing synchronises, so the
// pass can be captured in a graph.
int run_pass(PFDjob* J, PassInfo& P, bool discover, hipEvent_t* ev) {
    int rc = 0;
    if (P.numPassTris == 0 && P.segs.empty()) return 0;
    const uint32_t n = P.numPassTris;
    const uint32_t nb = grid_for(n, 256);
    if ((rc = ensure(J, J->cnt, (size_t)n * 4 + 4)) || (rc = ensure(J, J->off, (size_t)n * 4 + 4))) return rc;
    if ((rc = ensure(J, J->blockRT, (size_t)nb * 4 + 4)) || (rc = ensure(J, J->blockPairs, (size_t)nb * 4 + 4))) return rc;
    if ((rc = ensure(J, J->scanTotal, 16))) return rc;
    if ((rc = ensure(J, J->clipList, (size_t)n * 16 + 16)) || (rc = ensure(J, J->clipCount, 16))) return rc;
    if ((rc = ensure(J, J->clearSnap, J->frames.size() * 32))) return rc;
    const size_t fragSlots = (size_t)J->numTiles * 4;
    const size_t passIdx = (size_t)(&P - J->passes.data());
    if (discover) {
        // Slots no wave writes (TS = 8 uses one of four) must read as zero.
        if ((rc = ensure(J, J->tileFrags, fragSlots * J->passes.size() * 4 + 4))) return rc;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(fragSlots, 256 * 8)), dim3(256), 0, J->stream,
                           (uint32_t*)J->tileFrags.p + passIdx * fragSlots, (uint32_t)fragSlots, 0u);
    }
    if ((rc = ensure(J, J->tileBegin, (size_t)J->numTiles * 4 + 4)) || (rc = ensure(J, J->tileEnd, (size_t)J->numTiles * 4 + 4)))
        return rc;
    if (!discover) {
        // sizes known from the sizing run; buffers already hold them
    }
    SetupArgs sa;
    memset(&sa, 0, sizeof sa);
    sa.pos = (const float4*)J->pos.p;
    sa.col = (const uint32_t*)J->col.p;
    sa.uv = J->hasUV ? (const float2*)J->uv.p : nullptr;
    sa.nrm = J->hasNrm ? (const float*)J->nrm.p : nullptr;
    sa.info = (const uint32_t*)J->info.p;
    sa.states = (const PFFstate*)J->states.p;
    sa.texs = (const DTex*)J->texTable.p;
    sa.segs = P.dSegs;
    sa.frames = (const DFrame*)J->dFrames.p;
    sa.numSegs = (uint32_t)P.segs.size();
    sa.blockSeg = P.dBlockSeg;
    sa.numPassTris = n;
    sa.tileShift = J->tileShift;
    sa.numTiles = J->numTiles;
    sa.cntRT = (uint32_t*)J->cnt.p;
    sa.cntPairs = (uint32_t*)J->off.p;
    sa.blockRT = (uint32_t*)J->blockRT.p;
    sa.blockPairs = (uint32_t*)J->blockPairs.p;
    sa.tileBegin = (uint32_t*)J->tileBegin.p;
    sa.tileEnd = (uint32_t*)J->tileEnd.p;
    sa.clipList = (uint4*)J->clipList.p;
    sa.clipCount = (uint32_t*)J->clipCount.p;
    sa.clipCap = n;
    // A sized pass lists the same triangles on every run (the same inputs).
    sa.clipKnown = discover ? ~0u : P.numClip;
    sa.clearSnap = (uint32_t*)J->clearSnap.p;

    if (ev) HIP_TRY(hipEventRecord(ev[0], J->stream), PFD_E_LAUNCH);
    // Replays with few blocks leave the block sums raw for k_scan_local<true>.
    const bool fusedScan = !discover && nb <= 1024;
    // Triangles to clip: counted once the pass is sized, estimated before.
    const uint32_t clip0 = discover ? n / 64 : P.numClip;
    for (;;) {
        // Every pass leaves the clip-list counter at zero (k_tile_bounds, or
        // the fill at its end); a discovering run clears it here as well.
        if (discover)
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, J->stream, (uint32_t*)J->clipCount.p, 1u, 0u);
        hipLaunchKernelGGL((k_setup<0, false>), dim3(nb), dim3(256), 0, J->stream, sa);
        if (discover || P.numClip) launch_clip<0, false, VF_POS>(sa, clip0, J->stream);
        if (!fusedScan)
            hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, J->stream,
                               (uint32_t*)J->blockRT.p, (uint32_t*)J->blockPairs.p, nb, (uint32_t*)J->scanTotal.p);
        HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
        if (!discover) break;
        uint32_t tot[2] = {0, 0}, clipped = 0;
        HIP_TRY(hipMemcpyAsync(tot, J->scanTotal.p, 8, hipMemcpyDeviceToHost, J->stream), PFD_E_LAUNCH);
        HIP_TRY(hipMemcpyAsync(&clipped, J->clipCount.p, 4, hipMemcpyDeviceToHost, J->stream), PFD_E_LAUNCH);
        HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
        P.numRT = tot[0];
        P.numPairs = tot[1];
        P.numClip = std::min(clipped, n);
        // A job with an automatic configuration chooses its tiles from the
        // first pass's counts, before anything is drawn: the counts are
        // taken again at the chosen size (the clip list is refilled).
        if (!J->autoTiles || passIdx != 0) break;
        J->autoTiles = false;
        const uint32_t ts = auto_config(J, tot[0], tot[1]);
        if (ts == J->tileShift) break;
        J->tileShift = ts;
        layout_tiles(J);
        if ((rc = upload_frames_table(J))) return rc;
        sa.tileShift = J->tileShift;
        sa.numTiles = J->numTiles;
    }
    sa.clipKnown = P.numClip;                            // counted by now
    const uint32_t numRT = P.numRT, numPairs = P.numPairs;
    if ((rc = ensure(J, J->rts, (size_t)numRT * sizeof(RTri) + sizeof(RTri)))) return rc;
    if (J->anyPhong && (rc = ensure(J, J->rtl, (size_t)numRT * sizeof(RTriL) + sizeof(RTriL)))) return rc;
    if (J->diff && (rc = ensure(J, J->rtd, (size_t)numRT * sizeof(RTriD) + sizeof(RTriD)))) return rc;
    const size_t pb = (size_t)numPairs * 4 + 4;
    if ((rc = ensure(J, J->keysA, pb)) || (rc = ensure(J, J->valsA, pb)) || (rc = ensure(J, J->keysB, pb)) || (rc = ensure(J, J->valsB, pb)))
        return rc;
    sa.rts = (RTri*)J->rts.p;
    sa.rtl = (RTriL*)J->rtl.p;
    sa.rtd = J->diff ? (RTriD*)J->rtd.p : nullptr;
    sa.keys = (uint32_t*)J->keysA.p;
    sa.vals = (uint32_t*)J->valsA.p;
    if (numRT > PAIR_RT_MASK) {
        snprintf(g_err, sizeof g_err, "pass of %u raster triangles: at most %u", numRT, PAIR_RT_MASK);
        return PFD_E_INVALID;
    }
    sa.capRT = numRT;
    sa.capPairs = numPairs;
    // Raster triangles of more than 32 tile pairs on average (large
    // triangles: 800x600 frame of 10k, 510 each): the setup records one
    // rectangle per raster triangle and k_expand_pairs writes the pairs,
    // one thread per four, instead of the setup's lanes looping over them.
    const bool expand = numRT && numPairs / numRT > 32u;
    if (expand) {
        if ((rc = ensure(J, J->pairOff, (size_t)numRT * 4 + 4)) || (rc = ensure(J, J->pairRect, (size_t)numRT * 16 + 16)))
            return rc;
        sa.pairOff = (uint32_t*)J->pairOff.p;
        sa.pairRect = (uint4*)J->pairRect.p;
    }
    if (numRT) {
        // Slots per triangle, then the clipped triangles and the others set
        // up (both read the slots only).  Many clipped triangles (one lane
        // each) are clipped on the side stream beside k_setup<1>; a few (lane
        // groups) by k_setup<1>'s own leading workgroups in unlit passes
        // without float records, else by a kernel on the main stream before
        // it, where the fork and join would cost more than they overlap
        // (1080p frame: setup 65 -> 53 us); none, not at all.
        if (fusedScan) hipLaunchKernelGGL(k_scan_local<true>, dim3(nb), dim3(256), 0, J->stream, sa);
        else hipLaunchKernelGGL(k_scan_local<false>, dim3(nb), dim3(256), 0, J->stream, sa);
        const bool useSide = J->side && P.numClip > 16384u;
        hipStream_t clipStream = useSide ? J->side : J->stream;
        if (useSide) {
            HIP_TRY(hipEventRecord(J->fork, J->stream), PFD_E_LAUNCH);
            HIP_TRY(hipStreamWaitEvent(J->side, J->fork, 0), PFD_E_LAUNCH);
        }
        if (J->anyLit) {
            if (P.numClip) launch_clip<1, true, VF_FULL>(sa, P.numClip, clipStream);
            hipLaunchKernelGGL((k_setup<1, true>), dim3(nb), dim3(256), 0, J->stream, sa);
        } else if (!J->diff && P.numClip && P.numClip <= 16384u) {
            // A few clipped triangles: clipped by k_setup<1>'s leading
            // workgroups (1080p frame, 2,468 clipped: setup 50 -> 44 us).
            sa.clipBlocks = (P.numClip + 15) / 16;
            hipLaunchKernelGGL((k_setup<1, false, false, true>), dim3(nb + sa.clipBlocks), dim3(256), 0, J->stream, sa);
        } else {
            if (P.numClip && J->diff) launch_clip<1, false, VF_FULL>(sa, P.numClip, clipStream);
            else if (P.numClip) launch_clip<1, false, VF_SLIM>(sa, P.numClip, clipStream);
            if (J->diff) hipLaunchKernelGGL((k_setup<1, false, true>), dim3(nb), dim3(256), 0, J->stream, sa);
            else hipLaunchKernelGGL((k_setup<1, false, false>), dim3(nb), dim3(256), 0, J->stream, sa);
        }
        if (useSide) {
            HIP_TRY(hipEventRecord(J->join, J->side), PFD_E_LAUNCH);
            HIP_TRY(hipStreamWaitEvent(J->stream, J->join, 0), PFD_E_LAUNCH);
        }
    }
    if (expand)
        hipLaunchKernelGGL(k_expand_pairs, dim3(grid_for((numPairs + 3) / 4, 256)), dim3(256), 0, J->stream,
                           (const uint32_t*)J->pairOff.p, (const uint4*)J->pairRect.p, numRT, numPairs,
                           (uint32_t*)J->keysA.p, (uint32_t*)J->valsA.p);
    HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
    if (ev) HIP_TRY(hipEventRecord(ev[1], J->stream), PFD_E_LAUNCH);

    // Stable sort of the pairs by tile, then per-tile ranges.
    uint32_t* keys = (uint32_t*)J->keysA.p;
    uint32_t* vals = (uint32_t*)J->valsA.p;
    if (numPairs) {
        uint32_t bits = 1;
        while (bits < 32 && (1u << bits) < J->numTiles) bits++;
        const uint32_t rounds = radix_rounds(numPairs);
        const uint32_t rb = (numPairs + RADIX_THREADS * rounds - 1) / (RADIX_THREADS * rounds);
        const uint32_t db = radix_digit_bits(bits);
        const size_t nh = (size_t)rb << db;
        if ((rc = ensure(J, J->hist, nh * 4 + 4)) || (rc = ensure(J, J->histScan, ((size_t)4 << db) + 4))) return rc;
        uint32_t* kIn = keys; uint32_t* vIn = vals;
        uint32_t* kOut = (uint32_t*)J->keysB.p; uint32_t* vOut = (uint32_t*)J->valsB.p;
        for (uint32_t shift = 0; shift < bits; shift += db) {
            switch (rounds) {
            case 4:  launch_radix_pass_db<4>(db, J, kIn, vIn, kOut, vOut, numPairs, shift, rb); break;
            case 8:  launch_radix_pass_db<8>(db, J, kIn, vIn, kOut, vOut, numPairs, shift, rb); break;
            default: launch_radix_pass_db<16>(db, J, kIn, vIn, kOut, vOut, numPairs, shift, rb); break;
            }
            std::swap(kIn, kOut);
            std::swap(vIn, vOut);
        }
        keys = kIn;
        vals = vIn;
        hipLaunchKernelGGL(k_tile_bounds, dim3(grid_for(numPairs, 256 * 4)), dim3(256), 0, J->stream,
                           keys, numPairs, J->numTiles, (uint32_t*)J->tileBegin.p, (uint32_t*)J->tileEnd.p,
                           (uint32_t*)J->clipCount.p);
    }
    if (!numPairs)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, J->stream, (uint32_t*)J->clipCount.p, 1u, 0u);
    HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
    if (ev) HIP_TRY(hipEventRecord(ev[2], J->stream), PFD_E_LAUNCH);

    RasterArgs ra;
    ra.rts = (const RTri*)J->rts.p;
    ra.rtl = (const RTriL*)J->rtl.p;
    ra.states = (const PFFstate*)J->states.p;
    ra.tileFrame = (const uint32_t*)J->dTileFrame.p;
    ra.tileDesc = P.dTileDesc;
    ra.tileBegin = (const uint32_t*)J->tileBegin.p;
    ra.tileEnd = (const uint32_t*)J->tileEnd.p;
    ra.vals = vals;
    ra.numTiles = J->numTiles;
    ra.tileFrags = (uint32_t*)J->tileFrags.p + passIdx * fragSlots;
    ra.winner = J->diff ? (uint32_t*)J->winner.p : nullptr;
    ra.clearSnap = (const uint32_t*)J->clearSnap.p;
    switch (J->cfg.rasterMode) {
    case 1: launch_raster_ts<1, 1>(ra, J->tileShift, J->numTiles, J->anyPhong, target_class(J), J->stream); break;
    case 2: launch_raster_ts<0, 4>(ra, J->tileShift, J->numTiles, J->anyPhong, target_class(J), J->stream); break;
    case 3: launch_raster_ts<1, 4>(ra, J->tileShift, J->numTiles, J->anyPhong, target_class(J), J->stream); break;
    default: launch_raster_ts<0, 1>(ra, J->tileShift, J->numTiles, J->anyPhong, target_class(J), J->stream); break;
    }
    if (!P.pix.empty())
        hipLaunchKernelGGL(k_draw_pixels, dim3(P.pixGrid, (uint32_t)P.pix.size()), dim3(256), 0, J->stream,
                           P.dPix);
    if (!P.pts.empty())
        hipLaunchKernelGGL(k_points, dim3(P.ptTiles, (uint32_t)P.pts.size()), dim3(256), 0, J->stream,
                           P.dPts, (const DPointScr*)J->pointScr.p);
    if (!P.lns.empty())
        hipLaunchKernelGGL(k_lines, dim3(P.lnTiles, (uint32_t)P.lns.size()), dim3(256), 0, J->stream,
                           P.dLns, (const DLineScr*)J->lineScr.p);
    HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
    if (ev) HIP_TRY(hipEventRecord(ev[3], J->stream), PFD_E_LAUNCH);
    P.sized = true;
    return 0;
}

int run_all(PFDjob* J, bool discover) {
    int rc = 0;
    for (auto& P : J->passes)
        if ((rc = run_pass(J, P, discover || !P.sized, nullptr))) {
            // A run that stopped part-way may leave device counters (the
            // clip list's) off zero; the next run rediscovers, which clears
            // them before use.
            for (auto& Q : J->passes) Q.sized = false;
            return rc;
        }
    return 0;
}

// Device time of `iters` runs launched straight onto the job's stream (no
// graph), after a sizing run if one is due.
int time_direct(PFDjob* J, uint32_t iters, float* ms) {
    int rc = 0;
    bool allSized = true;
    for (auto& P : J->passes) allSized = allSized && P.sized;
    if (!allSized && (rc = run_all(J, true))) return rc;
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, J->stream);
    for (uint32_t i = 0; e == hipSuccess && !rc && i < iters; i++) rc = run_all(J, false);
    if (e == hipSuccess && !rc) e = hipEventRecord(b, J->stream);
    if (e == hipSuccess && !rc) e = hipEventSynchronize(b);
    float t = 0.0f;
    if (e == hipSuccess && !rc) e = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) { set_error("time_direct", e); return PFD_E_LAUNCH; }
    *ms = t;
    return PFD_OK;
}

void layout_tiles(PFDjob* J) {
    uint32_t ts = J->tileShift, base = 0;
    for (auto& F : J->frames) {
        F.tilesX = (F.W + (1u << ts) - 1) >> ts;
        F.tilesY = (F.H + (1u << ts) - 1) >> ts;
        F.tileBase = base;
        base += F.tilesX * F.tilesY;
    }
    J->numTiles = base;
}

// Frame table, tile -> frame map and the per-pass tile descriptors; all
// depend on the tile size, so they are uploaded again when it changes.
int upload_frames_table(PFDjob* J) {
    int rc = ensure(J, J->dFrames, J->frames.size() * sizeof(DFrame));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(J->dFrames.p, J->frames.data(), J->frames.size() * sizeof(DFrame), hipMemcpyHostToDevice), PFD_E_LAUNCH);
    std::vector<uint32_t> tf(J->numTiles ? J->numTiles : 1, 0u);
    for (uint32_t f = 0; f < J->frames.size(); f++) {
        const DFrame& F = J->frames[f];
        for (uint32_t t = 0; t < F.tilesX * F.tilesY; t++) tf[F.tileBase + t] = f;
    }
    if ((rc = ensure(J, J->dTileFrame, tf.size() * 4))) return rc;
    HIP_TRY(hipMemcpy(J->dTileFrame.p, tf.data(), tf.size() * 4, hipMemcpyHostToDevice), PFD_E_LAUNCH);
    const size_t nf = J->frames.size();
    std::vector<DTileFrame> d(J->passes.size() * nf);
    memset(d.data(), 0, d.size() * sizeof(DTileFrame));
    for (size_t p = 0; p < J->passes.size(); p++) {
        const PassInfo& P = J->passes[p];
        for (size_t f = 0; f < nf; f++) {
            DTileFrame& t = d[p * nf + f];
            t.F = J->frames[f];
            const int32_t si = f < P.frameSeg.size() ? P.frameSeg[f] : -1;
            if (si < 0) continue;
            const DSeg& sg = P.segs[si];
            t.active = 1;
            t.clear = sg.clear;
            t.clearColor = t.F.kind >= PFF_TARGET_FIRST_WIDE ? sg.clearColor      // wide: the RGBA8 value
                                                             : tgt_set_simd(t.F.kind, sg.clearColor);   // stored form
            t.clearDepth = sg.clearDepth;
        }
    }
    if (d.empty()) return 0;
    if ((rc = ensure(J, J->tileDescs, d.size() * sizeof(DTileFrame)))) return rc;
    HIP_TRY(hipMemcpy(J->tileDescs.p, d.data(), d.size() * sizeof(DTileFrame), hipMemcpyHostToDevice), PFD_E_LAUNCH);
    for (size_t p = 0; p < J->passes.size(); p++)
        J->passes[p].dTileDesc = (const DTileFrame*)J->tileDescs.p + p * nf;
    return 0;
}

void drop_graph(PFDjob* J) {
    if (J->graphExec) { (void)hipGraphExecDestroy(J->graphExec); J->graphExec = nullptr; }
    if (J->graph) { (void)hipGraphDestroy(J->graph); J->graph = nullptr; }
}

int job_shell(PFDjob** out, bool withSide);

// State of pfdExecuteFrameNoAlloc, per device: the address of its copy of
// `scratch`, the arena over it and a job (host tables and one stream) made
// once at set-up.  Guarded by g_noAllocLock.
constexpr int PFD_MAX_DEVICES = 64;
uint8_t*   g_scratchPtr[PFD_MAX_DEVICES];
Arena      g_noAllocArena[PFD_MAX_DEVICES];
PFDjob*    g_noAllocJob[PFD_MAX_DEVICES];
std::mutex g_noAllocLock;

// Set-up of the current device `dev` for pfdExecuteFrameNoAlloc, once per
// device, with g_noAllocLock held.
int no_alloc_setup(int dev) {
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) { set_error("hipGetSymbolAddress(scratch)", e); return PFD_E_NODEVICE; }
        g_scratchPtr[dev] = (uint8_t*)p;
    }
    if (!g_noAllocJob[dev]) {
        PFDjob* J = nullptr;
        int rc = job_shell(&J, false);    // the clipper then runs on the job's one stream
        if (rc) return rc;
        Arena& A = g_noAllocArena[dev];
        A.base = g_scratchPtr[dev];
        A.cap = PFD_SCRATCH_BYTES;
        A.used = 0;
        A.owner = "pfdExecuteFrameNoAlloc";
        J->arena = &A;
        g_noAllocJob[dev] = J;
    }
    return PFD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int pfdInit(void) {
    if (g_initState == 1) return PFD_OK;
    if (g_initState == -1) return PFD_E_NODEVICE;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("hipGetDeviceCount (no GPU visible)", e);
        g_initState = -1;
        return PFD_E_NODEVICE;
    }
    e = hipSetDevice(0);
    if (e != hipSuccess) { set_error("hipSetDevice", e); g_initState = -1; return PFD_E_NODEVICE; }
    {
        std::lock_guard<std::mutex> lk(g_noAllocLock);
        const int rc = no_alloc_setup(0);
        if (rc) { g_initState = -1; return rc; }
    }
    g_initState = 1;
    return PFD_OK;
}

const char* pfdLastError(void) { return g_err; }

static PFDconfig g_defaultCfg = {0u, 0u, {0, 0, 0, 0, 0, 0}};   // automatic (auto_config)

int pfdSetDefaultConfig(const PFDconfig* cfg) {
    if (!cfg) return PFD_E_INVALID;
    g_defaultCfg = *cfg;
    return PFD_OK;
}

}  // extern "C"

namespace {

// A job with no frames yet: its stream and, with `withSide`, the side
// stream and the fork/join events of the clipper.
int job_shell(PFDjob** out, bool withSide) {
    PFDjob* J = new PFDjob();
    hipError_t e = hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking);
    if (e == hipSuccess && withSide) {
        if ((e = hipStreamCreateWithFlags(&J->side, hipStreamNonBlocking)) == hipSuccess
            && (e = hipEventCreateWithFlags(&J->fork, hipEventDisableTiming)) == hipSuccess)
            e = hipEventCreateWithFlags(&J->join, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        set_error("hipStreamCreate", e);
        pfdJobDestroy(J);
        return PFD_E_NODEVICE;
    }
    *out = J;
    return PFD_OK;
}

// Forgets the frames of a job and keeps its streams and device buffers
// (which the next plan reuses; an arena job's buffers are re-carved).
void job_reset(PFDjob* J) {
    if (J->stream) (void)hipStreamSynchronize(J->stream);
    if (J->side) (void)hipStreamSynchronize(J->side);
    drop_graph(J);
    J->src.clear();
    J->leadClear.clear();
    J->srcTex.clear();
    J->srcStates.clear();
    J->frames.clear();
    J->passes.clear();
    J->numIn = 0;
    J->hasUV = J->hasNrm = J->anyPhong = J->anyLit = false;
    J->anyWide = J->anyNarrow = false;
    J->numTiles = 0;
    J->autoTiles = false;
    J->sortedInB = 0;
    J->lastFragments = J->totalRT = J->totalPairs = 0;
    J->diff = J->diffRan = false;
    J->numVertices = 0;
    J->paramTex = -1;
    if (J->arena) {
        J->arena->used = 0;
        for_each_buf(J, [](DevBuf& b) { b = DevBuf(); });
    }
}

#define PLAN_FAIL(code, ...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return (code); } while (0)

// Calls f(begin, end) for the triangle ranges of a frame that the host
// recorded itself, i.e. outside its array draws (whose slots the host
// leaves unwritten and k_deindex fills).
template <typename F>
void for_host_ranges(const PFFframe& fr, F f) {
    uint32_t at = 0;
    for (uint32_t k = 0; k < fr.numArrays; k++) {
        const PFFarrays& a = fr.arrays[k];
        if (a.triFirst > at) f(at, a.triFirst);
        at = a.triFirst + a.triCount;
    }
    if (fr.tris.count > at) f(at, fr.tris.count);
}

size_t elem_bytes(uint32_t type) {
    switch (type) {
    case PFF_UBYTE: return 1;
    case PFF_USHORT: case PFF_SHORT: return 2;
    case PFF_DOUBLE: return 8;
    default: return 4;
    }
}

// Checks an array draw of a frame with `nt` triangles and `ns` states
// against the rules of PFFarrays; returns a message or null.
const char* check_arrays(const PFFarrays& a, uint32_t prevEnd, uint32_t nt, uint32_t ns) {
    if (a.mode < 2u || a.mode > 7u) return "array draw: bad mode";
    if (a.numFaces < 1u || a.numFaces > 2u || a.state >= ns) return "array draw: bad faces or state";
    if (!a.position || a.posSize < 2 || a.posSize > 4
        || !(a.posType == PFF_SHORT || a.posType == PFF_INT || a.posType == PFF_FLOAT || a.posType == PFF_DOUBLE))
        return "array draw: bad positions";
    if (a.normal && !(a.nrmType == PFF_FLOAT || a.nrmType == PFF_DOUBLE)) return "array draw: bad normals";
    if (a.texcoord && !(a.uvType == PFF_FLOAT || a.uvType == PFF_DOUBLE)) return "array draw: bad texcoords";
    if (a.colors && (a.colSize < 3 || a.colSize > 4
                     || !(a.colType == PFF_UBYTE || a.colType == PFF_USHORT || a.colType == PFF_UINT
                          || a.colType == PFF_FLOAT || a.colType == PFF_DOUBLE)))
        return "array draw: bad colours";
    if (a.indices && !(a.idxType == PFF_UBYTE || a.idxType == PFF_USHORT || a.idxType == PFF_UINT))
        return "array draw: bad index type";
    if (a.numElements == 0u || (!a.indices && a.count > a.numElements)) return "array draw: too few elements";
    const uint32_t buf = a.mode == 2u ? 3u : (a.mode == 6u || a.mode == 7u) ? 6u : 4u;
    const uint32_t keep = (a.mode == 3u || a.mode == 4u) ? 1u : (a.mode == 6u || a.mode == 7u) ? 2u : 0u;
    const uint32_t per = a.mode == 2u ? 1u : (a.mode == 6u || a.mode == 7u) ? 4u : 2u;
    const uint64_t fills = a.count < buf ? 0 : 1 + (uint64_t)(a.count - buf) / (buf - keep);
    if (fills * a.numFaces * per != a.triCount) return "array draw: triangle count does not match";
    if (a.triFirst < prevEnd || (uint64_t)a.triFirst + a.triCount > nt) return "array draw: bad triangle range";
    return nullptr;
}

// Uploads the frames into the job's buffers and builds its passes:
// triangle streams, states and textures of all frames concatenated, render
// targets, decoded pfDrawPixels images, points and lines (projected once),
// and per pass its segments (one per frame at most) and pixel / point /
// line operations.  The job's configuration must be set.
int job_plan(PFDjob* J, const PFFframe* frames, uint32_t count, bool wantVid) {
    int rc = 0;
    hipError_t e = hipSuccess;

    // Sizes and validation, before anything is copied.
    uint32_t totalTris = 0, totalStates = 0;
    bool anyUV = false, anyNrm = false;
    size_t targetBytes = 0, texBytes = 0, pixBytes = 0, stageBytes = 0, arrBytes = 0;
    uint64_t arrTris = 0;
    uint32_t numDraws = 0;
    for (uint32_t f = 0; f < count; f++) {
        const PFFframe& fr = frames[f];
        if (!fr.color || !fr.depth || fr.width == 0 || fr.height == 0)
            PLAN_FAIL(PFD_E_INVALID, "frame %u has no render target", f);
        if (fr.tris.count && (!fr.tris.position || !fr.tris.color || !fr.tris.info))
            PLAN_FAIL(PFD_E_INVALID, "frame %u: triangle streams missing", f);
        if (fr.numArrays && !fr.arrays) PLAN_FAIL(PFD_E_INVALID, "frame %u: array draws missing", f);
        uint32_t prevEnd = 0;
        for (uint32_t k = 0; k < fr.numArrays; k++) {
            const PFFarrays& a = fr.arrays[k];
            const char* why = check_arrays(a, prevEnd, fr.tris.count, fr.numStates);
            if (why) PLAN_FAIL(PFD_E_INVALID, "frame %u: %s", f, why);
            prevEnd = a.triFirst + a.triCount;
            const size_t n = a.numElements;
            if (!a.device)
                arrBytes += align256(n * a.posSize * elem_bytes(a.posType)) + align256(a.normal ? n * 3 * elem_bytes(a.nrmType) : 0)
                          + align256(a.texcoord ? n * 2 * elem_bytes(a.uvType) : 0)
                          + align256(a.colors ? n * a.colSize * elem_bytes(a.colType) : 0)
                          + align256(a.indices ? (size_t)a.count * elem_bytes(a.idxType) : 0);
            else if (a.indices)
                PLAN_FAIL(PFD_E_INVALID, "frame %u: a device-resident array draw has indices", f);
            arrTris += a.triCount;
            numDraws++;
        }
        bool badState = false;
        for_host_ranges(fr, [&](uint32_t b, uint32_t e2) {
            for (uint32_t t = b; t < e2; t++) badState |= (fr.tris.info[t] & PFF_STATE_MASK) >= fr.numStates;
        });
        if (badState) PLAN_FAIL(PFD_E_INVALID, "bad state index");
        totalTris += fr.tris.count;
        totalStates += fr.numStates;
        anyUV |= fr.tris.texcoord != nullptr;
        anyNrm |= fr.tris.normal != nullptr;
        if (pff_target_bytes(fr.target) == 0u) PLAN_FAIL(PFD_E_INVALID, "frame %u: unknown render target format %u", f, fr.target);
        targetBytes += align256((size_t)fr.width * fr.height * pff_target_bytes(fr.target) + 16)
                     + align256((size_t)fr.width * fr.height * 4);
        for (uint32_t k = 0; k < fr.numTextures; k++) {
            const PFFtexture& T = fr.textures[k];
            if (T.width > 0xFFFFu || T.height > 0xFFFFu)        // raster records keep 16-bit sides
                PLAN_FAIL(PFD_E_INVALID, "texture %ux%u: sides above 65535 are not supported", T.width, T.height);
            const uint32_t srcBytes = T.format == PFF_TEX_SOURCE ? pff_source_texel_bytes(T.source) : 0u;
            if (T.format > PFF_TEX_SOURCE || (T.format == PFF_TEX_SOURCE && srcBytes == 0u))
                PLAN_FAIL(PFD_E_INVALID, "texture format %u / source 0x%x not supported", T.format, T.source);
            const size_t texels = (size_t)T.width * T.height;
            texBytes += align256(texels * (T.format == PFF_TEX_RGBAF32 ? 16 : 4));
            if (T.format == PFF_TEX_SOURCE) stageBytes = std::max(stageBytes, texels * srcBytes + 4);
        }
        for (uint32_t c = 0; c < fr.numCmds; c++) {
            const PFFcmd& cmd = fr.cmds[c];
            if (cmd.kind != PFF_CMD_PIXELS) continue;
            if (!fr.pixelOps || cmd.first >= fr.numPixelOps) PLAN_FAIL(PFD_E_INVALID, "bad pixel command");
            const PFFpixels& op = fr.pixelOps[cmd.first];
            const uint32_t bpp = pff_source_texel_bytes(op.source);
            if (bpp == 0u || op.width == 0u || op.height == 0u || !op.pixels) PLAN_FAIL(PFD_E_INVALID, "bad pixel image");
            const size_t texels = (size_t)op.width * op.height;
            pixBytes += align256(texels * 4);
            stageBytes = std::max(stageBytes, texels * bpp + 4);
        }
    }
    if ((rc = ensure(J, J->targetPool, targetBytes)) || (rc = ensure(J, J->texPool, texBytes))
        || (pixBytes && (rc = ensure(J, J->pixPool, pixBytes))) || (stageBytes && (rc = ensure(J, J->stage, stageBytes))))
        return rc;

    J->numIn = totalTris;
    J->hasUV = anyUV;
    J->hasNrm = anyNrm;
    std::vector<PFFstate> hstates(totalStates ? totalStates : 1);
    std::vector<DTex> htex;
    std::vector<PFFpoint> hpts;
    std::vector<PFFline> hlns;
    uint32_t triBase = 0, stateBase = 0, ptBase = 0, lnBase = 0;
    size_t targetOff = 0, texOff = 0, pixOff = 0;
    J->src.assign(frames, frames + count);
    J->leadClear.resize(count);
    for (uint32_t f = 0; f < count; f++)
        J->leadClear[f] = frames[f].numCmds > 0 && frames[f].cmds[0].kind == PFF_CMD_CLEAR;
    J->srcTex.assign(frames[0].textures, frames[0].textures + frames[0].numTextures);
    J->srcStates.assign(frames[0].states, frames[0].states + frames[0].numStates);
    J->frames.resize(count);
    for (uint32_t f = 0; f < count; f++) {
        const PFFframe& fr = frames[f];
        const size_t nt = fr.tris.count;
        const uint32_t texBase = (uint32_t)htex.size();
        for (uint32_t k = 0; k < fr.numTextures; k++) {
            const PFFtexture& T = fr.textures[k];
            DTex d;
            memset(&d, 0, sizeof d);
            d.w = T.width; d.h = T.height; d.format = T.format; d.wrap = T.wrap; d.filter = T.filter;
            const size_t texels = (size_t)T.width * T.height;
            const size_t bytes = texels * (T.format == PFF_TEX_RGBAF32 ? 16 : 4);
            uint8_t* dst = (uint8_t*)J->texPool.p + texOff;
            texOff += align256(bytes);
            if (bytes && T.format == PFF_TEX_SOURCE) {
                // Raw texels + 4 zero bytes (the getters' word reads run up
                // to 3 bytes past the last texel), decoded to RGBA8.
                const size_t rawBytes = texels * pff_source_texel_bytes(T.source);
                e = staged_h2d(J, J->stage.p, T.pixels, rawBytes);
                if (e == hipSuccess) e = hipMemsetAsync((uint8_t*)J->stage.p + rawBytes, 0, 4, J->stream);
                if (e == hipSuccess) {
                    hipLaunchKernelGGL(k_decode_texels, dim3((uint32_t)((texels + 255) / 256)), dim3(256), 0, J->stream,
                                       (const uint8_t*)J->stage.p, texels, T.source, (uint32_t*)dst);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipStreamSynchronize(J->stream);   // the stage is reused
                if (e != hipSuccess) { set_error("texture decode", e); return PFD_E_LAUNCH; }
            } else if (bytes) {
                e = staged_h2d(J, dst, T.pixels, bytes);
                if (e != hipSuccess) { set_error("texture upload", e); return PFD_E_LAUNCH; }
            }
            if (T.format == PFF_TEX_SOURCE) d.format = PFF_TEX_RGBA8;
            d.pixels = dst;
            htex.push_back(d);
        }
        for (uint32_t s = 0; s < fr.numStates; s++) {
            PFFstate st = fr.states[s];
            if ((st.flags & PFF_LIGHTING) && st.lightMode == 1 && st.numLights > 0) J->anyPhong = true;
            if (st.flags & PFF_LIGHTING) J->anyLit = true;
            if (st.texture >= 0) {
                if ((uint32_t)st.texture >= fr.numTextures) st.texture = -1;
                else st.texture += (int32_t)texBase;
            }
            hstates[stateBase + s] = st;
        }
        DFrame& D = J->frames[f];
        memset(&D, 0, sizeof D);
        D.W = fr.width;
        D.H = fr.height;
        const size_t npx = (size_t)fr.width * fr.height;
        D.kind = fr.target;
        D.bpp = pff_target_bytes(fr.target);
        J->anyWide = J->anyWide || fr.target >= PFF_TARGET_FIRST_WIDE;
        J->anyNarrow = J->anyNarrow || (fr.target != PFF_TARGET_RGBA8 && fr.target < PFF_TARGET_FIRST_WIDE);
        D.color = (uint8_t*)J->targetPool.p + targetOff;
        // 16 spare bytes: the wide targets' getters read whole words past a
        // pixel's last half component.
        targetOff += align256(npx * D.bpp + 16);
        D.depth = (float*)((uint8_t*)J->targetPool.p + targetOff);
        targetOff += align256(npx * 4);

        for (uint32_t k = 0; k < fr.numPoints; k++) {
            PFFpoint q = fr.points[k];
            if (q.state >= fr.numStates) PLAN_FAIL(PFD_E_INVALID, "bad point state");
            q.state += stateBase;
            hpts.push_back(q);
        }
        for (uint32_t k = 0; k < fr.numLines; k++) {
            PFFline q = fr.lines[k];
            if (q.state >= fr.numStates) PLAN_FAIL(PFD_E_INVALID, "bad line state");
            q.state += stateBase;
            if (k == 0) q.groupStart = 1u;             // groups never span frames
            hlns.push_back(q);
        }
        // Commands -> segments; the k-th segment of a frame runs in pass k.
        uint32_t segNo = 0;
        DSeg cur;
        memset(&cur, 0, sizeof cur);
        cur.frame = f;
        bool open = false, hasDraw = false;
        auto flush_seg = [&](void) {
            if (!open) return;
            if (J->passes.size() <= segNo) J->passes.emplace_back();
            J->passes[segNo].segs.push_back(cur);
            segNo++;
            memset(&cur, 0, sizeof cur);
            cur.frame = f;
            open = false;
            hasDraw = false;
        };
        // A command that runs alone after the raster kernel of its own pass.
        auto own_pass = [&](void) -> PassInfo& {
            flush_seg();
            cur.triFirst = cur.triLast = triBase;
            open = true;
            if (J->passes.size() <= segNo) J->passes.emplace_back();
            return J->passes[segNo];
        };
        for (uint32_t c = 0; c < fr.numCmds; c++) {
            const PFFcmd& cmd = fr.cmds[c];
            if (cmd.kind == PFF_CMD_CLEAR) {
                if (hasDraw) flush_seg();
                cur.clear = 1;
                cur.clearColor = cmd.clearColor;
                cur.clearDepth = cmd.clearDepth;
                open = true;
            } else if (cmd.kind == PFF_CMD_LINES) {
                // The run's lines are applied in order.
                if (!fr.lines || cmd.first > cmd.last || cmd.last > fr.numLines) PLAN_FAIL(PFD_E_INVALID, "bad lines command");
                DLnOp d;
                memset(&d, 0, sizeof d);
                d.color = D.color; d.depth = D.depth; d.target = D.kind; d.W = D.W; d.H = D.H;
                d.tilesX = (D.W + 15) / 16; d.tilesY = (D.H + 15) / 16;
                d.first = lnBase + cmd.first; d.last = lnBase + cmd.last;
                own_pass().lns.push_back(d);
                flush_seg();
            } else if (cmd.kind == PFF_CMD_POINTS) {
                // The run's points are applied in order.
                if (!fr.points || cmd.first > cmd.last || cmd.last > fr.numPoints) PLAN_FAIL(PFD_E_INVALID, "bad points command");
                DPtOp d;
                memset(&d, 0, sizeof d);
                d.color = D.color; d.depth = D.depth; d.target = D.kind; d.W = D.W; d.H = D.H;
                d.tilesX = (D.W + 15) / 16; d.tilesY = (D.H + 15) / 16;
                d.first = ptBase + cmd.first; d.last = ptBase + cmd.last;
                own_pass().pts.push_back(d);
                flush_seg();
            } else if (cmd.kind == PFF_CMD_FOG) {
                if (!fr.pixelOps || cmd.first >= fr.numPixelOps) PLAN_FAIL(PFD_E_INVALID, "bad fog command");
                const PFFpixels& op = fr.pixelOps[cmd.first];
                DPixOp d;
                memset(&d, 0, sizeof d);
                d.kind = 1u;
                d.color = D.color; d.depth = D.depth; d.target = D.kind; d.W = D.W; d.H = D.H;
                d.fogMode = op.fogMode; d.fogColor = op.fogColor;
                d.fogDensity = op.fogDensity; d.fogStart = op.fogStart; d.fogEnd = op.fogEnd;
                d.xMin = 0; d.yMin = 0; d.xMax = (int32_t)D.W - 1; d.yMax = (int32_t)D.H - 1;
                PassInfo& P = own_pass();
                P.pix.push_back(d);
                cur.pixOp = (uint32_t)P.pix.size();
                flush_seg();
            } else if (cmd.kind == PFF_CMD_PIXELS) {
                // No triangles: the call runs after the pass's raster
                // kernel, so it lands between the draws recorded before and
                // after it.  Validated in the size pass above.
                const PFFpixels& op = fr.pixelOps[cmd.first];
                const uint32_t bpp = pff_source_texel_bytes(op.source);
                const size_t texels = (size_t)op.width * op.height;
                uint32_t* img = (uint32_t*)((uint8_t*)J->pixPool.p + pixOff);
                pixOff += align256(texels * 4);
                e = hipMemcpy(J->stage.p, op.pixels, texels * bpp, hipMemcpyHostToDevice);
                if (e == hipSuccess) {
                    hipLaunchKernelGGL(k_decode_image, dim3((uint32_t)((texels + 255) / 256)), dim3(256), 0, J->stream,
                                       (const uint8_t*)J->stage.p, texels, op.source, img);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipStreamSynchronize(J->stream);   // the stage is reused
                if (e != hipSuccess) { set_error("pixel image decode", e); return PFD_E_LAUNCH; }
                DPixOp d;
                memset(&d, 0, sizeof d);
                d.texels = img;
                d.color = D.color; d.depth = D.depth; d.target = D.kind; d.W = D.W; d.H = D.H;
                d.w = op.width; d.h = op.height;
                d.xScreen = op.xScreen; d.yScreen = op.yScreen;
                d.xMin = op.xMin; d.yMin = op.yMin; d.xMax = op.xMax; d.yMax = op.yMax;
                d.zPos = op.zPos; d.invXLen = op.invXLen; d.invYLen = op.invYLen;
                d.flags = op.flags; d.depthFunc = op.depthFunc; d.blendMode = op.blendMode;
                PassInfo& P = own_pass();
                P.pix.push_back(d);
                cur.pixOp = (uint32_t)P.pix.size();
                flush_seg();
            } else if (cmd.kind == PFF_CMD_DRAW) {
                if (cmd.first > cmd.last || cmd.last > fr.tris.count) PLAN_FAIL(PFD_E_INVALID, "bad draw range");
                if (hasDraw && cur.triLast != triBase + cmd.first) flush_seg();
                if (!hasDraw) { cur.triFirst = triBase + cmd.first; cur.triLast = triBase + cmd.first; }
                cur.triLast = triBase + cmd.last;
                hasDraw = true;
                open = true;
            }
        }
        flush_seg();
        triBase += (uint32_t)nt;
        stateBase += fr.numStates;
        ptBase += fr.numPoints;
        lnBase += fr.numLines;
    }

    // Device copies of the inputs.
    auto up = [&](DevBuf& b, const void* h, size_t bytes) -> int {
        int r = ensure(J, b, bytes ? bytes : 4);
        if (r) return r;
        if (bytes) {
            hipError_t ee = hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice);
            if (ee != hipSuccess) { set_error("upload", ee); return PFD_E_LAUNCH; }
        }
        return 0;
    };
    const bool vidOn = wantVid && count == 1 && frames[0].tris.vertexId != nullptr;
    if (count == 1) {
        // One frame: the host-recorded ranges of its streams go up as they
        // are (state indices need no offset and were checked above).
        const PFFframe& fr = frames[0];
        const PFFtriangles& T = fr.tris;
        const size_t n = T.count;
        if ((rc = ensure(J, J->pos, n * 48 + 4)) || (rc = ensure(J, J->col, n * 12 + 4)) || (rc = ensure(J, J->info, n * 4 + 4))
            || (rc = ensure(J, J->uv, T.texcoord ? n * 24 + 4 : 4)) || (rc = ensure(J, J->nrm, T.normal ? n * 36 + 4 : 4))
            || (vidOn && (rc = ensure(J, J->vid, n * 12 + 4))))
            return rc;
        // Asynchronous on the job's stream: the host library records into
        // page-locked memory, so these are DMA transfers queued back to back
        // (no round trip each); the stream is synchronised at the end of
        // this function, before the caller may touch the streams again.
        for_host_ranges(fr, [&](uint32_t b, uint32_t e2) {
            auto part = [&](DevBuf& d, const void* h, size_t per) {
                if (e == hipSuccess && h)
                    e = hipMemcpyAsync((uint8_t*)d.p + b * per, (const uint8_t*)h + b * per, (size_t)(e2 - b) * per,
                                       hipMemcpyHostToDevice, J->stream);
            };
            part(J->pos, T.position, 48);
            part(J->col, T.color, 12);
            part(J->info, T.info, 4);
            part(J->uv, T.texcoord, 24);
            part(J->nrm, T.normal, 36);
            if (vidOn) part(J->vid, T.vertexId, 12);
        });
        if (e != hipSuccess) { set_error("upload", e); return PFD_E_LAUNCH; }
    } else {
        std::vector<float> hpos((size_t)totalTris * 12);
        std::vector<uint32_t> hcol((size_t)totalTris * 3), hinfo(totalTris);
        std::vector<float> huv(anyUV ? (size_t)totalTris * 6 : 0), hnrm(anyNrm ? (size_t)totalTris * 9 : 0);
        uint32_t tb = 0, sb = 0;
        for (uint32_t f = 0; f < count; f++) {
            const PFFframe& fr = frames[f];
            for_host_ranges(fr, [&](uint32_t b, uint32_t e2) {
                const size_t nt = e2 - b, o = (size_t)tb + b;
                memcpy(&hpos[o * 12], fr.tris.position + (size_t)b * 12, nt * 12 * sizeof(float));
                memcpy(&hcol[o * 3], fr.tris.color + (size_t)b * 3, nt * 3 * sizeof(uint32_t));
                if (anyUV && fr.tris.texcoord) memcpy(&huv[o * 6], fr.tris.texcoord + (size_t)b * 6, nt * 6 * sizeof(float));
                if (anyNrm && fr.tris.normal) memcpy(&hnrm[o * 9], fr.tris.normal + (size_t)b * 9, nt * 9 * sizeof(float));
                for (uint32_t t = b; t < e2; t++)
                    hinfo[tb + t] = (fr.tris.info[t] & ~PFF_STATE_MASK) | ((fr.tris.info[t] & PFF_STATE_MASK) + sb);
            });
            tb += fr.tris.count;
            sb += fr.numStates;
        }
        if ((rc = up(J->pos, hpos.data(), hpos.size() * 4)) || (rc = up(J->col, hcol.data(), hcol.size() * 4))
            || (rc = up(J->info, hinfo.data(), hinfo.size() * 4)) || (rc = up(J->uv, huv.data(), huv.size() * 4))
            || (rc = up(J->nrm, hnrm.data(), hnrm.size() * 4)))
            return rc;
    }
    if (numDraws) {
        // Array draws: element copies into one pool, then their triangles.
        if ((rc = ensure(J, J->arrPool, arrBytes)) || (rc = ensure(J, J->arrTable, (size_t)numDraws * sizeof(DArrays))))
            return rc;
        std::vector<DArrays> tab;
        tab.reserve(numDraws);
        bool usedResident = false;
        uint8_t* dst = (uint8_t*)J->arrPool.p;
        uint32_t tb = 0, sb = 0, cum = 0;
        auto put = [&](const void* h, size_t bytes) -> const uint8_t* {
            if (!h) return nullptr;
            uint8_t* d = dst;
            if (bytes && e == hipSuccess) e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
            dst += align256(bytes);
            return d;
        };
        for (uint32_t f = 0; f < count; f++) {
            const PFFframe& fr = frames[f];
            for (uint32_t k = 0; k < fr.numArrays; k++) {
                const PFFarrays& a = fr.arrays[k];
                const size_t n = a.numElements;
                DArrays d;
                memset(&d, 0, sizeof d);
                if (a.device) {
                    // Resident copy (a render list's vertices): same layout as on the host.
                    const uint8_t* dev = (const uint8_t*)a.device;
                    const uint8_t* h0 = (const uint8_t*)a.position;
                    auto at = [&](const void* h) -> const uint8_t* {
                        return h ? dev + ((const uint8_t*)h - h0) : nullptr;
                    };
                    d.pos = dev;
                    d.nrm = at(a.normal);
                    d.uv = at(a.texcoord);
                    d.col = at(a.colors);
                    usedResident = true;
                } else {
                    d.pos = put(a.position, n * a.posSize * elem_bytes(a.posType));
                    d.nrm = put(a.normal, n * 3 * elem_bytes(a.nrmType));
                    d.uv = put(a.texcoord, n * 2 * elem_bytes(a.uvType));
                    d.col = put(a.colors, n * a.colSize * elem_bytes(a.colType));
                    d.idx = put(a.indices, (size_t)a.count * elem_bytes(a.idxType));
                }
                d.flags = a.flags;
                memcpy(d.uvm, a.uvMatrix, sizeof d.uvm);
                d.posSize = (uint32_t)a.posSize; d.posType = a.posType; d.nrmType = a.nrmType; d.uvType = a.uvType;
                d.colSize = (uint32_t)a.colSize; d.colType = a.colType; d.idxType = a.idxType;
                d.mode = a.mode;
                d.perBuf = a.mode == 2u ? 1u : (a.mode == 6u || a.mode == 7u) ? 4u : 2u;
                d.stride = (a.mode == 2u || a.mode == 3u || a.mode == 4u) ? 3u : 4u;
                d.numFaces = a.numFaces;
                d.facePass0 = a.facePasses[0];
                d.facePass1 = a.facePasses[1];
                d.state = a.state + sb;
                d.color = a.color;
                d.elementBase = a.elementBase;
                d.numElements = a.numElements;
                d.slot0 = tb + a.triFirst;
                d.cum = cum;
                cum += a.triCount;
                tab.push_back(d);
            }
            tb += fr.tris.count;
            sb += fr.numStates;
        }
        if (e == hipSuccess) e = hipMemcpy(J->arrTable.p, tab.data(), tab.size() * sizeof(DArrays), hipMemcpyHostToDevice);
        if (e != hipSuccess) { set_error("array draw upload", e); return PFD_E_LAUNCH; }
        if (arrTris) {
            hipLaunchKernelGGL(k_deindex, dim3((uint32_t)((arrTris + 255) / 256)), dim3(256), 0, J->stream,
                               (const DArrays*)J->arrTable.p, numDraws, (uint32_t)arrTris, (float4*)J->pos.p,
                               (uint32_t*)J->col.p, anyUV ? (float2*)J->uv.p : nullptr,
                               anyNrm ? (float*)J->nrm.p : nullptr, (uint32_t*)J->info.p,
                               vidOn ? (uint32_t*)J->vid.p : nullptr);
            HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
        }
        // Resident copies belong to their lists, which may go once the
        // frame has been handed over: read them before returning.
        if (usedResident) HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    }
    if ((rc = up(J->states, hstates.data(), hstates.size() * sizeof(PFFstate)))
        || (rc = up(J->texTable, htex.data(), htex.size() * sizeof(DTex))))
        return rc;
    if (!hpts.empty()) {
        // Points are projected once; every replay reads the projections.
        if ((rc = up(J->points, hpts.data(), hpts.size() * sizeof(PFFpoint)))
            || (rc = ensure(J, J->pointScr, hpts.size() * sizeof(DPointScr))))
            return rc;
        const uint32_t np = (uint32_t)hpts.size();
        hipLaunchKernelGGL(k_point_setup, dim3((np + 255) / 256), dim3(256), 0, J->stream, (const PFFpoint*)J->points.p,
                           np, (const PFFstate*)J->states.p, (DPointScr*)J->pointScr.p);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
        if (e != hipSuccess) { set_error("point setup", e); return PFD_E_LAUNCH; }
    }
    if (!hlns.empty()) {
        // Lines are projected and clipped once; every replay reads them.
        if ((rc = up(J->lines, hlns.data(), hlns.size() * sizeof(PFFline)))
            || (rc = ensure(J, J->lineScr, hlns.size() * sizeof(DLineScr))))
            return rc;
        const uint32_t nl = (uint32_t)hlns.size();
        hipLaunchKernelGGL(k_line_setup, dim3((nl + 255) / 256), dim3(256), 0, J->stream, (const PFFline*)J->lines.p,
                           nl, (const PFFstate*)J->states.p, (DLineScr*)J->lineScr.p);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
        if (e != hipSuccess) { set_error("line setup", e); return PFD_E_LAUNCH; }
    }
    // Per-pass tables, all passes in one buffer (one copy).
    std::vector<size_t> offs;
    size_t tabBytes = 0;
    auto place = [&](size_t bytes) { offs.push_back(tabBytes); tabBytes += align256(bytes); };
    for (auto& P : J->passes) {
        uint32_t acc = 0;
        P.frameSeg.assign(count, -1);
        for (size_t s = 0; s < P.segs.size(); s++) {
            P.segs[s].passTriBase = acc;
            acc += P.segs[s].triLast - P.segs[s].triFirst;
            P.frameSeg[P.segs[s].frame] = (int32_t)s;
        }
        P.numPassTris = acc;
        P.blockSeg.assign((acc + 255) / 256 + 1, 0u);
        for (size_t b = 0, sgi = 0; b < P.blockSeg.size(); b++) {
            while (sgi + 1 < P.segs.size() && P.segs[sgi + 1].passTriBase <= b * 256) sgi++;
            P.blockSeg[b] = (uint32_t)sgi;
        }
        if (!P.pix.empty()) {
            uint64_t maxArea = 1;
            for (const DPixOp& d : P.pix) {
                if (d.xMax < d.xMin || d.yMax < d.yMin) continue;
                const uint64_t a = (uint64_t)(uint32_t)(d.xMax - d.xMin + 1) * (uint64_t)(uint32_t)(d.yMax - d.yMin + 1);
                if (a > maxArea) maxArea = a;
            }
            const uint64_t g = (maxArea + 255) / 256;
            P.pixGrid = (uint32_t)(g < 4096 ? g : 4096);
        }
        P.ptTiles = 1;
        for (const DPtOp& d : P.pts) P.ptTiles = std::max(P.ptTiles, d.tilesX * d.tilesY);
        P.lnTiles = 1;
        for (const DLnOp& d : P.lns) P.lnTiles = std::max(P.lnTiles, d.tilesX * d.tilesY);
        place(P.segs.size() * sizeof(DSeg));
        place(P.blockSeg.size() * 4);
        place(P.pix.size() * sizeof(DPixOp));
        place(P.pts.size() * sizeof(DPtOp));
        place(P.lns.size() * sizeof(DLnOp));
    }
    if (tabBytes) {
        std::vector<uint8_t> tab(tabBytes, 0);
        size_t k = 0;
        for (auto& P : J->passes) {
            memcpy(&tab[offs[k + 0]], P.segs.data(), P.segs.size() * sizeof(DSeg));
            memcpy(&tab[offs[k + 1]], P.blockSeg.data(), P.blockSeg.size() * 4);
            if (!P.pix.empty()) memcpy(&tab[offs[k + 2]], P.pix.data(), P.pix.size() * sizeof(DPixOp));
            if (!P.pts.empty()) memcpy(&tab[offs[k + 3]], P.pts.data(), P.pts.size() * sizeof(DPtOp));
            if (!P.lns.empty()) memcpy(&tab[offs[k + 4]], P.lns.data(), P.lns.size() * sizeof(DLnOp));
            k += 5;
        }
        if ((rc = up(J->passTab, tab.data(), tabBytes))) return rc;
        const uint8_t* base = (const uint8_t*)J->passTab.p;
        k = 0;
        for (auto& P : J->passes) {
            P.dSegs = (const DSeg*)(base + offs[k + 0]);
            P.dBlockSeg = (const uint32_t*)(base + offs[k + 1]);
            P.dPix = (const DPixOp*)(base + offs[k + 2]);
            P.dPts = (const DPtOp*)(base + offs[k + 3]);
            P.dLns = (const DLnOp*)(base + offs[k + 4]);
            k += 5;
        }
    }
    layout_tiles(J);
    if ((rc = upload_frames_table(J))) return rc;
    // The caller may change its streams once this returns: the asynchronous
    // uploads above must have read them.
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

#undef PLAN_FAIL

}  // namespace

extern "C" {

int pfdJobCreate(const PFFframe* frames, uint32_t count, const PFDconfig* cfg, PFDjob** out) {
    if (!out || !frames || count == 0) { snprintf(g_err, sizeof g_err, "pfdJobCreate: bad arguments"); return PFD_E_INVALID; }
    *out = nullptr;
    int rc = pfdInit();
    if (rc) return rc;
    PFDjob* J = nullptr;
    if ((rc = job_shell(&J, true))) return rc;
    J->cfg = cfg ? *cfg : g_defaultCfg;
    J->tileShift = (J->cfg.tileShift >= 3 && J->cfg.tileShift <= 6) ? J->cfg.tileShift : 3;
    if ((rc = job_plan(J, frames, count, false))) { pfdJobDestroy(J); return rc; }
    *out = J;
    return PFD_OK;
}

int pfdJobSetConfig(PFDjob* J, const PFDconfig* cfg) {
    if (!J || !cfg) return PFD_E_INVALID;
    if (cfg->rasterMode != J->cfg.rasterMode) drop_graph(J);
    J->cfg = *cfg;
    uint32_t ts = (cfg->tileShift >= 3 && cfg->tileShift <= 6) ? cfg->tileShift : 3;
    if (ts != J->tileShift) {
        J->tileShift = ts;
        layout_tiles(J);
        int rc = upload_frames_table(J);
        if (rc) return rc;
        for (auto& P : J->passes) P.sized = false;
        drop_graph(J);
    }
    return PFD_OK;
}

// Host render targets -> device.  A frame whose first command is a clear
// needs only pixels 0..7 (all other pixels are overwritten or copy pixel 0).
int pfdJobUploadTargets(PFDjob* J) {
    if (!J) return PFD_E_INVALID;
    for (size_t f = 0; f < J->frames.size(); f++) {
        const PFFframe& fr = J->src[f];
        size_t npx = (size_t)fr.width * fr.height;
        bool leadClear = J->leadClear[f] != 0;
        size_t n = leadClear ? (npx < 8 ? npx : 8) : npx;
        HIP_TRY(hipMemcpyAsync(J->frames[f].depth, fr.depth, n * 4, hipMemcpyHostToDevice, J->stream), PFD_E_LAUNCH);
        HIP_TRY(staged_h2d(J, J->frames[f].color, fr.color, n * pff_target_bytes(fr.target)), PFD_E_LAUNCH);
    }
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

int pfdJobLaunch(PFDjob* J) {
    if (!J) return PFD_E_INVALID;
    bool allSized = true;
    for (auto& P : J->passes) allSized = allSized && P.sized;
    if (!allSized) return run_all(J, true);
    if (!J->graphExec) {
        HIP_TRY(hipStreamBeginCapture(J->stream, hipStreamCaptureModeThreadLocal), PFD_E_LAUNCH);
        int rc = run_all(J, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(J->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess)
            for (auto& P : J->passes) P.sized = false;
        if (e != hipSuccess) { set_error("hipStreamEndCapture", e); return PFD_E_LAUNCH; }
        J->graph = g;
        e = hipGraphInstantiate(&J->graphExec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { set_error("hipGraphInstantiate", e); drop_graph(J); return PFD_E_LAUNCH; }
    }
    HIP_TRY(hipGraphLaunch(J->graphExec, J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

int pfdJobSync(PFDjob* J) {
    if (!J) return PFD_E_INVALID;
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

int pfdJobDownload(PFDjob* J) {
    if (!J) return PFD_E_INVALID;
    for (size_t f = 0; f < J->frames.size(); f++) {
        const PFFframe& fr = J->src[f];
        size_t npx = (size_t)fr.width * fr.height;
        // Depth is the library's page-locked buffer: queued first, it
        // overlaps the staged copy of an application's pageable colour.
        HIP_TRY(hipMemcpyAsync(fr.depth, J->frames[f].depth, npx * 4, hipMemcpyDeviceToHost, J->stream), PFD_E_LAUNCH);
        HIP_TRY(staged_d2h(J, fr.color, J->frames[f].color, npx * pff_target_bytes(fr.target)), PFD_E_LAUNCH);
    }
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

int pfdJobRunTimed(PFDjob* J, PFDtimes* t) {
    if (!J || !t) return PFD_E_INVALID;
    memset(t, 0, sizeof *t);
    hipEvent_t ev[4];
    for (int i = 0; i < 4; i++) HIP_TRY(hipEventCreate(&ev[i]), PFD_E_LAUNCH);
    int rc = 0;
    for (auto& P : J->passes) {
        rc = run_pass(J, P, !P.sized, ev);
        if (rc) break;
        HIP_TRY(hipEventSynchronize(ev[3]), PFD_E_LAUNCH);
        float a = 0, b = 0, c = 0;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        (void)hipEventElapsedTime(&c, ev[2], ev[3]);
        t->setupMs += a; t->binMs += b; t->rasterMs += c;
    }
    for (int i = 0; i < 4; i++) (void)hipEventDestroy(ev[i]);
    t->totalMs = t->setupMs + t->binMs + t->rasterMs;
    return rc;
}

int pfdJobTimeReplays(PFDjob* J, uint32_t iters, float* msTotal) {
    if (!J || !msTotal) return PFD_E_INVALID;
    int rc = pfdJobLaunch(J);              // sizes (first call) or warm graph
    if (rc) return rc;
    if ((rc = pfdJobLaunch(J))) return rc; // instantiate the graph if needed
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), PFD_E_LAUNCH);
    HIP_TRY(hipEventCreate(&b), PFD_E_LAUNCH);
    HIP_TRY(hipEventRecord(a, J->stream), PFD_E_LAUNCH);
    for (uint32_t i = 0; i < iters; i++) HIP_TRY(hipGraphLaunch(J->graphExec, J->stream), PFD_E_LAUNCH);
    HIP_TRY(hipEventRecord(b, J->stream), PFD_E_LAUNCH);
    HIP_TRY(hipEventSynchronize(b), PFD_E_LAUNCH);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *msTotal = ms;
    return PFD_OK;
}

int pfdJobTarget(PFDjob* J, uint32_t frame, void** color, float** depth, uint32_t* width, uint32_t* height) {
    if (!J || frame >= J->src.size()) return PFD_E_INVALID;
    const PFFframe& fr = J->src[frame];
    if (color) *color = fr.color;
    if (depth) *depth = fr.depth;
    if (width) *width = fr.width;
    if (height) *height = fr.height;
    return PFD_OK;
}

int pfdJobFrameCount(PFDjob* J) { return J ? (int)J->src.size() : PFD_E_INVALID; }

int pfdJobStats(PFDjob* J, PFDstats* st) {
    if (!J || !st) return PFD_E_INVALID;
    memset(st, 0, sizeof *st);
    unsigned long long frags = 0;
    if (J->tileFrags.p) {
        const size_t n = (size_t)J->numTiles * 4 * J->passes.size();
        std::vector<uint32_t> tf(n);
        HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
        HIP_TRY(hipMemcpy(tf.data(), J->tileFrags.p, n * 4, hipMemcpyDeviceToHost), PFD_E_LAUNCH);
        for (uint32_t v : tf) frags += v;
    }
    st->fragments = frags;
    for (auto& P : J->passes) {
        st->rasterTriangles += P.numRT;
        st->tilePairs += P.numPairs;
        st->clippedTriangles += P.numClip;
    }
    st->inputTriangles = J->numIn;
    st->numTiles = J->numTiles;
    st->tileShift = J->tileShift;
    size_t bytes = 0;
    for_each_buf(J, [&](DevBuf& b) { bytes += b.bytes; });
    st->deviceBytes = bytes;
    if (J->tileBegin.p && J->numTiles) {
        std::vector<uint32_t> b(J->numTiles), e(J->numTiles);
        HIP_TRY(hipMemcpy(b.data(), J->tileBegin.p, (size_t)J->numTiles * 4, hipMemcpyDeviceToHost), PFD_E_LAUNCH);
        HIP_TRY(hipMemcpy(e.data(), J->tileEnd.p, (size_t)J->numTiles * 4, hipMemcpyDeviceToHost), PFD_E_LAUNCH);
        for (uint32_t i = 0; i < J->numTiles; i++) {
            uint32_t n = e[i] - b[i];
            if (n) st->busyTiles++;
            if (n > st->maxTileList) st->maxTileList = n;
        }
    }
    return PFD_OK;
}

// Times every tile size and raster mode on the job (one sizing run plus a
// few runs each) and keeps the fastest.  The target buffers are restored
// from the host afterwards so the job's output is unchanged.  Candidates are
// timed as runs launched straight onto the stream, not as graphs: a graph
// instantiated and destroyed per candidate (14 per job) bought nothing in
// the ranking, and the one graph of the chosen configuration is built by
// the next pfdJobLaunch.
int pfdJobAutotune(PFDjob* J, PFDconfig* chosen) {
    if (!J) return PFD_E_INVALID;
    float best = 1e30f;
    PFDconfig bestCfg = J->cfg;
    for (uint32_t mode = 0; mode <= 3; mode++)
        for (uint32_t ts = (mode >= 2 ? 4 : 3); ts <= 6; ts++) {   // four pixels per lane: 16x16 tiles
            PFDconfig c = J->cfg;
            c.tileShift = ts;
            c.rasterMode = mode;
            int rc = pfdJobSetConfig(J, &c);
            if (rc) return rc;
            if ((rc = pfdJobUploadTargets(J))) return rc;
            // Replays worth about 10 ms, best of three trials: single short
            // timings mis-ranked configurations a few percent apart.
            float est = 0;
            if ((rc = time_direct(J, 3, &est))) return rc;
            const float per = est / 3.0f;
            const uint32_t iters = per > 0.0f ? (uint32_t)fminf(50.0f, fmaxf(3.0f, 10.0f / per)) : 3u;
            float ms = 1e30f;
            for (int trial = 0; trial < 3; trial++) {
                float t = 0;
                if ((rc = time_direct(J, iters, &t))) return rc;
                ms = fminf(ms, t / (float)iters);
            }
            if (ms < best) { best = ms; bestCfg = c; }
        }
    PFDconfig c = bestCfg;
    int rc = pfdJobSetConfig(J, &c);
    if (rc) return rc;
    if ((rc = pfdJobUploadTargets(J))) return rc;
    if (chosen) *chosen = J->cfg;
    return PFD_OK;
}

// Host buffers of the host library.  A 64-byte header in front of the
// caller's bytes records how they were allocated, so pfdHostFree needs no
// table and the caller's pointer keeps 64-byte alignment.
namespace {
constexpr uint64_t HOST_MAGIC = 0x50464448u;   // "PFDH"
struct HostHdr {
    uint64_t magic;
    uint64_t pinned;
    uint64_t pad[6];
};
static_assert(sizeof(HostHdr) == 64, "host header");
}  // namespace

void* pfdHostAlloc(size_t bytes) {
    void* base = nullptr;
    uint64_t pinned = 0;
    if (pfdInit() == PFD_OK && hipHostMalloc(&base, bytes + sizeof(HostHdr), hipHostMallocDefault) == hipSuccess)
        pinned = 1;
    else if (!(base = malloc(bytes + sizeof(HostHdr))))
        return nullptr;
    HostHdr* h = static_cast<HostHdr*>(base);
    h->magic = HOST_MAGIC;
    h->pinned = pinned;
    return h + 1;
}

void pfdHostFree(void* p) {
    if (!p) return;
    HostHdr* h = static_cast<HostHdr*>(p) - 1;
    if (h->magic != HOST_MAGIC) return;            // not ours: left alone rather than corrupting the heap
    h->magic = 0;
    if (h->pinned) (void)hipHostFree(h);
    else free(h);
}

struct PFDresident {
    void*  p;
    size_t bytes;
};

int pfdResidentCreate(const void* host, size_t bytes, PFDresident** out) {
    if (!out || (!host && bytes)) return PFD_E_INVALID;
    *out = nullptr;
    int rc = pfdInit();
    if (rc) return rc;
    PFDresident* r = new PFDresident();
    r->bytes = bytes;
    hipError_t e = hipMalloc(&r->p, bytes ? bytes : 16);
    if (e != hipSuccess) { set_error("hipMalloc", e); delete r; return PFD_E_NOMEM; }
    if (bytes && (e = hipMemcpy(r->p, host, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
        set_error("resident upload", e);
        (void)hipFree(r->p);
        delete r;
        return PFD_E_LAUNCH;
    }
    *out = r;
    return PFD_OK;
}

const void* pfdResidentPointer(const PFDresident* r) { return r ? r->p : nullptr; }

void pfdResidentDestroy(PFDresident* r) {
    if (!r) return;
    (void)hipFree(r->p);
    delete r;
}

void pfdJobDestroy(PFDjob* J) {
    if (!J) return;
    if (J->stream) (void)hipStreamSynchronize(J->stream);
    if (J->side) (void)hipStreamSynchronize(J->side);
    drop_graph(J);
    for_each_buf(J, [](DevBuf& b) { release(b); });
    for (int i = 0; i < 2; i++) {
        if (J->hstage[i]) (void)hipHostFree(J->hstage[i]);
        if (J->hstageEv[i]) (void)hipEventDestroy(J->hstageEv[i]);
    }
    if (J->fork) (void)hipEventDestroy(J->fork);
    if (J->join) (void)hipEventDestroy(J->join);
    if (J->side) (void)hipStreamDestroy(J->side);
    if (J->stream) (void)hipStreamDestroy(J->stream);
    delete J;
}

}  // extern "C"

namespace {

// Plans and runs frames once on a reset job: upload, one discovering run
// (no graph), download.
// Shape of a set of frames for reusing an automatic configuration: a frame
// of the same targets, triangle and command counts as the last one is
// taken to want the same tiles (a performance choice only).
uint64_t frame_shape(const PFFframe* frames, uint32_t count) {
    uint64_t h = 1469598103934665603ull ^ count;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    for (uint32_t f = 0; f < count; f++) {
        mix(frames[f].width); mix(frames[f].height); mix(frames[f].tris.count);
        mix(frames[f].numCmds); mix(frames[f].numArrays); mix(frames[f].target);
    }
    return h | 1u;
}

int run_one_shot(PFDjob* J, const PFFframe* frames, uint32_t count) {
    J->cfg = g_defaultCfg;
    J->tileShift = (J->cfg.tileShift >= 3 && J->cfg.tileShift <= 6) ? J->cfg.tileShift : 3;
    J->autoTiles = J->cfg.tileShift == 0;
    const uint64_t shape = J->autoTiles ? frame_shape(frames, count) : 0;
    if (J->autoTiles && shape == J->autoKey) {
        // The previous frame had this shape: its choice, without counting
        // the first pass twice.
        J->autoTiles = false;
        J->tileShift = J->autoShift;
        J->cfg.rasterMode = J->autoMode;
    }
    const bool choosing = J->autoTiles;
    int rc = job_plan(J, frames, count, false);
    if (!rc && !(rc = pfdJobUploadTargets(J)) && !(rc = run_all(J, true)) && !(rc = pfdJobSync(J)))
        rc = pfdJobDownload(J);
    if (!rc && choosing) {
        J->autoKey = shape;
        J->autoShift = J->tileShift;
        J->autoMode = J->cfg.rasterMode;
    }
    return rc;
}

// The job of pfdExecuteFrame(s), kept between calls: its streams, device
// buffers (which only grow) and host tables serve the next frame, so a
// frame of the same size as an earlier one costs no hipMalloc, hipFree or
// stream creation.  A call made while another thread runs one uses a job
// of its own.
std::mutex g_oneShotLock;
PFDjob*    g_oneShot = nullptr;

}  // namespace

extern "C" {

int pfdExecuteFrames(const PFFframe* frames, uint32_t count) {
    if (!frames || count == 0) { snprintf(g_err, sizeof g_err, "pfdExecuteFrames: bad arguments"); return PFD_E_INVALID; }
    int rc = pfdInit();
    if (rc) return rc;
    std::unique_lock<std::mutex> lk(g_oneShotLock, std::try_to_lock);
    if (lk.owns_lock()) {
        if (!g_oneShot && (rc = job_shell(&g_oneShot, true))) return rc;
        job_reset(g_oneShot);
        return run_one_shot(g_oneShot, frames, count);
    }
    PFDjob* J = nullptr;
    if ((rc = job_shell(&J, true))) return rc;
    rc = run_one_shot(J, frames, count);
    pfdJobDestroy(J);
    return rc;
}

int pfdExecuteFrame(const PFFframe* frame) { return pfdExecuteFrames(frame, 1); }

int pfdExecuteFrameNoAlloc(const PFFframe* frame) {
    if (!frame) { snprintf(g_err, sizeof g_err, "pfdExecuteFrameNoAlloc: no frame"); return PFD_E_INVALID; }
    int rc = pfdInit();
    if (rc) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), PFD_E_NODEVICE);
    if (dev < 0 || dev >= PFD_MAX_DEVICES) {
        snprintf(g_err, sizeof g_err, "pfdExecuteFrameNoAlloc: device %d out of range", dev);
        return PFD_E_NODEVICE;
    }
    std::lock_guard<std::mutex> lk(g_noAllocLock);
    if ((rc = no_alloc_setup(dev))) return rc;
    PFDjob* J = g_noAllocJob[dev];
    job_reset(J);                       // also waits for the previous call's work on the stream
    rc = run_one_shot(J, frame, 1);
    if (rc) (void)hipStreamSynchronize(J->stream);
    return rc;
}

static PFDjob* g_captured = nullptr;

int pfdCaptureExecutor(const PFFframe* frame, void* user) {
    (void)user;
    PFDjob* J = nullptr;
    int rc = pfdJobCreate(frame, 1, nullptr, &J);
    if (rc) return rc;
    if (g_captured) pfdJobDestroy(g_captured);
    g_captured = J;
    return PFD_OK;
}

int pfdCaptureBatchExecutor(const PFFframe* frames, uint32_t count, void* user) {
    (void)user;
    PFDjob* J = nullptr;
    int rc = pfdJobCreate(frames, count, nullptr, &J);
    if (rc) return rc;
    if (g_captured) pfdJobDestroy(g_captured);
    g_captured = J;
    return PFD_OK;
}

PFDjob* pfdTakeCapturedJob(void) {
    PFDjob* J = g_captured;
    g_captured = nullptr;
    return J;
}

// ---------------------------------------------------------------------------
// Differentiable jobs
// ---------------------------------------------------------------------------

int pfdDiffCreate(const PFFframe* frame, PFDjob** out) {
    if (!frame || !out) return PFD_E_INVALID;
    *out = nullptr;
    if (frame->target != PFF_TARGET_RGBA8) {
        snprintf(g_err, sizeof g_err, "differentiable frame: the render target must be RGBA / UNSIGNED_BYTE");
        return PFD_E_INVALID;
    }
    // One pass: at most one clear, before every draw.
    for (uint32_t c = 0; c < frame->numCmds; c++)
        if ((frame->cmds[c].kind == PFF_CMD_CLEAR && c != 0) || frame->cmds[c].kind == PFF_CMD_PIXELS
            || frame->cmds[c].kind == PFF_CMD_FOG || frame->cmds[c].kind == PFF_CMD_POINTS
            || frame->cmds[c].kind == PFF_CMD_LINES) {
            snprintf(g_err, sizeof g_err, "differentiable frame: %s",
                     frame->cmds[c].kind == PFF_CMD_CLEAR ? "a clear after the first command"
                                                          : "points, lines, pfDrawPixels and pfFogProcess are not differentiable here");
            return PFD_E_INVALID;
        }
    for (uint32_t i = 0; i < frame->numStates; i++) {
        const PFFstate& st = frame->states[i];
        const char* why = nullptr;
        if (st.flags & PFF_BLEND) why = "blending";
        else if (st.shadeModel != 1) why = "flat shading";
        else if ((st.flags & PFF_LIGHTING) && st.lightMode != 1) why = "Gouraud lighting";
        if (why) {
            snprintf(g_err, sizeof g_err, "differentiable frame: %s is not differentiable here", why);
            return PFD_E_INVALID;
        }
    }
    int rc = pfdInit();
    if (rc) return rc;
    PFDjob* J = nullptr;
    if ((rc = job_shell(&J, true))) return rc;
    J->cfg = g_defaultCfg;
    J->tileShift = (J->cfg.tileShift >= 3 && J->cfg.tileShift <= 6) ? J->cfg.tileShift : 3;
    // The corners' colour-array indices (host-recorded or formed from array
    // draws) address the colour parameters.
    if ((rc = job_plan(J, frame, 1, true))) { pfdJobDestroy(J); return rc; }
    J->diff = true;
    const size_t npx = (size_t)frame->width * frame->height;
    if ((rc = ensure(J, J->winner, npx * 4)) || (rc = ensure(J, J->image, npx * 16)) || (rc = ensure(J, J->dImage, npx * 16))) {
        pfdJobDestroy(J);
        return rc;
    }
    *out = J;
    return PFD_OK;
}

}  // extern "C"

namespace {

// Uploads the parameters (each call: they change between optimisation
// steps) and fills the kernel arguments.
int diff_args(PFDjob* J, const PFDdiffParams* p, DiffArgs& a) {
    if (!J || !J->diff || !p) { snprintf(g_err, sizeof g_err, "not a differentiable job"); return PFD_E_INVALID; }
    int rc;
    memset(&a, 0, sizeof a);
    a.paramTex = -1;
    if (p->texture >= 0) {
        if ((uint32_t)p->texture >= J->srcTex.size()) { snprintf(g_err, sizeof g_err, "bad parameter texture"); return PFD_E_INVALID; }
        const PFFtexture& T = J->srcTex[p->texture];
        const float* texels = p->texels;
        if (!texels) {
            if (T.format != PFF_TEX_RGBAF32) {
                snprintf(g_err, sizeof g_err, "parameter texture must be RGBA float or come with texels");
                return PFD_E_INVALID;
            }
            texels = (const float*)T.pixels;
        }
        const size_t bytes = (size_t)T.width * T.height * 16;
        if ((rc = ensure(J, J->texParams, bytes))) return rc;
        HIP_TRY(hipMemcpyAsync(J->texParams.p, texels, bytes, hipMemcpyHostToDevice, J->stream), PFD_E_LAUNCH);
        a.paramTex = p->texture;
        a.texParams = (const float*)J->texParams.p;
    }
    if (p->vertexColors && p->numVertices) {
        const size_t bytes = (size_t)p->numVertices * 16;
        if ((rc = ensure(J, J->colParams, bytes))) return rc;
        HIP_TRY(hipMemcpyAsync(J->colParams.p, p->vertexColors, bytes, hipMemcpyHostToDevice, J->stream), PFD_E_LAUNCH);
        a.colParams = (const float*)J->colParams.p;
        a.numVertices = p->numVertices;
    }
    J->numVertices = a.numVertices;
    J->paramTex = a.paramTex;
    if (p->lightDiffuse || p->lightSpecular) {
        // Recorded colours / 255 for whichever of the two is not given.
        float lp[DIFF_LIGHT_VALS];
        for (int id = 0; id < PFF_MAX_LIGHTS; id++)
            for (int c = 0; c < 3; c++) { lp[id * 3 + c] = 0.0f; lp[PFF_MAX_LIGHTS * 3 + id * 3 + c] = 0.0f; }
        for (const PFFstate& st : J->srcStates)
            for (uint32_t l = 0; l < st.numLights; l++) {
                const PFFlight& L = st.lights[l];
                for (int c = 0; c < 3; c++) {
                    lp[(L.id & 7u) * 3 + c] = (float)((L.diffuse >> (8 * c)) & 0xFFu) / 255.0f;
                    lp[PFF_MAX_LIGHTS * 3 + (L.id & 7u) * 3 + c] = (float)((L.specular >> (8 * c)) & 0xFFu) / 255.0f;
                }
            }
        if (p->lightDiffuse) memcpy(lp, p->lightDiffuse, PFF_MAX_LIGHTS * 3 * sizeof(float));
        if (p->lightSpecular) memcpy(lp + PFF_MAX_LIGHTS * 3, p->lightSpecular, PFF_MAX_LIGHTS * 3 * sizeof(float));
        if ((rc = ensure(J, J->lightParams, sizeof lp))) return rc;
        HIP_TRY(hipMemcpyAsync(J->lightParams.p, lp, sizeof lp, hipMemcpyHostToDevice, J->stream), PFD_E_LAUNCH);
        HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);   // lp is on this stack frame
        a.lightParams = (const float*)J->lightParams.p;
    }
    a.rts = (const RTri*)J->rts.p;
    a.rtl = (const RTriL*)J->rtl.p;
    a.rtd = (const RTriD*)J->rtd.p;
    a.winner = (const uint32_t*)J->winner.p;
    a.color = (const uint32_t*)J->frames[0].color;
    a.srcCol = (const uint32_t*)J->col.p;
    a.vid = J->vid.p ? (const uint32_t*)J->vid.p : nullptr;
    a.states = (const PFFstate*)J->states.p;
    a.texs = (const DTex*)J->texTable.p;
    a.W = J->frames[0].W;
    a.H = J->frames[0].H;
    a.image = (float*)J->image.p;
    a.dImage = (const float*)J->dImage.p;
    return PFD_OK;
}

dim3 diff_grid(const DiffArgs& a) { return dim3((a.W + 15) / 16, (a.H + 15) / 16); }

// The integer pass that fixes coverage and winners; runs once per job.
int diff_raster(PFDjob* J) {
    if (J->diffRan) return PFD_OK;
    int rc;
    if ((rc = pfdJobUploadTargets(J)) || (rc = run_all(J, true))) return rc;
    if ((rc = pfdJobDownload(J))) return rc;
    J->diffRan = true;
    return PFD_OK;
}

int diff_backward_launch(PFDjob* J, DiffArgs& a, bool wantTex, bool wantCol, bool wantLight) {
    int rc;
    const dim3 grid = diff_grid(a);
    const uint32_t blocks = grid.x * grid.y;
    if (wantTex && a.paramTex >= 0) {
        const PFFtexture& T = J->srcTex[a.paramTex];
        const size_t n = (size_t)T.width * T.height * 4;
        if ((rc = ensure(J, J->gTex, n * 4))) return rc;
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256 * 8)), dim3(256), 0, J->stream, (float*)J->gTex.p, n);
        a.gTex = (float*)J->gTex.p;
    }
    if (wantCol && a.colParams) {
        const size_t n = (size_t)a.numVertices * 4;
        if ((rc = ensure(J, J->gCol, n * 4))) return rc;
        hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256 * 8)), dim3(256), 0, J->stream, (float*)J->gCol.p, n);
        a.gCol = (float*)J->gCol.p;
    }
    if (wantLight) {
        if ((rc = ensure(J, J->gLightPart, (size_t)blocks * DIFF_LIGHT_VALS * 4)) || (rc = ensure(J, J->gLight, DIFF_LIGHT_VALS * 4)))
            return rc;
        a.gLightPart = (float*)J->gLightPart.p;
    }
    hipLaunchKernelGGL(k_diff<true>, grid, dim3(256), 0, J->stream, a);
    if (wantLight)
        hipLaunchKernelGGL(k_diff_light_sum, dim3(DIFF_LIGHT_VALS), dim3(256), 0, J->stream,
                           (const float*)J->gLightPart.p, blocks, (float*)J->gLight.p);
    HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
    return PFD_OK;
}

}  // namespace

extern "C" {

int pfdDiffForward(PFDjob* J, const PFDdiffParams* p, float* image) {
    DiffArgs a;
    int rc;
    if ((rc = diff_raster(J)) || (rc = diff_args(J, p, a))) return rc;
    hipLaunchKernelGGL(k_diff<false>, diff_grid(a), dim3(256), 0, J->stream, a);
    HIP_TRY(hipGetLastError(), PFD_E_LAUNCH);
    if (image)
        HIP_TRY(hipMemcpyAsync(image, J->image.p, (size_t)a.W * a.H * 16, hipMemcpyDeviceToHost, J->stream), PFD_E_LAUNCH);
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    return PFD_OK;
}

int pfdDiffBackward(PFDjob* J, const PFDdiffParams* p, const float* dImage, PFDdiffGrads* g) {
    if (!dImage || !g) return PFD_E_INVALID;
    DiffArgs a;
    int rc;
    if ((rc = diff_raster(J)) || (rc = diff_args(J, p, a))) return rc;
    HIP_TRY(hipMemcpyAsync(J->dImage.p, dImage, (size_t)a.W * a.H * 16, hipMemcpyHostToDevice, J->stream), PFD_E_LAUNCH);
    const bool wantLight = g->lightDiffuse || g->lightSpecular;
    if ((rc = diff_backward_launch(J, a, g->texels != nullptr, g->vertexColors != nullptr, wantLight))) return rc;
    if (g->texels && a.gTex) {
        const PFFtexture& T = J->srcTex[a.paramTex];
        HIP_TRY(hipMemcpyAsync(g->texels, a.gTex, (size_t)T.width * T.height * 16, hipMemcpyDeviceToHost, J->stream),
                PFD_E_LAUNCH);
    }
    if (g->vertexColors && a.gCol)
        HIP_TRY(hipMemcpyAsync(g->vertexColors, a.gCol, (size_t)a.numVertices * 16, hipMemcpyDeviceToHost, J->stream),
                PFD_E_LAUNCH);
    float lg[DIFF_LIGHT_VALS];
    if (wantLight)
        HIP_TRY(hipMemcpyAsync(lg, J->gLight.p, sizeof lg, hipMemcpyDeviceToHost, J->stream), PFD_E_LAUNCH);
    HIP_TRY(hipStreamSynchronize(J->stream), PFD_E_LAUNCH);
    if (g->lightDiffuse) memcpy(g->lightDiffuse, lg, PFF_MAX_LIGHTS * 3 * sizeof(float));
    if (g->lightSpecular) memcpy(g->lightSpecular, lg + PFF_MAX_LIGHTS * 3, PFF_MAX_LIGHTS * 3 * sizeof(float));
    return PFD_OK;
}

int pfdDiffTime(PFDjob* J, const PFDdiffParams* p, uint32_t iters, PFDdiffTimes* t) {
    if (!t || iters == 0) return PFD_E_INVALID;
    DiffArgs a;
    int rc;
    if ((rc = diff_raster(J)) || (rc = diff_args(J, p, a))) return rc;
    // dL/dimage: leave whatever the device buffer holds (timing only).
    hipEvent_t ev[4];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e), PFD_E_LAUNCH);
    float raster = 0.0f, fwd = 0.0f, bwd = 0.0f;
    for (uint32_t i = 0; i < iters && rc == 0; i++) {
        (void)hipEventRecord(ev[0], J->stream);
        rc = pfdJobLaunch(J);
        (void)hipEventRecord(ev[1], J->stream);
        hipLaunchKernelGGL(k_diff<false>, diff_grid(a), dim3(256), 0, J->stream, a);
        (void)hipEventRecord(ev[2], J->stream);
        DiffArgs b = a;
        if (!rc) rc = diff_backward_launch(J, b, a.paramTex >= 0, a.colParams != nullptr, true);
        (void)hipEventRecord(ev[3], J->stream);
        if (hipEventSynchronize(ev[3]) != hipSuccess) rc = PFD_E_LAUNCH;
        float m0 = 0, m1 = 0, m2 = 0;
        (void)hipEventElapsedTime(&m0, ev[0], ev[1]);
        (void)hipEventElapsedTime(&m1, ev[1], ev[2]);
        (void)hipEventElapsedTime(&m2, ev[2], ev[3]);
        raster += m0; fwd += m1; bwd += m2;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    t->rasterMs = raster / iters;
    t->forwardMs = fwd / iters;
    t->backwardMs = bwd / iters;
    return PFD_OK;
}

}  // extern "C"
