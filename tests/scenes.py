"""Named test scenes shared by the golden-fixture generator and the tests.

Each entry is (name, keyword arguments of pixelforge_amd.scene()).  Sizes
are small so that the oracle and the compiled reference finish in well
under a second per scene.  The goldens of the `BILINEAR` scenes come from the
reference built with its bilinear colour lerp fixed (color.h:141 reads an
uninitialised vector; oracle/Makefile `ref-bilinear`).
"""
import pixelforge_amd as pf

W, H = 128, 96

# Feature sets of the benchmark configurations at fixture size.
CONFIG_SCENES = [
    # config 1: random NDC triangles, Gouraud-interpolated colours, LESS
    ("c1_ndc_less", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=400, seed=11)),
    # config 2 feature set with nearest filtering (bilinear has no golden)
    ("c2_heightfield_nearest", dict(kind=pf.HEIGHTFIELD, width=W, height=H, gridX=60, gridZ=12, seed=12,
                                    texture=1, texW=64, texH=64, wrap=pf.PF_REPEAT, filter=pf.PF_NEAREST,
                                    depthFunc=pf.PF_LEQUAL, blendFraction=0.3, cull=-1)),
    # config 3: spheres, Phong with 4 point lights, nearest + clamp texture
    ("c3_spheres_phong", dict(kind=pf.SPHERES, width=W, height=H, spheres=6, sphereLat=10, sphereLon=20, seed=13,
                              lighting=2, numLights=4, texture=1, texW=33, texH=33, wrap=pf.PF_CLAMP_TO_EDGE,
                              filter=pf.PF_NEAREST, cull=pf.PF_BACK)),
]

FEATURE_SCENES = [
    ("flat_shading", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=200, seed=21, shade=pf.PF_FLAT)),
    ("no_depth_test", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=200, seed=22, depthTest=0)),
    ("no_clear", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=150, seed=23, clearFirst=0)),
    ("odd_size_clear", dict(kind=pf.RANDOM_NDC, width=W - 3, height=H - 5, ntris=150, seed=24)),
    # W*H = 7 mod 8: the clear copies pixel 0 into the last 7 pixels, and
    # triangles of the same frame then draw over pixel 0 itself
    ("odd_clear_pixel0_3d", dict(kind=pf.RANDOM_3D, width=125, height=91, ntris=300, seed=85)),
    ("clip_3d", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=400, seed=25)),
    ("clip_3d_cull_front", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=400, seed=26, cull=pf.PF_FRONT)),
    ("clip_3d_cull_back", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=400, seed=27, cull=pf.PF_BACK)),
    ("gouraud_lights", dict(kind=pf.SPHERES, width=W, height=H, spheres=6, sphereLat=10, sphereLon=20, seed=28,
                            lighting=1, numLights=3, cull=pf.PF_BACK)),
    ("gouraud_spot_att", dict(kind=pf.SPHERES, width=W, height=H, spheres=6, sphereLat=10, sphereLon=20, seed=29,
                              lighting=1, numLights=2, spot=1, attenuation=1, cull=-1)),
    ("phong_spot_att", dict(kind=pf.SPHERES, width=W, height=H, spheres=6, sphereLat=10, sphereLon=20, seed=30,
                            lighting=2, numLights=2, spot=1, attenuation=1, cull=-1)),
    ("phong_3d_random", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=200, seed=31, lighting=2, numLights=2)),
    ("gouraud_color_material", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=200, seed=32, lighting=1,
                                    numLights=2, colorMaterial=1)),
    ("texture_float", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=200, seed=33, texture=2, texW=32, texH=16)),
    ("multi_frame", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=300, seed=34, flushEvery=70)),
]

BLEND_SCENES = [
    (f"blend_{m}", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=150, seed=40 + m, blend=1, blendMode=m))
    for m in range(8)
]

DEPTH_SCENES = [
    (f"depth_{f}", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=150, seed=50 + f, depthFunc=f))
    for f in range(6)
]

# Odd texture sizes: with an even size, MIRRORED and CLAMP_TO_EDGE map a
# coordinate of exactly 1.0 to index size (rne(size - 1 + 0.5) rounds up),
# and the reference then reads past the texture (sampler.h:236-255).
WRAP_SCENES = [
    (f"wrap_{w}_2d", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=150, seed=60 + w, texture=1,
                          texW=49, texH=41, wrap=w, filter=pf.PF_NEAREST))
    for w in range(3)
] + [
    (f"wrap_{w}_3d", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=250, seed=63 + w, texture=1,
                          texW=65, texH=33, wrap=w, filter=pf.PF_NEAREST))
    for w in range(3)
]

# Even sizes in MIRRORED / CLAMP_TO_EDGE: out-of-texture reads, which the
# reference leaves undefined and this library clamps to the last texel.
# Bilinear CLAMP_TO_EDGE at an even side reads past the texture as soon as
# u + 1/w reaches 1 (the second tap's index rounds to w).
EDGE_READ_SCENES = [
    (f"edge_read_{w}", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=250, seed=66 + w, texture=1,
                            texW=48, texH=40, wrap=w, filter=pf.PF_NEAREST))
    for w in (1, 2)
] + [
    ("edge_read_bilinear_2", dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=250, seed=72, texture=1,
                                  texW=64, texH=64, wrap=pf.PF_CLAMP_TO_EDGE, filter=pf.PF_BILINEAR)),
]

# Bilinear filtering in every wrap mode (odd sides for MIRRORED and
# CLAMP_TO_EDGE; CLAMP_TO_EDGE at even sides on the 2D path with vertex
# coordinates below 1 - 2/w, so that no tap leaves the texture -- the 2D
# path interpolates them affinely, while the 3D path divides by the
# interpolated 1/z of clip space, which can leave the vertices' range),
# float texels, the config-2 height field and a batch-style additive
# context.  Goldens: the reference with the colour-lerp fix (module doc).
_BL = dict(kind=pf.RANDOM_3D, width=W, height=H, ntris=250, texture=1, filter=pf.PF_BILINEAR)
BILINEAR_SCENES = [
    ("bilinear_0", dict(_BL, seed=70, texW=64, texH=64, wrap=pf.PF_REPEAT)),
    ("bilinear_1", dict(_BL, seed=71, texW=63, texH=47, wrap=pf.PF_MIRRORED_REPEAT)),
    ("bilinear_1_even", dict(_BL, seed=75, texW=64, texH=64, wrap=pf.PF_MIRRORED_REPEAT)),
    ("bilinear_2", dict(_BL, seed=73, texW=63, texH=47, wrap=pf.PF_CLAMP_TO_EDGE)),
    ("bilinear_2_even_inner", dict(kind=pf.RANDOM_NDC, width=W, height=H, ntris=200, seed=76, texture=1,
                                   filter=pf.PF_BILINEAR, texW=64, texH=48, wrap=pf.PF_CLAMP_TO_EDGE,
                                   uvMin=-0.5, uvMax=1.0 - 3.0 / 64.0)),
    ("bilinear_2d", dict(kind=pf.RANDOM_NDC, width=W - 7, height=H - 9, ntris=200, seed=77, texture=1, texW=45,
                         texH=29, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR)),
    ("bilinear_float", dict(_BL, seed=78, texture=2, texW=33, texH=17, wrap=pf.PF_REPEAT)),
    ("c2_heightfield_bilinear", dict(kind=pf.HEIGHTFIELD, width=W, height=H, gridX=60, gridZ=12, seed=74,
                                     texture=1, texW=64, texH=64, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR,
                                     depthFunc=pf.PF_LEQUAL, blendFraction=0.3, cull=-1)),
    ("bilinear_batch_add", dict(kind=pf.HEIGHTFIELD, width=113, height=87, gridX=40, gridZ=6, seed=79, texture=1,
                                texW=96, texH=72, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR,
                                depthFunc=pf.PF_LEQUAL, cull=pf.PF_BACK, blend=1, blendMode=pf.PF_BLEND_ADD,
                                blendFraction=1.0)),
]
BILINEAR_GOLDEN_NAMES = frozenset(n for n, _ in BILINEAR_SCENES)

# Vertex arrays (pfDrawElements / pfDrawArrays) against the same mesh drawn
# in immediate mode; every index, position and colour type.
_MESH = dict(kind=pf.MESH, width=W, height=H, gridX=12, gridZ=10, texture=1, texW=33, texH=33,
             filter=pf.PF_NEAREST, lighting=2, numLights=2, cull=-1)
ARRAY_SCENES = [
    ("mesh_immediate", dict(_MESH, seed=80, arrays=0)),
    ("mesh_elements_uint", dict(_MESH, seed=80, arrays=1)),
    ("mesh_elements_ubyte", dict(_MESH, seed=80, arrays=1, indexType=1)),
    ("mesh_elements_ushort", dict(_MESH, seed=80, arrays=1, indexType=2)),
    ("mesh_draw_arrays", dict(_MESH, seed=80, arrays=2)),
    ("mesh_double_positions", dict(_MESH, seed=80, arrays=1, posType=1)),
    ("mesh_colors_ubyte", dict(_MESH, seed=81, arrays=1, colType=1)),
    ("mesh_colors_ushort", dict(_MESH, seed=81, arrays=1, colType=2)),
    ("mesh_colors_uint", dict(_MESH, seed=81, arrays=1, colType=3)),
    ("mesh_colors_double", dict(_MESH, seed=81, arrays=1, colType=4)),
    ("mesh_colors_rgb", dict(_MESH, seed=82, arrays=1, colSize=3)),
    ("mesh_gouraud_cull", dict(_MESH, seed=83, arrays=1, lighting=1, cull=pf.PF_BACK)),
    # every other triangle mode, one draw call per row of cells (strip order,
    # or the cells' corners for PF_QUADS), culling off, front or back
    ("mesh_strip_elements", dict(_MESH, seed=84, arrays=1, arrayMode=pf.PF_TRIANGLE_STRIP)),
    ("mesh_strip_cull_back", dict(_MESH, seed=84, arrays=1, arrayMode=pf.PF_TRIANGLE_STRIP, cull=pf.PF_BACK)),
    ("mesh_fan_cull_front", dict(_MESH, seed=85, arrays=1, arrayMode=pf.PF_TRIANGLE_FAN, cull=pf.PF_FRONT,
                                 indexType=2)),
    ("mesh_quads_elements", dict(_MESH, seed=86, arrays=1, arrayMode=pf.PF_QUADS)),
    ("mesh_quads_arrays_cull", dict(_MESH, seed=86, arrays=2, arrayMode=pf.PF_QUADS, cull=pf.PF_BACK)),
    ("mesh_quad_strip_arrays", dict(_MESH, seed=87, arrays=2, arrayMode=pf.PF_QUAD_STRIP, lighting=1)),
    ("mesh_quad_fan_elements", dict(_MESH, seed=88, arrays=1, arrayMode=pf.PF_QUAD_FAN, colType=1)),
]

# Texture formats: every (format, type) pair with a texel getter other than
# RGBA8 and RGBA float (covered above), sampled NEAREST at a small size.
# Names: tex_<format>_<type>.  Half and float texels include values above 1,
# negative, infinite and NaN.
_FORMATS = dict(red=0, green=1, blue=2, alpha=3, lum=4, lumalpha=5, rgb=6, rgba=7, bgr=8, bgra=9)
_TYPES = dict(ubyte=0, u565=2, u5551=3, u4444=4, half=9, float=10)
_PAIRS = ([(f, t) for f in ("red", "green", "blue", "alpha", "lum", "lumalpha") for t in ("ubyte", "half", "float")]
          + [(f, t) for f in ("rgb", "bgr") for t in ("ubyte", "u565", "half", "float")]
          + [("rgba", t) for t in ("u5551", "u4444", "half")]
          + [("bgra", t) for t in ("ubyte", "u5551", "u4444", "half", "float")])
FORMAT_SCENES = [
    (f"tex_{f}_{t}", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=200 + k, texture=3,
                          texFormat=_FORMATS[f], texType=_TYPES[t], texW=13, texH=11, filter=pf.PF_NEAREST))
    for k, (f, t) in enumerate(_PAIRS)
]

# pfDrawPixels: three calls (plain; zoomed and blended; zoomed without the
# depth test) between two halves of the triangles, images in every (format,
# type) pair with a pixel getter; one 3D scene.
_ALL_PAIRS = _PAIRS + [("rgba", "ubyte"), ("rgba", "float")]
PIXEL_SCENES = [
    (f"pixels_{f}_{t}", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=300 + k, pixelOps=3,
                             pixFormat=_FORMATS[f], pixType=_TYPES[t], depthFunc=pf.PF_LEQUAL))
    for k, (f, t) in enumerate(_ALL_PAIRS)
] + [
    ("pixels_3d_rgba_ubyte", dict(kind=pf.RANDOM_3D, width=96, height=64, ntris=120, seed=340, pixelOps=6,
                                  pixFormat=7, pixType=0)),
]

# pfFogProcess after lit spheres (depths 2.5..4.3, background at FLT_MAX):
# each fog mode, an opaque fog colour, the setters' quirks, and fog between
# two halves of a 3D scene's triangles.
_FOG = dict(kind=pf.SPHERES, width=64, height=48, spheres=5, sphereLat=8, sphereLon=14, lighting=1, numLights=2,
            cull=pf.PF_BACK, fogStart=2.6, fogEnd=3.4, fogAlpha=200)
FOG_SCENES = [
    ("fog_linear", dict(_FOG, seed=400, fog=1)),
    ("fog_exp", dict(_FOG, seed=401, fog=2, fogStart=2.4, fogEnd=3.0)),
    ("fog_exp2", dict(_FOG, seed=402, fog=3)),
    ("fog_opaque", dict(_FOG, seed=403, fog=1, fogAlpha=255)),
    ("fog_setter_quirks", dict(_FOG, seed=404, fog=4)),
    ("fog_between_draws_3d", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=80, seed=405, fog=1,
                                  fogStart=-0.2, fogEnd=1.5, fogAlpha=180)),
]

# Points: PF_POINTS between two halves of the triangles (single pixels,
# discs, blended discs), and the PF_POINT polygon mode for triangles (front,
# both faces twice, culled) and for quads (the render-list sphere's lower
# half).
POINT_SCENES = [
    ("points_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=500, points=300,
                        depthFunc=pf.PF_LEQUAL)),
    ("points_discs_3d", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=501, points=120, pointSize=5.5)),
    ("points_blend", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=502, points=150, pointSize=3.0,
                          blend=1, blendMode=3)),
    ("poly_point_front", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=80, seed=503, polyPoint=1,
                              pointSize=2.0)),
    ("poly_point_both", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=80, seed=504, polyPoint=3, blend=1,
                             blendMode=2)),
    ("poly_point_culled", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=80, seed=505, polyPoint=2,
                               cull=pf.PF_FRONT, pointSize=4.0)),
    ("poly_point_quads_list", dict(kind=pf.LIST, width=W, height=H, texture=1, texW=33, texH=33,
                                   filter=pf.PF_NEAREST, seed=506, polyPoint=3, pointSize=3.0, cull=-1)),
]

# Lines: PF_LINES between two halves of the triangles (thin, thick, depth
# test off, blending; 2D and 3D clipping), and the PF_LINE polygon mode for
# triangles and quads.
LINE_SCENES = [
    ("lines_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=600, lines=80,
                       depthFunc=pf.PF_LEQUAL)),
    ("lines_3d", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=601, lines=80)),
    ("lines_thick_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=602, lines=40,
                             lineWidth=4.5)),
    ("lines_thick_nodepth", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=603, lines=40,
                                 lineWidth=3.0, depthTest=0)),
    ("lines_blend_3d", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=604, lines=60, lineWidth=2.5,
                            blend=1, blendMode=5)),
    ("poly_line_both", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=605, polyLine=3)),
    ("poly_line_back_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=60, seed=606, polyLine=2,
                                lineWidth=2.0)),
    ("lines_many_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=20, seed=608, lines=600,
                            blend=1, blendMode=1)),
    ("poly_line_quads_list", dict(kind=pf.LIST, width=W, height=H, texture=1, texW=33, texH=33,
                                  filter=pf.PF_NEAREST, seed=607, polyLine=3, cull=-1)),
]

# pfRectf fills between two halves of the triangles (2D, and 3D where the
# corners go through a perspective MVP without division).
RECT_SCENES = [
    ("rects_ndc", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=700, rects=6)),
    ("rects_3d", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60, seed=701, rects=5)),
]

# Render targets other than RGBA8 (the render-target pairs of the
# library): per pair, a textured 3D scene drawn blended into a 61x45
# offscreen framebuffer (SIMD getter and setter of the raster, the clear
# and its one-pixel tail), a 2D scene with pfDrawPixels, points, lines,
# rects and fog between its halves (the scalar accessors), and an
# unblended 3D scene in the context's main buffer created in that pair.
_TARGETS = [("rgb", "ubyte"), ("bgr", "ubyte"), ("rgb", "u565"), ("bgr", "u565"),
            ("rgba", "u5551"), ("bgra", "u5551"), ("rgba", "u4444"), ("bgra", "u4444")]
# Targets in half and float components.
_WIDE_TARGETS = [(f, t) for f in ("lumalpha", "rgb", "bgr", "rgba", "bgra") for t in ("half", "float")]
TARGET_SCENES = [
    (f"target_{f}_{t}_raster", dict(kind=pf.RANDOM_3D, width=61, height=45, ntris=150, seed=800 + k, texture=1,
                                    texW=17, texH=13, blend=1, blendMode=1 + k % 7,
                                    fbFormat=_FORMATS[f], fbType=_TYPES[t]))
    for k, (f, t) in enumerate(_TARGETS + _WIDE_TARGETS)
] + [
    (f"target_{f}_{t}_prims", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=40, seed=820 + k, pixelOps=3,
                                   pixFormat=7, pixType=0, points=120, pointSize=3.0, lines=40, lineWidth=2.0,
                                   rects=3, fog=1, fogStart=-0.5, fogEnd=0.8, fogAlpha=200, blend=1, blendMode=1,
                                   depthFunc=pf.PF_LEQUAL, fbFormat=_FORMATS[f], fbType=_TYPES[t]))
    for k, (f, t) in enumerate(_TARGETS + _WIDE_TARGETS)
] + [
    (f"target_{f}_{t}_main", dict(kind=pf.RANDOM_3D, width=41, height=29, ntris=100, seed=840 + k,
                                  fbFormat=_FORMATS[f], fbType=_TYPES[t], fbMain=1))
    for k, (f, t) in enumerate(_TARGETS + [("rgba", "ubyte")] + _WIDE_TARGETS)
]

# Golden frames rendered by the reference with extra scene parameters.
# BGRA / UNSIGNED_BYTE: the reference's AVX2 getter shuffles with a mask that
# repeats 2,1,0,3 at every 4 bytes of a 128-bit register (pixel.h:2916-2920,
# simd.h:564-574), so each group of four SIMD lanes gets its first lane's
# texel; this library decodes every texel on its own (as the scalar getter,
# pixel.h:702-706), which equals the reference sampling the same texels
# handed over as RGBA / UNSIGNED_BYTE.
GOLDEN_REF_OVERRIDES = {"tex_bgra_ubyte": dict(texSwizzle=1)}

# Render lists (pfNewList / pfEndList / pfCallList): spheres recorded once and
# called at several transforms, a list recorded by calling another, state
# restored at pfEndList; unlit, Gouraud, and Phong with colour material.
_LIST = dict(kind=pf.LIST, width=W, height=H, texture=1, texW=33, texH=33, filter=pf.PF_NEAREST)
LIST_SCENES = [
    ("list_unlit", dict(_LIST, seed=100, cull=pf.PF_BACK)),
    ("list_gouraud", dict(_LIST, seed=101, lighting=1, numLights=2, cull=pf.PF_BACK)),
    ("list_phong_color_material", dict(_LIST, seed=102, lighting=2, numLights=2, colorMaterial=1, cull=-1)),
    ("list_flat_blend", dict(_LIST, seed=103, shade=pf.PF_FLAT, blend=1, blendMode=1, cull=pf.PF_FRONT)),
    # called under a texture matrix and PF_NORMALIZE (both applied when the
    # calls replay), culling off, list A re-recorded before pfFinish
    ("list_texmat_normalize", dict(_LIST, seed=104, lighting=2, numLights=2, cull=-1, listOptions=7)),
    ("list_texmat_gouraud", dict(_LIST, seed=105, lighting=1, numLights=3, cull=pf.PF_BACK, listOptions=3)),
]

# Targets smaller than a block or a tile, with sides that are no multiple of
# any tile size, and a frame that only clears (GPU against the oracle).
SIZE_SCENES = [
    ("tiny_1x1", dict(kind=pf.RANDOM_NDC, width=1, height=1, ntris=20, seed=90)),
    ("tiny_7x3", dict(kind=pf.RANDOM_NDC, width=7, height=3, ntris=20, seed=91)),
    ("clear_only", dict(kind=pf.RANDOM_NDC, width=64, height=48, ntris=0, seed=92)),
    ("odd_129x77_3d", dict(kind=pf.RANDOM_3D, width=129, height=77, ntris=300, seed=93, texture=1, texW=31,
                           texH=17, filter=pf.PF_BILINEAR)),
]

GOLDEN_SCENES = (CONFIG_SCENES + FEATURE_SCENES + BLEND_SCENES + DEPTH_SCENES + WRAP_SCENES + ARRAY_SCENES
                 + LIST_SCENES + FORMAT_SCENES + PIXEL_SCENES + FOG_SCENES + POINT_SCENES + LINE_SCENES
                 + RECT_SCENES + BILINEAR_SCENES + TARGET_SCENES)
# Many triangles crossing the clip planes: random 3D triangles around the
# camera, about half of them clipped.  The library clips a pass's listed
# triangles 16 lanes per triangle up to 16k of them and one lane per
# triangle beyond; these scenes take one path each (about 11k and 34k
# clipped triangles).
CLIP_LOAD_SCENES = [
    ("clip_load_20k", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=20000, seed=77)),
    ("clip_load_60k", dict(kind=pf.RANDOM_3D, width=64, height=48, ntris=60000, seed=77)),
]
ALL_SCENES = GOLDEN_SCENES + EDGE_READ_SCENES + SIZE_SCENES + CLIP_LOAD_SCENES


def params(kw):
    return pf.scene(**kw)


# Benchmark configurations at full size.  Their goldens are SHA-256 digests
# of the reference's colour and depth bytes (tests/golden/full_sizes.json):
# the images are megabytes each.  Config 2 samples bilinearly, so its digest
# comes from the reference with the colour-lerp fix; config 3 is not here,
# because its even-sided CLAMP texture is read past its end by the
# reference at coordinate 1.0 (deviation D3).
FULL_SIZE_SCENES = [
    ("c1_full_800x600", dict(kind=pf.RANDOM_NDC, width=800, height=600, ntris=10000, seed=1, cull=-1,
                             depthFunc=pf.PF_LESS)),
    ("c2_full_1080p", dict(kind=pf.HEIGHTFIELD, width=1920, height=1080, seed=2, gridX=1000, gridZ=100,
                           texture=1, texW=1024, texH=1024, wrap=pf.PF_REPEAT, filter=pf.PF_BILINEAR,
                           depthFunc=pf.PF_LEQUAL, blendFraction=0.3, cull=pf.PF_BACK)),
]
FULL_SIZE_BILINEAR = frozenset({"c2_full_1080p"})


def digest(color, depth):
    """(colour, depth) SHA-256 hex digests of the raw bytes."""
    import hashlib
    return (hashlib.sha256(color.tobytes()).hexdigest(), hashlib.sha256(depth.tobytes()).hexdigest())
