/* Binary scene format: a recorded list of BLContext calls.  Written by the
 * Python scene writer (blend2d_amd/scene.py); played through the MI355X
 * backend by scene_player.cpp and blend2d_amd/replay.py, through the
 * reference library by oracle/ref_render.cpp and through the CPU oracle by
 * oracle/orc_scene.c, so the same byte string drives all three.  All fields
 * are little endian.
 *
 *   SceneHeader
 *   n_images x { SceneImage, pixel bytes (h * stride, padded to 8) }
 *   n_cmds   x { SceneCmd, payload (size bytes, multiple of 8) }
 *
 * The rendering library (libblend2d_amd.so) never includes this header.
 */
#ifndef B2AMD_SCENE_FORMAT_H
#define B2AMD_SCENE_FORMAT_H

#include <stdint.h>

#define SCENE_MAGIC 0x43533242u /* "B2SC" */
#define SCENE_VERSION 1u

/* Pixel formats, values equal to BLFormat (core/format.h). */
enum { SCENE_FMT_NONE = 0, SCENE_FMT_PRGB32 = 1, SCENE_FMT_XRGB32 = 2, SCENE_FMT_A8 = 3 };

typedef struct SceneHeader {
  uint32_t magic;
  uint32_t version;
  uint32_t width;
  uint32_t height;
  uint32_t format;
  uint32_t n_images;
  uint32_t n_cmds;
  uint32_t clear_value;   /* raw pixel value the canvas starts with (A8 uses the low byte) */
} SceneHeader;

typedef struct SceneImage {
  uint32_t width;
  uint32_t height;
  uint32_t format;
  uint32_t stride;        /* bytes per row, the pixel rows follow */
} SceneImage;

typedef struct SceneCmd {
  uint32_t op;
  uint32_t size;          /* payload size in bytes */
} SceneCmd;

enum SceneOp {
  SCENE_OP_SET_COMP_OP        = 1,  /* u32 comp_op, u32 pad */
  SCENE_OP_SET_FILL_RULE      = 2,  /* u32 fill_rule, u32 pad */
  SCENE_OP_SET_TRANSFORM      = 3,  /* f64[6] user transform (m00 m01 m10 m11 m20 m21) */
  SCENE_OP_SET_GLOBAL_ALPHA   = 4,  /* f64 alpha */
  SCENE_OP_SET_STYLE_SOLID    = 5,  /* u32 rgba32 (straight alpha, 0xAARRGGBB), u32 pad */
  SCENE_OP_SET_STYLE_GRADIENT = 6,  /* SceneGradient + n_stops x SceneStop */
  SCENE_OP_SET_STYLE_PATTERN  = 7,  /* ScenePattern */
  SCENE_OP_SET_PATTERN_QUALITY = 8, /* u32 quality (0 nearest, 1 bilinear), u32 pad */
  SCENE_OP_FILL_PATH          = 9,  /* u32 n, u32 pad, u8 cmds[n] padded to 8, f64 vtx[2n] */
  SCENE_OP_FILL_POLYGON       = 10, /* u32 n, u32 pad, f64 vtx[2n] */
  SCENE_OP_FILL_RECT_D        = 11, /* f64 x, y, w, h */
  SCENE_OP_FILL_RECT_I        = 12, /* i32 x, y, w, h */
  SCENE_OP_FILL_ALL           = 13, /* no payload */
  SCENE_OP_SET_FILL_ALPHA     = 14, /* f64 alpha */
  SCENE_OP_SET_STYLE_RGBA64   = 15, /* u64 rgba64 (16 bits per channel, straight alpha) */
  SCENE_OP_SET_STYLE_RGBA     = 16, /* f32 r, g, b, a */
  SCENE_OP_CLEAR_ALL          = 17, /* no payload */
  SCENE_OP_CLEAR_RECT_I       = 18, /* i32 x, y, w, h */
  SCENE_OP_CLEAR_RECT_D       = 19, /* f64 x, y, w, h */
  SCENE_OP_FILL_GEOMETRY      = 20, /* SceneGeometry + geometry data (see below) */
  SCENE_OP_FILL_ALL_RGBA32    = 21, /* u32 rgba32, u32 pad */
  SCENE_OP_FILL_RECT_I_RGBA32 = 22, /* i32 x, y, w, h, u32 rgba32, u32 pad */
  SCENE_OP_FILL_RECT_D_RGBA32 = 23, /* f64 x, y, w, h, u32 rgba32, u32 pad */
  SCENE_OP_FILL_PATH_EX       = 24, /* ScenePathEx, u8 cmds[n] padded to 8, f64 vtx[2n] */
  SCENE_OP_FILL_EXPLICIT      = 25, /* SceneExplicit, style payload, target payload (see below) */
  SCENE_OP_BLIT_IMAGE         = 26, /* SceneBlit */
  SCENE_OP_FILL_MASK          = 27, /* SceneMask */
  SCENE_OP_SET_GRADIENT_QUALITY = 28 /* u32 quality (0 nearest, 1 smooth, 2 dither), u32 pad */
};

/* FILL_MASK: fill_mask_i / _d (kinds 0, 1) with the context's fill style of
 * image `image_index` as the mask, its `area` (x, y, w, h) when has_area,
 * else the whole image, placed at dst = x, y (integers for kind 0).  The
 * same struct is the target payload of FILL_EXPLICIT target 5. */
typedef struct SceneMask {
  uint32_t kind;
  uint32_t image_index;
  uint32_t has_area;
  uint32_t pad;
  int32_t area[4];
  double dst[2];
} SceneMask;

/* BLIT_IMAGE: blit_image_i / _d (kinds 0, 1: dst = x, y) and
 * blit_scaled_image_i / _d (kinds 2, 3: dst = x, y, w, h) of image
 * `image_index`; `area` (x, y, w, h in the image) when has_area, else the
 * whole image.  The _i kinds carry integers in `dst`. */
typedef struct SceneBlit {
  uint32_t kind;
  uint32_t image_index;
  uint32_t has_area;
  uint32_t pad;
  int32_t area[4];
  double dst[4];
} SceneBlit;

/* FILL_EXPLICIT: a fill with an explicit style for this call only -- the
 * fill_*_rgba64 (style kind 1) and fill_*_ext (kinds 0-3) entry points of
 * core/context.h:735-758.  After the header comes the style payload (kind 2:
 * SceneGradient + n_stops x SceneStop; kind 3: ScenePattern; kinds 0 and 1:
 * nothing, the colour is in `rgba`), then the target payload: target 0
 * (fill_all) nothing, 1 (fill_rect_i) i32 x, y, w, h, 2 (fill_rect_d)
 * f64 x, y, w, h, 3 (fill_path_d) u8 cmds[count] padded to 8 and
 * f64 vtx[2 count], 4 (fill_geometry) the geometry data as FILL_GEOMETRY
 * (struct, or `count` array elements), 5 (fill_mask) a SceneMask. */
typedef struct SceneExplicit {
  uint32_t target;        /* 0 all, 1 rect_i, 2 rect_d, 3 path, 4 geometry, 5 mask */
  uint32_t style_kind;    /* 0 rgba32, 1 rgba64, 2 gradient, 3 pattern */
  uint32_t geom_type;     /* target 4: BLGeometryType */
  uint32_t count;         /* target 3: vertices; target 4: array elements */
  uint64_t rgba;          /* kind 0: rgba32, kind 1: rgba64 (straight alpha) */
} SceneExplicit;

/* FILL_GEOMETRY: BLGeometryType `type`; for array types (POLYGON*, POLYLINE*,
 * ARRAY_VIEW_*) `count` elements follow (BLPointI / BLPoint / BLBoxI / BLBox /
 * BLRectI / BLRect), otherwise the geometry struct itself. */
typedef struct SceneGeometry {
  uint32_t type;
  uint32_t has_rgba32;    /* 1: explicit colour (fill_geometry_rgba32) */
  uint32_t rgba32;
  uint32_t count;
} SceneGeometry;

typedef struct ScenePathEx {
  uint32_t n;
  uint32_t flags;         /* bit 0: origin present, bit 1: explicit colour */
  uint32_t rgba32;
  uint32_t pad;
  double origin[2];
} ScenePathEx;

typedef struct SceneGradient {
  uint32_t type;          /* 0 linear, 1 radial, 2 conic (BLGradientType) */
  uint32_t extend;        /* 0 pad, 1 repeat, 2 reflect (BLExtendMode) */
  uint32_t n_stops;
  uint32_t has_transform;
  double values[6];       /* linear: x0 y0 x1 y1; radial: x0 y0 x1 y1 r0 r1; conic: x0 y0 angle repeat */
  double transform[6];
} SceneGradient;

typedef struct SceneStop {
  double offset;
  uint32_t rgba32;        /* straight alpha 0xAARRGGBB */
  uint32_t pad;
} SceneStop;

typedef struct ScenePattern {
  uint32_t image_index;
  uint32_t extend;        /* BLExtendMode, simple modes 0..2 */
  uint32_t has_transform;
  uint32_t pad;
  double transform[6];
} ScenePattern;

#endif /* B2AMD_SCENE_FORMAT_H */
