/* blend2d_amd -- MI355X (gfx950) backend for Blend2D's fill path.
 *
 * C ABI of libblend2d_amd.so.  Every entry point returns a BLResult-compatible
 * code (0 = success, errors from 0x10000, core/api.h:1143-1152 of the
 * reference) and never throws.  Types are plain: geometry arrives as the
 * arrays a BLPath stores (command bytes + double vertex pairs,
 * core/path.h:22-39, 411), colours as 0xAARRGGBB straight-alpha rgba32,
 * matrices as six doubles [m00 m01 m10 m11 m20 m21] (core/matrix.h).
 *
 * Each entry cites the reference C API it replaces.  The context records
 * calls and renders them on the GPU at flush/end, the asynchronous model of
 * the reference (BLContextCreateInfo::thread_count > 0, context.h:103-110),
 * so pixels become visible after bl_amd_context_flush(SYNC) or end().
 */
#ifndef BLEND2D_AMD_H
#define BLEND2D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint32_t BLAmdResult;
typedef struct BLAmdImage BLAmdImage;
typedef struct BLAmdContext BLAmdContext;
typedef struct BLAmdBatch BLAmdBatch;

enum {
  BL_AMD_SUCCESS = 0,
  BL_AMD_ERROR_OUT_OF_MEMORY = 0x10000u,
  BL_AMD_ERROR_INVALID_VALUE = 0x10001u,
  BL_AMD_ERROR_INVALID_STATE = 0x10002u,
  BL_AMD_ERROR_NOT_IMPLEMENTED = 0x10007u,
  BL_AMD_ERROR_DATA_TOO_LARGE = 0x10032u,
  BL_AMD_ERROR_INVALID_GEOMETRY = 0x10034u,
  BL_AMD_ERROR_DEVICE = 0x1F000u
};

/* BLContextErrorFlags (context.h:156-189) that rendering can raise. */
enum {
  BL_AMD_CONTEXT_ERROR_FLAG_INVALID_GEOMETRY = 0x00000004u,
  BL_AMD_CONTEXT_ERROR_FLAG_OUT_OF_MEMORY = 0x40000000u
};

/* Pixel formats: BLFormat (core/format.h). */
enum { BL_AMD_FORMAT_PRGB32 = 1, BL_AMD_FORMAT_XRGB32 = 2, BL_AMD_FORMAT_A8 = 3 };

/* BLGeometryType values (core/geometry.h:40-86). */
enum {
  BL_AMD_GEOMETRY_NONE = 0, BL_AMD_GEOMETRY_BOXI = 1, BL_AMD_GEOMETRY_BOXD = 2, BL_AMD_GEOMETRY_RECTI = 3,
  BL_AMD_GEOMETRY_RECTD = 4, BL_AMD_GEOMETRY_CIRCLE = 5, BL_AMD_GEOMETRY_ELLIPSE = 6, BL_AMD_GEOMETRY_ROUND_RECT = 7,
  BL_AMD_GEOMETRY_ARC = 8, BL_AMD_GEOMETRY_CHORD = 9, BL_AMD_GEOMETRY_PIE = 10, BL_AMD_GEOMETRY_LINE = 11,
  BL_AMD_GEOMETRY_TRIANGLE = 12, BL_AMD_GEOMETRY_POLYLINEI = 13, BL_AMD_GEOMETRY_POLYLINED = 14,
  BL_AMD_GEOMETRY_POLYGONI = 15, BL_AMD_GEOMETRY_POLYGOND = 16, BL_AMD_GEOMETRY_ARRAY_VIEW_BOXI = 17,
  BL_AMD_GEOMETRY_ARRAY_VIEW_BOXD = 18, BL_AMD_GEOMETRY_ARRAY_VIEW_RECTI = 19, BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD = 20,
  BL_AMD_GEOMETRY_PATH = 21, BL_AMD_GEOMETRY_MAX_VALUE = 21
};

/* BLArrayView<T> (core/api.h:1850-1853): pointer + element count. */
typedef struct BLAmdArrayView { const void* data; size_t size; } BLAmdArrayView;

/* ---- runtime --------------------------------------------------------- */

/* Library version: major << 16 | minor << 8 | patch. */
uint32_t bl_amd_version(void);
/* Returns the number of HIP devices (0 when none is usable). */
int bl_amd_device_count(void);

/* ---- images (BLImage, core/image.h) ---------------------------------- */

/* bl_image_create (core/image.h:157): device-resident image. */
BLAmdResult bl_amd_image_create(BLAmdImage** out, int w, int h, uint32_t format);
/* bl_image_destroy (core/image.h:152).  Drops the caller's reference: the
 * pixels stay allocated while a context renders into the image, a fill style
 * samples it or a recorded or replayable frame uses it. */
BLAmdResult bl_amd_image_destroy(BLAmdImage* img);
/* Copies host pixels into the image (bl_image_make_mutable, core/image.h:160, + write). */
BLAmdResult bl_amd_image_upload(BLAmdImage* img, const void* pixels, intptr_t stride);
/* Copies the image to host memory (bl_image_get_data, core/image.h:159, + read). */
BLAmdResult bl_amd_image_download(const BLAmdImage* img, void* pixels, intptr_t stride);
/* Fills every pixel with a raw value (A8 uses the low byte). */
BLAmdResult bl_amd_image_fill_raw(BLAmdImage* img, uint32_t value);
/* BLImageData: size, format, device address of row 0, stride in bytes. */
BLAmdResult bl_amd_image_get_data(const BLAmdImage* img, int* w, int* h, uint32_t* format,
                                  void** device_pixels, intptr_t* stride);

/* ---- rendering context (BLContext, core/context.h:631-767) ------------- */

/* bl_context_init (context.h:631) / bl_context_destroy (context.h:635). */
BLAmdResult bl_amd_context_create(BLAmdContext** out);
BLAmdResult bl_amd_context_destroy(BLAmdContext* ctx);
/* bl_context_begin (context.h:646): attach a target, reset the state. */
BLAmdResult bl_amd_context_begin(BLAmdContext* ctx, BLAmdImage* target);
/* bl_context_end (context.h:647): flush synchronously and detach. */
BLAmdResult bl_amd_context_end(BLAmdContext* ctx);
/* bl_context_reset (context.h:637): renders the pending calls and detaches
 * the target when one is attached (the reference maps end() to reset()),
 * then restores the default state. */
BLAmdResult bl_amd_context_reset(BLAmdContext* ctx);
/* bl_context_flush (context.h:649); flags: BL_CONTEXT_FLUSH_SYNC = 0x80000000 (context.h:107). */
BLAmdResult bl_amd_context_flush(BLAmdContext* ctx, uint32_t flags);

/* Errors found while the device rendered (edge building happens there, so a
 * fill reports them at the flush, as the reference's asynchronous mode
 * reports errors of deferred work, context.h:156-189): a fill whose line
 * clipper meets NaN coordinates, or a curve with a non-finite control
 * point, draws nothing and raises BL_AMD_CONTEXT_ERROR_FLAG_INVALID_GEOMETRY
 * (raster/edgebuilder_p.h:1494-1496, raster/rastercontextops.cpp:38-45); a
 * fill one of whose curve pieces flattens to more than 2^22 segments draws
 * nothing and raises BL_AMD_CONTEXT_ERROR_FLAG_OUT_OF_MEMORY. */
/* BLContext::accumulated_error_flags() (BLContextState, context.h:581-615):
 * flags accumulated since begin(). */
BLAmdResult bl_amd_context_get_error_flags(const BLAmdContext* ctx, uint32_t* flags);
/* The first rendering call of the last flush that failed on the device:
 * `call_index` counts the context's fill_* / clear_* calls from begin()
 * (0-based, every such call counts), `code` the BLResult the reference's
 * synchronous mode returns from that call (BL_AMD_ERROR_INVALID_GEOMETRY or
 * BL_AMD_ERROR_OUT_OF_MEMORY); code 0 when none failed. */
BLAmdResult bl_amd_context_get_deferred_error(const BLAmdContext* ctx, uint64_t* call_index, uint32_t* code);

/* bl_context_set_comp_op (context.h:697): BLCompOp values; SRC_OVER, SRC_COPY,
 * PLUS, MULTIPLY, SCREEN render, others report NOT_IMPLEMENTED at fill time. */
BLAmdResult bl_amd_context_set_comp_op(BLAmdContext* ctx, uint32_t comp_op);
/* bl_context_set_fill_rule (context.h:700): 0 non-zero, 1 even-odd. */
BLAmdResult bl_amd_context_set_fill_rule(BLAmdContext* ctx, uint32_t fill_rule);
/* bl_context_set_global_alpha (context.h:694). */
BLAmdResult bl_amd_context_set_global_alpha(BLAmdContext* ctx, double alpha);
/* bl_context_apply_transform_op(BL_TRANSFORM_OP_ASSIGN) (context.h:658). */
BLAmdResult bl_amd_context_set_transform(BLAmdContext* ctx, const double m[6]);
/* bl_context_set_hint(BL_CONTEXT_HINT_PATTERN_QUALITY) (context.h:408-422):
 * 0 nearest, 1 bilinear (default). */
BLAmdResult bl_amd_context_set_pattern_quality(BLAmdContext* ctx, uint32_t quality);
/* BLContext::set_hint(BL_CONTEXT_HINT_GRADIENT_QUALITY) (context.h,
 * rastercontext.cpp:1816-1821): 0 nearest, 1 smooth (rendered as nearest),
 * 2 dither; other values are BL_AMD_ERROR_INVALID_VALUE.  A gradient style
 * takes the hint when it is set (explicit styles: when the call is made); a
 * dithered gradient samples a 64-bit table of twice the size (at most 1024
 * entries) and adds the 16 x 16 Bayer threshold of the pixel before cutting
 * each channel to 8 bits.  A8 targets never dither. */
BLAmdResult bl_amd_context_set_gradient_quality(BLAmdContext* ctx, uint32_t quality);

/* bl_context_set_fill_style_rgba32 (context.h:674). */
BLAmdResult bl_amd_context_set_fill_style_rgba32(BLAmdContext* ctx, uint32_t rgba32);
/* bl_context_set_fill_style_rgba64 (context.h:675). */
BLAmdResult bl_amd_context_set_fill_style_rgba64(BLAmdContext* ctx, uint64_t rgba64);
/* bl_context_set_fill_style_rgba (context.h:673): BLRgba {r, g, b, a} floats. */
BLAmdResult bl_amd_context_set_fill_style_rgba(BLAmdContext* ctx, const float rgba[4]);
/* bl_context_set_fill_alpha (context.h:678). */
BLAmdResult bl_amd_context_set_fill_alpha(BLAmdContext* ctx, double alpha);
/* bl_context_set_fill_style(BLGradient) (context.h:671) with the gradient
 * passed by value: type 0 linear [x0 y0 x1 y1], 1 radial [x0 y0 x1 y1 r0 r1],
 * 2 conic [x0 y0 angle repeat]; extend 0 pad, 1 repeat, 2 reflect; stops as
 * BLGradient::add_stop would receive them; transform may be NULL. */
BLAmdResult bl_amd_context_set_fill_style_gradient(BLAmdContext* ctx, uint32_t type, const double* values,
                                                   uint32_t extend, const double* stop_offsets,
                                                   const uint32_t* stop_rgba32, size_t n_stops,
                                                   const double* transform);
/* bl_context_set_fill_style(BLPattern) (context.h:671): image + extend mode +
 * optional pattern transform. */
BLAmdResult bl_amd_context_set_fill_style_pattern(BLAmdContext* ctx, const BLAmdImage* image, uint32_t extend,
                                                  const double* transform);

/* bl_context_fill_path_d (context.h:750): BLPath command bytes and vertices;
 * origin may be NULL (same as fill_geometry(BL_GEOMETRY_TYPE_PATH)). */
BLAmdResult bl_amd_context_fill_path_d(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                       const double* vertices, size_t n);
/* bl_context_fill_geometry(BL_GEOMETRY_TYPE_POLYGOND) (context.h:755). */
BLAmdResult bl_amd_context_fill_polygon_d(BLAmdContext* ctx, const double* vertices, size_t n);
/* bl_context_fill_rect_d (context.h:745). */
BLAmdResult bl_amd_context_fill_rect_d(BLAmdContext* ctx, double x, double y, double w, double h);
/* bl_context_fill_rect_i (context.h:740). */
BLAmdResult bl_amd_context_fill_rect_i(BLAmdContext* ctx, int x, int y, int w, int h);
/* bl_context_fill_all (context.h:735). */
BLAmdResult bl_amd_context_fill_all(BLAmdContext* ctx);

/* Explicit-colour fills: the colour replaces the fill style for this call
 * only (BL_CONTEXT_RESOLVE_EXPLICIT_SOLID_OP, rastercontext.cpp:1258-1275). */
/* bl_context_fill_all_rgba32 (context.h:736). */
BLAmdResult bl_amd_context_fill_all_rgba32(BLAmdContext* ctx, uint32_t rgba32);
/* bl_context_fill_rect_i_rgba32 (context.h:741). */
BLAmdResult bl_amd_context_fill_rect_i_rgba32(BLAmdContext* ctx, int x, int y, int w, int h, uint32_t rgba32);
/* bl_context_fill_rect_d_rgba32 (context.h:746). */
BLAmdResult bl_amd_context_fill_rect_d_rgba32(BLAmdContext* ctx, double x, double y, double w, double h, uint32_t rgba32);
/* bl_context_fill_path_d_rgba32 (context.h:751). */
BLAmdResult bl_amd_context_fill_path_d_rgba32(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                              const double* vertices, size_t n, uint32_t rgba32);

/* _rgba64 forms (context.h:737, 742, 747, 752, 757): the colour (16 bits
 * per channel, straight alpha) replaces the fill style for this call; its
 * high bytes make the rgba32 colour, as the reference's conversion does. */
BLAmdResult bl_amd_context_fill_all_rgba64(BLAmdContext* ctx, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_rect_i_rgba64(BLAmdContext* ctx, int x, int y, int w, int h, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_rect_d_rgba64(BLAmdContext* ctx, double x, double y, double w, double h, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_path_d_rgba64(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                              const double* vertices, size_t n, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_geometry_rgba64(BLAmdContext* ctx, uint32_t type, const void* data, uint64_t rgba64);

/* A style for one call (the BLVar that the _ext entry points of context.h:
 * 738-758 take): a colour, or a gradient / pattern described as
 * bl_amd_context_set_fill_style_gradient / _pattern receive them. */
enum { BL_AMD_STYLE_RGBA32 = 0, BL_AMD_STYLE_RGBA64 = 1, BL_AMD_STYLE_GRADIENT = 2, BL_AMD_STYLE_PATTERN = 3 };
typedef struct BLAmdStyle {
  uint32_t kind;                 /* BL_AMD_STYLE_* */
  uint32_t rgba32;               /* RGBA32 */
  uint64_t rgba64;               /* RGBA64 */
  uint32_t gradient_type;        /* GRADIENT: 0 linear, 1 radial, 2 conic */
  uint32_t extend;               /* GRADIENT, PATTERN: 0 pad, 1 repeat, 2 reflect */
  const double* values;          /* GRADIENT: gradient values */
  const double* stop_offsets;    /* GRADIENT: n_stops offsets */
  const uint32_t* stop_rgba32;   /* GRADIENT: n_stops straight-alpha colours */
  size_t n_stops;
  const double* transform;       /* GRADIENT, PATTERN: own transform [6] or NULL */
  const BLAmdImage* image;       /* PATTERN */
} BLAmdStyle;

/* _ext forms (context.h:738, 743, 748, 753, 758): `style` replaces the fill
 * style for this call only and is set up with the user transform of the
 * moment (resolve_explicit_style_op, rastercontext.cpp:1277-1328); the
 * context's own style is not changed. */
BLAmdResult bl_amd_context_fill_all_ext(BLAmdContext* ctx, const BLAmdStyle* style);
BLAmdResult bl_amd_context_fill_rect_i_ext(BLAmdContext* ctx, int x, int y, int w, int h, const BLAmdStyle* style);
BLAmdResult bl_amd_context_fill_rect_d_ext(BLAmdContext* ctx, double x, double y, double w, double h, const BLAmdStyle* style);
BLAmdResult bl_amd_context_fill_path_d_ext(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                           const double* vertices, size_t n, const BLAmdStyle* style);
BLAmdResult bl_amd_context_fill_geometry_ext(BLAmdContext* ctx, uint32_t type, const void* data, const BLAmdStyle* style);

/* bl_context_fill_geometry (context.h:755) / _rgba32 (context.h:756).
 * `data` points to the reference's geometry struct for `type` (same memory
 * layout: BLBoxI {int x0,y0,x1,y1}, BLBox {double ...}, BLRectI {x,y,w,h},
 * BLRect, BLCircle {cx,cy,r}, BLEllipse {cx,cy,rx,ry}, BLRoundRect
 * {x,y,w,h,rx,ry}, BLLine {x0,y0,x1,y1}, BLTriangle {x0,y0,x1,y1,x2,y2},
 * BLArrayView {data, size} of BLPointI / BLPoint / boxes / rects).  ARC,
 * CHORD and PIE report BL_ERROR_NOT_IMPLEMENTED; paths go through
 * bl_amd_context_fill_path_d. */
BLAmdResult bl_amd_context_fill_geometry(BLAmdContext* ctx, uint32_t type, const void* data);
BLAmdResult bl_amd_context_fill_geometry_rgba32(BLAmdContext* ctx, uint32_t type, const void* data, uint32_t rgba32);

/* bl_context_clear_all (context.h:731), clear_rect_i (:732), clear_rect_d
 * (:733): SRC_COPY of transparent (opaque black on XRGB32) at full alpha. */
BLAmdResult bl_amd_context_clear_all(BLAmdContext* ctx);
BLAmdResult bl_amd_context_clear_rect_i(BLAmdContext* ctx, int x, int y, int w, int h);
BLAmdResult bl_amd_context_clear_rect_d(BLAmdContext* ctx, double x, double y, double w, double h);

/* bl_context_blit_image_i / _d (context.h:870-871, raster/rastercontext.cpp:
 * 3940-4066): composites `area` (x, y, w, h of the image; NULL: all of it)
 * of `image` with its top-left corner at (x, y), through the user
 * transform; the comp op and global alpha apply, the fill style and fill
 * alpha do not.  Fractional and transformed positions sample the area with
 * the pattern quality hint, extend REFLECT.  An area outside the image is
 * BL_AMD_ERROR_INVALID_VALUE.  As for pattern styles and masks, `image` must
 * not be the context's target: tiles are composited in parallel, so reading
 * pixels the same frame writes gives an undefined result. */
BLAmdResult bl_amd_context_blit_image_i(BLAmdContext* ctx, int x, int y, const BLAmdImage* image, const int32_t* area);
BLAmdResult bl_amd_context_blit_image_d(BLAmdContext* ctx, double x, double y, const BLAmdImage* image,
                                        const int32_t* area);
/* bl_context_blit_scaled_image_i / _d (context.h:872-873, rastercontext.cpp:
 * 4072-4150): the area scaled onto the rectangle (x, y, w, h). */
BLAmdResult bl_amd_context_blit_scaled_image_i(BLAmdContext* ctx, int x, int y, int w, int h, const BLAmdImage* image,
                                               const int32_t* area);
BLAmdResult bl_amd_context_blit_scaled_image_d(BLAmdContext* ctx, double x, double y, double w, double h,
                                               const BLAmdImage* image, const int32_t* area);

/* bl_context_fill_mask_i / _d and their _rgba32 / _rgba64 / _ext forms
 * (context.h:800-808, raster/rastercontext.cpp:3590-3780): fills the
 * pixels of `area` (x, y, w, h of `mask`; NULL: all of it) placed with its
 * top-left corner at (x, y) with the fill style (or the call's own colour /
 * style), each pixel's coverage being byte x of the mask area's row (an A8
 * mask's value; for a 32-bit mask the reference reads its bytes as an A8
 * row too), times the fill alpha.  As in the reference, only pixel-aligned
 * placements under a translation render: a fractional start or any other
 * transform is BL_AMD_ERROR_NOT_IMPLEMENTED (a fill that renders nothing
 * anyway returns success), an area outside the mask
 * BL_AMD_ERROR_INVALID_VALUE. */
BLAmdResult bl_amd_context_fill_mask_i(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask, const int32_t* area);
BLAmdResult bl_amd_context_fill_mask_d(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                       const int32_t* area);
BLAmdResult bl_amd_context_fill_mask_i_rgba32(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                              const int32_t* area, uint32_t rgba32);
BLAmdResult bl_amd_context_fill_mask_d_rgba32(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint32_t rgba32);
BLAmdResult bl_amd_context_fill_mask_i_rgba64(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                              const int32_t* area, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_mask_d_rgba64(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint64_t rgba64);
BLAmdResult bl_amd_context_fill_mask_i_ext(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                           const int32_t* area, const BLAmdStyle* style);
BLAmdResult bl_amd_context_fill_mask_d_ext(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                           const int32_t* area, const BLAmdStyle* style);

/* ---- statistics and replay (no counterpart in the reference) ----------- */

/* Counters of the last flush or replay. */
typedef struct BLAmdStats {
  uint64_t commands;   /* fill commands rendered */
  uint64_t sources;    /* source segments (lines, quads, cubics) */
  uint64_t segments;   /* 24.8 edge segments after clip + flatten */
  uint64_t binned;     /* (segment, band) pairs */
  uint64_t tiles;      /* 64x16 tiles in the render grid */
  uint64_t pixels;     /* pixels composited with a non-zero mask */
  uint64_t frame_bytes;  /* bytes of the device-resident frame (commands, geometry, styles, LUTs) */
  uint64_t lut_bytes;    /* gradient LUT bytes of the frame */
  uint64_t image_bytes;  /* pixel bytes of the pattern images the frame samples */
  uint64_t pred_segs;    /* learned sizing: predicted segments (0 otherwise) */
  uint64_t pred_bins;    /* learned sizing: predicted (segment, band) copies */
  uint32_t sizing;       /* BL_AMD_SIZING_* the frame used */
  uint32_t fallback;     /* 1: a pool overflowed and the frame was redone with exact counting */
  uint32_t syncs;        /* host round trips of the frame */
  uint32_t render_kernels; /* render kernels the frame launched: 1 k_render (bands with analytic
                              fills), 2 k_render_box_solid (solid box fills only), 4 k_render_box,
                              8 k_render_box_fast (solid SRC_OVER / SRC_COPY aligned boxes only) */
  double render_ms;    /* device time of the render stage (k_render, k_render_box_solid, k_render_box) */
  double total_ms;     /* host wall time of the device stages */
  double merge_ms;     /* flush: host time merging the batch's recorded frames (0 for one context) */
  double upload_ms;    /* flush: host time packing and uploading the frame (0 for a replay) */
  double pack_ms;      /* flush: the part of upload_ms spent packing the frame into staging memory */
  uint32_t schedule;   /* BL_AMD_SCHEDULE_* the analytic tiles ran with (AUTO resolved) */
  uint32_t queue_wgs;  /* queue schedules: workgroups per CU of the persistent grid (0 for GRID) */
  uint32_t graph;      /* 1: a replay launched as a captured HIP graph (render_ms then covers every device stage) */
  uint32_t reserved;
} BLAmdStats;

BLAmdResult bl_amd_context_get_stats(const BLAmdContext* ctx, BLAmdStats* out);

/* Executes the last flushed frame again from its device-resident copy
 * (commands, geometry, styles), with no recording or upload: a retained
 * display list.  The target and every pattern image the frame used must
 * still be alive; the frame is replaced by the next flush that renders. */
BLAmdResult bl_amd_context_replay(BLAmdContext* ctx);

/* No counterpart in the reference.  With graph replay on (process-wide,
 * default off), bl_amd_context_replay and bl_amd_batch_replay run the
 * resident frame's device stages as one captured HIP graph: flushes then
 * size exactly and record the frame's counts (segments, binned segments,
 * list entries, tiles and tile groups), the first replay captures every
 * launch with those counts in place of the read-backs, and it and later
 * replays launch the captured graph (one launch, one synchronization).  A new
 * upload, a different schedule or tile model, or a grown buffer captures it
 * again.  Pixels are those of the eager replay. */
BLAmdResult bl_amd_set_graph_replay(uint32_t on);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bl_amd_context_replay with exact sizing, whose
 * intermediate buffers (edge work items and segments, binned segments,
 * band matrices and lists, tile headers, counters) are slices of one
 * 64 MiB device region that the library holds per device for the life of
 * the process; it calls no hipMalloc / hipFree.  A frame whose buffers do
 * not fit fails with BL_AMD_ERROR_DATA_TOO_LARGE before anything is
 * composited.  Every call on a device shares the region and leaves it dirty
 * (the call clears what must start at zero); calls that overlap on one
 * device race. */
BLAmdResult bl_amd_context_replay_noalloc(BLAmdContext* ctx);

/* ---- batching (no counterpart in the reference) ------------------------
 * Renders the pending calls of many contexts (canvases of any size and
 * format) with one set of kernel launches: one edge pass, one binning pass
 * and one render grid covering every tile of every canvas. */
BLAmdResult bl_amd_batch_create(BLAmdBatch** out);
BLAmdResult bl_amd_batch_destroy(BLAmdBatch* batch);
/* A context can be in a batch once, and no two contexts of a batch may
 * render into the same image (their canvases are composited concurrently):
 * such an add fails with BL_AMD_ERROR_INVALID_VALUE. */
BLAmdResult bl_amd_batch_add(BLAmdBatch* batch, BLAmdContext* ctx);
BLAmdResult bl_amd_batch_clear(BLAmdBatch* batch);
BLAmdResult bl_amd_batch_flush(BLAmdBatch* batch);
/* Same as bl_amd_context_replay for the batch's last merged frame. */
BLAmdResult bl_amd_batch_replay(BLAmdBatch* batch);
BLAmdResult bl_amd_batch_get_stats(const BLAmdBatch* batch, BLAmdStats* out);

/* ---- learned sizing (no counterpart in the reference) --------------------
 * A frame must size its edge-segment and bin buffers before the kernels that
 * fill them run.  BL_AMD_SIZING_EXACT runs a counting pass of the edge
 * builder and reads the sizes back; BL_AMD_SIZING_LEARNED predicts them with
 * a small MLP (16 per-command features -> 64 -> 64 -> 2, fp32, trained on
 * the GPU by the kernels below) and runs the edge builder in one kernel
 * with no host round trip (each work item counted, then written at a slot
 * range taken per wave);
 * BL_AMD_SIZING_HISTORY reuses the totals of the last exact frame.  In both
 * predicted modes an overflow is detected on the device, nothing is
 * composited, and the frame is redone exactly: pixels never depend on the
 * model. */
enum { BL_AMD_SIZING_EXACT = 0, BL_AMD_SIZING_LEARNED = 1, BL_AMD_SIZING_HISTORY = 2 };
enum { BL_AMD_COST_FEATURES = 16, BL_AMD_COST_OUTPUTS = 2, BL_AMD_COST_PARAMS = 5378 };

/* Process-wide sizing mode (default EXACT). */
BLAmdResult bl_amd_set_sizing(uint32_t mode);
/* Installs the model parameters (BL_AMD_COST_PARAMS floats: W1[16][64] b1[64]
 * W2[64][64] b2[64] W3[64][2] b3[2]); NULL removes it (LEARNED then counts
 * exactly). */
BLAmdResult bl_amd_cost_model_set(const float* params);
/* Re-executes the context's resident frame with exact counting and returns,
 * per fill command with source geometry, its 16 features and the targets
 * log2(1 + segments), log2(1 + (segment, band) copies).  `n_out` receives
 * the number of rows (rows beyond `cap` are not written). */
BLAmdResult bl_amd_context_collect_costs(BLAmdContext* ctx, float* features, float* targets, size_t cap, size_t* n_out);
/* MLP forward pass over n rows on the GPU. */
BLAmdResult bl_amd_cost_model_forward(const float* features, size_t n, const float* params, float* outputs);
/* Gradient of the mean squared error over n rows (and the loss) on the GPU. */
BLAmdResult bl_amd_cost_model_grad(const float* features, const float* targets, size_t n, const float* params,
                                   float* grad, float* loss);
/* Adam training on the GPU, mini-batches of `batch` rows in the given row
 * order; `params` holds the initial parameters and receives the trained
 * ones; `epoch_loss` (optional, `epochs` floats) the mean loss per epoch. */
BLAmdResult bl_amd_cost_model_train(const float* features, const float* targets, size_t n, float* params,
                                    uint32_t epochs, uint32_t batch, float lr, float* epoch_loss);

/* ---- tile scheduling and autotuning (no counterpart in the reference) ----
 * The tiles of bands with path fills (the heavy part of the render stage)
 * run in one of these schedules; none changes pixels (tiles own disjoint
 * pixels):
 *   GRID             one workgroup per tile, tiles dealt to the XCDs in
 *                    contiguous ranges;
 *   QUEUE            a persistent grid (wgs_per_cu workgroups per CU)
 *                    draining a device work queue of the tiles in band order;
 *   QUEUE_HEURISTIC  the queue sorted longest first by each tile's binned
 *                    segments + 16 x commands;
 *   QUEUE_LEARNED    the queue sorted longest first by the tile cost model:
 *                    an MLP of the sizing model's shape (16 tile features ->
 *                    64 -> 64 -> 2; outputs log2(1 + wall-clock ticks of the
 *                    tile), log2(1 + composited pixels)) trained on tile
 *                    timings logged by the render kernel (without a model it
 *                    runs as QUEUE_HEURISTIC);
 *   AUTO (default)   the schedule the autotuner picked for frames of this
 *                    shape (tile, command and canvas counts, each by its
 *                    power of two), else GRID.
 * bl_amd_context_replay_noalloc always runs GRID. */
enum { BL_AMD_SCHEDULE_GRID = 0, BL_AMD_SCHEDULE_QUEUE = 1, BL_AMD_SCHEDULE_QUEUE_HEURISTIC = 2,
       BL_AMD_SCHEDULE_QUEUE_LEARNED = 3, BL_AMD_SCHEDULE_AUTO = 4 };
enum { BL_AMD_TILE_FEATURES = 16, BL_AMD_TILE_OUTPUTS = 2 };

/* Process-wide schedule (default AUTO); wgs_per_cu 1..4 for the queue
 * schedules (0: 4). */
BLAmdResult bl_amd_set_render_schedule(uint32_t schedule, uint32_t wgs_per_cu);
/* Installs the tile cost model (BL_AMD_COST_PARAMS floats, the layout of
 * bl_amd_cost_model_set); NULL removes it. */
BLAmdResult bl_amd_tile_model_set(const float* params);
/* Re-executes the resident frame with exact sizing on the QUEUE schedule
 * with tile timing and returns, per analytic tile, its 16 features and the
 * targets log2(1 + wall-clock ticks at 100 MHz), log2(1 + composited
 * pixels).  `n_out` receives the number of rows (rows beyond `cap` are not
 * written). */
BLAmdResult bl_amd_context_collect_tile_costs(BLAmdContext* ctx, float* features, float* targets, size_t cap,
                                              size_t* n_out);
BLAmdResult bl_amd_batch_collect_tile_costs(BLAmdBatch* batch, float* features, float* targets, size_t cap,
                                            size_t* n_out);

typedef struct BLAmdTune {
  uint32_t schedule;    /* the fastest schedule (BL_AMD_SCHEDULE_*) */
  uint32_t wgs_per_cu;  /* its workgroups per CU (queue schedules) */
  uint32_t candidates;  /* schedules timed */
  uint32_t reserved;
  double best_ms;       /* its render-stage time, best of the replays */
  double grid_ms;       /* the GRID schedule's, for comparison */
} BLAmdTune;

/* Autotuner: times the render stage of the resident frame under every
 * schedule (GRID; QUEUE, QUEUE_HEURISTIC and, with a tile model,
 * QUEUE_LEARNED at 2, 3 and 4 workgroups per CU), best of `reps` replays
 * each, records the fastest for frames of this shape (what AUTO then runs)
 * and returns it. */
BLAmdResult bl_amd_context_autotune(BLAmdContext* ctx, uint32_t reps, BLAmdTune* out);
BLAmdResult bl_amd_batch_autotune(BLAmdBatch* batch, uint32_t reps, BLAmdTune* out);
/* The autotuner's record, one line per tuned frame shape with every
 * candidate's time; copies up to cap - 1 bytes and a terminating zero,
 * returns the full length. */
size_t bl_amd_autotune_log(char* buf, size_t cap);
/* Forgets every tuned shape. */
BLAmdResult bl_amd_autotune_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* BLEND2D_AMD_H */
