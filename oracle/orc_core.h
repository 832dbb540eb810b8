/* CPU oracle for the Blend2D fill path -- shared internal definitions.
 *
 * TEST INFRASTRUCTURE.  The oracle is a plain-C restatement of the reference's
 * non-JIT fill path (edge builder, analytic rasterizer, fill loops, fetchers,
 * comp ops, style setup).  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker.  The MI355X product
 * library never links or calls it.
 *
 * Citations are relative to the reference repository root.
 */
#ifndef ORC_CORE_H
#define ORC_CORE_H

#include <stddef.h>
#include <stdint.h>

typedef struct { double x, y; } OrcPt;
typedef struct { double m00, m01, m10, m11, m20, m21; } OrcMat;
typedef struct { double x0, y0, x1, y1; } OrcBox;

/* BLTransformType (core/matrix.h). */
enum { ORC_TT_IDENTITY = 0, ORC_TT_TRANSLATE = 1, ORC_TT_SCALE = 2, ORC_TT_SWAP = 3, ORC_TT_AFFINE = 4, ORC_TT_INVALID = 5 };

/* FormatExt (core/format_p.h): PRGB32=1, XRGB32=2, A8=3, FRGB32=4, ZERO32=5. */
enum { ORC_FMT_NONE = 0, ORC_FMT_PRGB32 = 1, ORC_FMT_XRGB32 = 2, ORC_FMT_A8 = 3, ORC_FMT_FRGB32 = 4, ORC_FMT_ZERO32 = 5 };

/* BLCompOp values used by the scene (core/context.h) plus the internal
 * results of comp-op simplification (core/compop_p.h CompOpExt). */
enum {
  ORC_OP_SRC_OVER = 0, ORC_OP_SRC_COPY = 1, ORC_OP_SRC_IN = 2, ORC_OP_SRC_OUT = 3, ORC_OP_SRC_ATOP = 4,
  ORC_OP_DST_OVER = 5, ORC_OP_DST_COPY = 6, ORC_OP_DST_IN = 7, ORC_OP_DST_OUT = 8, ORC_OP_DST_ATOP = 9,
  ORC_OP_XOR = 10, ORC_OP_CLEAR = 11, ORC_OP_PLUS = 12, ORC_OP_MINUS = 13, ORC_OP_MODULATE = 14,
  ORC_OP_MULTIPLY = 15, ORC_OP_SCREEN = 16,
  ORC_OP_NOP = 255
};

/* CompOpSolidId (core/compopinfo_p.h). */
enum { ORC_SOLID_NONE = 0, ORC_SOLID_TRANSPARENT = 1, ORC_SOLID_OPAQUE_BLACK = 2, ORC_SOLID_OPAQUE_WHITE = 3 };

/* One emitted edge segment in 24.8 fixed point with y0 <= y1 and the sign of
 * the original direction (1 = the source went upwards), the unit the analytic
 * rasterizer consumes (raster/edgestorage_p.h EdgeVector point pairs). */
typedef struct { int32_t x0, y0, x1, y1; uint32_t sign; } OrcSeg;

typedef struct {
  OrcSeg* segs;
  size_t n, cap;
  int32_t bbox_y0, bbox_y1;   /* fixed-point y range of non-horizontal segments */
} OrcEdges;

/* Resolved fetch description (pipeline/pipedefs_p.h FetchData). */
enum { ORC_FETCH_SOLID = 0, ORC_FETCH_LINEAR = 1, ORC_FETCH_RADIAL = 2, ORC_FETCH_CONIC = 3,
       ORC_FETCH_PATTERN_AFFINE_NN = 4, ORC_FETCH_PATTERN_AFFINE_BI = 5,
       ORC_FETCH_PATTERN_ALIGNED = 6, ORC_FETCH_PATTERN_FXFY = 7 };

typedef struct {
  uint32_t type;
  uint32_t src_format;        /* FormatExt of the source as seen by comp-op simplification */
  uint32_t solid_prgb32;
  /* gradient */
  const uint32_t* lut;
  const uint64_t* lut64;      /* dithered gradients: the 64-bit table (else NULL) */
  uint32_t lut_size;
  uint32_t extend;            /* 0 pad, 1 repeat, 2 reflect */
  uint32_t maxi, rori;
  uint64_t lin_pt0, lin_dt, lin_dy;
  double rad_tx, rad_ty, rad_yx, rad_yy, rad_b0, rad_by, rad_dd0, rad_ddy;
  double rad_inv2a, rad_amul4, rad_sq_inv2a, rad_sq_fr;
  float rad_bd, rad_ddd;
  double con_tx, con_ty, con_yx, con_yy;
  float con_q[4], con_n1, con_n2, con_n4, con_offset, con_xx;
  int32_t con_maxi, con_rori;
  /* pattern */
  const uint8_t* img;
  intptr_t img_stride;
  uint32_t img_w, img_h, img_format;
  uint64_t pa_xx, pa_xy, pa_yx, pa_yy, pa_tx, pa_ty;
  int32_t pa_ox, pa_oy, pa_rx, pa_ry, pa_minx, pa_miny, pa_maxx, pa_maxy, pa_corx, pa_cory, pa_tw, pa_th;
  /* aligned / fractional-translation pattern (pipedefs.cpp:60-160): texel
   * offset, repeat-or-reflect periods (0 = pad) and bilinear weights */
  int32_t sp_tx, sp_ty, sp_rx, sp_ry;
  uint32_t sp_wa, sp_wb, sp_wc, sp_wd;
} OrcFetch;

/* ---- numeric helpers (support/math_p.h, core/api.h) ---- */
static inline double orc_min(double a, double b) { return b < a ? b : a; }
static inline double orc_max(double a, double b) { return a < b ? b : a; }
static inline double orc_clamp(double v, double lo, double hi) { return orc_min(hi, orc_max(lo, v)); }
static inline int orc_trunc(double x) { return (int)x; }

int orc_nearby_to_int(double x);
int orc_round_to_int(double x);
int64_t orc_floor_to_int64(double x);

/* ---- matrix (core/matrix.cpp, core/matrix_p.h) ---- */
int orc_mat_type(const OrcMat* m);
int orc_mat_invert(OrcMat* dst, const OrcMat* src);
void orc_mat_multiply(OrcMat* dst, const OrcMat* a, const OrcMat* b);
OrcPt orc_mat_map(const OrcMat* m, double x, double y);

/* ---- edges (orc_edges.c) ---- */
void orc_edges_init(OrcEdges* e);
void orc_edges_free(OrcEdges* e);
void orc_edges_clear(OrcEdges* e);
/* BL_ERROR_INVALID_GEOMETRY (core/api.h), returned by the two functions
 * below when the line clipper meets NaN (raster/edgebuilder_p.h:1494-1496,
 * 1584-1586); the caller then drops every edge of the call. */
#define ORC_ERR_INVALID_GEOMETRY 0x10034
/* Adds the edges of a path (BLPath command bytes + vertices), closed as for
 * filling (raster/rastercontextops.cpp add_filled_path_edges). */
int orc_edges_add_path(OrcEdges* e, const uint8_t* cmds, const OrcPt* vtx, size_t n,
                       const OrcMat* m, int tt, const OrcBox* clip_fixed, double tol_sq);
int orc_edges_add_poly(OrcEdges* e, const OrcPt* vtx, size_t n,
                       const OrcMat* m, int tt, const OrcBox* clip_fixed, double tol_sq);

/* ---- styles (orc_style.c) ---- */
typedef struct { double offset; uint64_t rgba64; } OrcStop;
uint32_t orc_premultiply(uint32_t argb32);
void orc_interpolate_lut(uint32_t* dst, uint32_t dst_size, const OrcStop* stops, size_t n);
void orc_interpolate_lut64(uint64_t* dst, uint32_t dst_size, const OrcStop* stops, size_t n);
/* Returns lut size and sets solid/format as ensure_info (core/gradient.cpp:202-287). */
uint32_t orc_gradient_info(const OrcStop* stops, size_t n, int* solid, uint32_t* format);
int orc_init_linear(OrcFetch* f, const double v[4], uint32_t extend, const OrcMat* t);
int orc_init_radial(OrcFetch* f, const double v[6], uint32_t extend, const OrcMat* t);
int orc_init_conic(OrcFetch* f, const double v[4], const OrcMat* t);
int orc_init_pattern(OrcFetch* f, uint32_t extend, uint32_t quality, const OrcMat* t);
/* comp-op simplification: returns executed op (or ORC_OP_NOP), sets dst/src
 * pipeline formats and solid override id. */
uint32_t orc_simplify(uint32_t op, uint32_t dst_fmt, uint32_t src_fmt, uint32_t* out_dst, uint32_t* out_src, uint32_t* solid_id);

/* ---- pipeline (orc_pipe.c) ---- */
typedef struct {
  uint8_t* pixels;
  intptr_t stride;
  uint32_t w, h, format;      /* destination format: PRGB32 / XRGB32 / A8 */
} OrcCanvas;

typedef struct {
  const OrcFetch* fetch;
  uint32_t op;                /* executed op after simplification */
  uint32_t dst_fmt, src_fmt;  /* pipeline formats after simplification */
  /* per-fill-call state */
  uint32_t y_start;           /* row the fill call started at (spanInitY / rect init) */
  double con_tp_x, con_tp_y;  /* conic per-row origin, accumulated from y_start */
  int32_t con_row;
} OrcPipe;

void orc_pipe_begin(OrcPipe* p, uint32_t y0);
void orc_pipe_row(OrcPipe* p, uint32_t y);
void orc_pipe_composite(OrcPipe* p, OrcCanvas* c, uint32_t x, uint32_t y, uint32_t m);
uint32_t orc_div255(uint32_t x);

/* ---- raster (orc_raster.c) ---- */
uint32_t orc_band_height(uint32_t width);
void orc_fill_analytic(OrcCanvas* c, OrcEdges* e, OrcPipe* p, uint32_t alpha, uint32_t fill_rule);
void orc_fill_box_a(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1);
void orc_fill_box_u(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1);
void orc_fill_box_masked(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1,
                         const uint8_t* mask, intptr_t mask_stride);

#endif /* ORC_CORE_H */
