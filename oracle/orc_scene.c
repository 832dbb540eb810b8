/* CPU oracle -- scene replay (the BLContext state machine of the sync path).
 *
 * TEST INFRASTRUCTURE (see orc_core.h).  Restates the frontend decisions of
 * raster/rastercontext.cpp for the fill path:
 *   resolve_*_style_op / formatFromRgba32     (:90-93, :1231-1342)
 *   init_non_solid_fetch_data                  (:363-494)
 *   fill_unclipped_path / polygon / box / rect (:2787-2917)
 *   set_global_alpha                           (:2000-2029)
 *   defaults after attach                      (:4383-4411)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_core.h"
#include "scene_format.h"

#define ERR_INVALID_VALUE 0x10001u
#define ERR_NOT_IMPLEMENTED 0x10007u

typedef struct { const uint8_t* p; const uint8_t* end; } Rd;

static int rd(Rd* r, void* out, size_t n) {
  if ((size_t)(r->end - r->p) < n) return 0;
  memcpy(out, r->p, n);
  r->p += n;
  return 1;
}
static size_t pad8(size_t n) { return (n + 7u) & ~(size_t)7u; }

typedef struct {
  uint32_t type;      /* 0 solid, 1 gradient, 2 pattern, 3 disabled (NaN colour), 4 premultiplied solid */
  uint32_t rgba32;
  SceneGradient g;
  OrcStop stops[64];
  size_t n_stops;
  ScenePattern pat;
  OrcMat xform;       /* user transform captured by set_fill_style (rastercontext.cpp:653-695) */
  uint32_t quality;   /* pattern quality hint captured by set_fill_style */
  uint32_t dither;    /* gradients: the gradient quality hint was DITHER at set_fill_style */
  uint32_t has_view;  /* blits: the pattern samples view = x, y, w, h of the image */
  int32_t view[4];
} Style;

typedef struct {
  OrcCanvas canvas;
  uint32_t comp_op, fill_rule, pattern_quality, gradient_quality;
  OrcMat user;
  double global_alpha;
  double fill_alpha;
  Style style;
  Style explicit_style;   /* mode 3: the style of an explicit-style call */
  uint32_t mode;          /* 0 fill with the style, 1 explicit colour, 2 clear, 3 explicit style, 4 blit */
  uint32_t explicit_rgba32;
  uint32_t call_all;      /* fill_all / clear_all: an invalid transform is no NOP */
  SceneImage img_info[16];
  const uint8_t* img_data[16];
  uint32_t n_images;
  OrcEdges edges;
  uint32_t first_error;
} Ctx;

static void note(Ctx* c, uint32_t err) { if (err && !c->first_error) c->first_error = err; }

/* styleAlphaI of the fill slot (rastercontext.cpp:1920-1939, :2000-2029);
 * clear calls use the full alpha (rastercontext.cpp:1385-1391). */
static uint32_t alpha_i(const Ctx* c) {
  if (c->mode == 2) return 255u;
  /* blits: global alpha alone (BL_CONTEXT_RESOLVE_BLIT_OP, rastercontext.cpp:1445-1453) */
  if (c->mode == 4) return (uint32_t)orc_round_to_int(c->global_alpha * 255.0);
  return (uint32_t)orc_round_to_int(c->global_alpha * 255.0 * c->fill_alpha);
}

/* core/gradient.cpp:751-800 (add_stop keeps stops sorted; a third stop at an
 * existing offset replaces the second one). */
static void add_stop(Style* s, double offset, uint32_t rgba32) {
  uint64_t v = 0;
  for (int i = 0; i < 4; i++) {
    uint64_t c = (rgba32 >> (i * 8)) & 0xFFu;
    v |= (c * 0x101u) << (i * 16);
  }
  size_t n = s->n_stops, i = 0;
  if (n && offset >= s->stops[0].offset) {
    /* last index with offset <= new offset */
    size_t k = 0;
    while (k + 1 < n && s->stops[k + 1].offset <= offset) k++;
    i = k;
    if (i > 0 && s->stops[i - 1].offset == offset) {
      s->stops[i].offset = offset;
      s->stops[i].rgba64 = v;
      return;
    }
    i++;
  }
  if (n >= 64) return;
  memmove(&s->stops[i + 1], &s->stops[i], (n - i) * sizeof(OrcStop));
  s->stops[i].offset = offset;
  s->stops[i].rgba64 = v;
  s->n_stops = n + 1;
}

/* set_style_impl for a gradient / pattern (rastercontext.cpp:653-695): the
 * style captures the user transform (and pattern quality) of the moment. */
static size_t parse_gradient_style(Style* st, const uint8_t* pl, const OrcMat* user, uint32_t quality) {
  memset(st, 0, sizeof(*st));
  st->type = 1;
  st->xform = *user;
  st->dither = quality >= 2u;   /* BL_GRADIENT_QUALITY_DITHER */
  memcpy(&st->g, pl, sizeof(SceneGradient));
  for (uint32_t k = 0; k < st->g.n_stops; k++) {
    SceneStop s;
    memcpy(&s, pl + sizeof(SceneGradient) + k * sizeof(SceneStop), sizeof(s));
    if (!(s.offset >= 0.0 && s.offset <= 1.0)) continue;
    add_stop(st, s.offset, s.rgba32);
  }
  return sizeof(SceneGradient) + st->g.n_stops * sizeof(SceneStop);
}

static size_t parse_pattern_style(Style* st, const uint8_t* pl, const OrcMat* user, uint32_t quality) {
  memset(st, 0, sizeof(*st));
  st->type = 2;
  st->xform = *user;
  st->quality = quality;
  memcpy(&st->pat, pl, sizeof(ScenePattern));
  return sizeof(ScenePattern);
}

/* rgba32FromRgba64 (core/rgba_p.h:67-74): the high byte of each channel */
static uint32_t rgba32_from_rgba64(uint64_t v) {
  return (uint32_t)(((v >> 56) & 0xFFu) << 24 | ((v >> 40) & 0xFFu) << 16 | ((v >> 24) & 0xFFu) << 8 | ((v >> 8) & 0xFFu));
}

/* Resolves style + comp op into a pipeline (fetch + op).  Returns 0 when the
 * call renders, 1 when it is a NOP, or an error code. */
static uint32_t resolve(Ctx* c, OrcFetch* f, uint32_t* lut_buf, OrcPipe* pipe) {
  memset(f, 0, sizeof(*f));
  memset(pipe, 0, sizeof(*pipe));
  uint32_t src_fmt = ORC_FMT_NONE;
  int solid = 0;
  uint32_t lut_size = 0;

  /* kNoUserTransform is part of every geometry call's NOP flags, not of
   * fill_all / clear_all (raster/rasterdefs_p.h:114-131) */
  if (!c->call_all && orc_mat_type(&c->user) == ORC_TT_INVALID) return 1;

  if (c->mode == 2) {
    /* resolve_clear_op (rastercontext.cpp:1232-1242) */
    uint32_t pd, ps, sid;
    uint32_t op = orc_simplify(ORC_OP_SRC_COPY, c->canvas.format, ORC_FMT_PRGB32, &pd, &ps, &sid);
    if (op == ORC_OP_NOP) return 1;
    if (op != ORC_OP_SRC_COPY) return ERR_NOT_IMPLEMENTED;
    pipe->op = op;
    pipe->dst_fmt = pd;
    pipe->src_fmt = ps;
    pipe->fetch = f;
    f->type = ORC_FETCH_SOLID;
    f->solid_prgb32 = sid == ORC_SOLID_OPAQUE_BLACK ? 0xFF000000u : sid == ORC_SOLID_OPAQUE_WHITE ? 0xFFFFFFFFu : 0u;
    return 0;
  }

  if (alpha_i(c) == 0) return 1;

  Style tmp;
  Style* st = &c->style;
  if (c->mode == 1) {
    memset(&tmp, 0, sizeof(tmp));
    tmp.type = 0;
    tmp.rgba32 = c->explicit_rgba32;
    st = &tmp;
  }
  else if (c->mode == 3 || c->mode == 4) {
    st = &c->explicit_style;
  }
  OrcMat ft = st->xform;  /* final_transform (user x identity meta) when the style was set */
  if (st->type == 3) return 1;
  if (st->type == 0 || st->type == 4) {
    uint32_t v = st->type == 0 ? orc_premultiply(st->rgba32) : st->rgba32;
    uint32_t fv = st->type == 0 ? st->rgba32 : v;
    src_fmt = fv == 0 ? ORC_FMT_ZERO32 : (fv >= 0xFF000000u ? ORC_FMT_FRGB32 : ORC_FMT_PRGB32);
    f->type = ORC_FETCH_SOLID;
    f->solid_prgb32 = v;
  }
  else if (st->type == 1) {
    if (st->n_stops == 0) return 1;
    lut_size = orc_gradient_info(st->stops, st->n_stops, &solid, &src_fmt);
    if (solid) {
      uint64_t last = st->stops[st->n_stops - 1].rgba64;
      uint32_t c32 = (uint32_t)(((last >> 56) & 0xFFu) << 24 | ((last >> 40) & 0xFFu) << 16 | ((last >> 24) & 0xFFu) << 8 | ((last >> 8) & 0xFFu));
      f->type = ORC_FETCH_SOLID;
      f->solid_prgb32 = orc_premultiply(c32);
    }
  }
  else {
    if (st->pat.image_index >= c->n_images) return ERR_INVALID_VALUE;
    src_fmt = c->img_info[st->pat.image_index].format;
  }

  uint32_t pd, ps, sid;
  uint32_t op = orc_simplify(c->comp_op, c->canvas.format, src_fmt, &pd, &ps, &sid);
  if (op == ORC_OP_NOP) return 1;
  if (!(op == ORC_OP_SRC_COPY || op == ORC_OP_SRC_OVER || op == ORC_OP_PLUS ||
        ((op == ORC_OP_MULTIPLY || op == ORC_OP_SCREEN) && pd == ORC_FMT_PRGB32)))
    return ERR_NOT_IMPLEMENTED;
  if (ps == ORC_FMT_A8 && pd != ORC_FMT_A8 && op != ORC_OP_SRC_COPY && op != ORC_OP_SRC_OVER && op != ORC_OP_PLUS)
    return ERR_NOT_IMPLEMENTED;

  pipe->op = op;
  pipe->dst_fmt = pd;
  pipe->src_fmt = ps;
  pipe->fetch = f;

  if (sid != ORC_SOLID_NONE) {
    f->type = ORC_FETCH_SOLID;
    f->solid_prgb32 = sid == ORC_SOLID_TRANSPARENT ? 0u : sid == ORC_SOLID_OPAQUE_BLACK ? 0xFF000000u : 0xFFFFFFFFu;
    return 0;
  }

  if (st->type == 1 && !solid) {
    OrcMat t = ft;
    if (st->g.has_transform) {
      OrcMat gm = { st->g.transform[0], st->g.transform[1], st->g.transform[2], st->g.transform[3], st->g.transform[4], st->g.transform[5] };
      if (orc_mat_type(&gm) != ORC_TT_IDENTITY) orc_mat_multiply(&t, &gm, &ft);
    }
    if (st->dither && c->canvas.format != ORC_FMT_A8) {
      /* a dithered gradient (not on A8 targets, rastercontext.cpp:456-467)
       * samples a 64-bit table of twice the size, at most 1024 entries */
      static uint64_t lut64[1024];
      lut_size = lut_size * 2u < 1024u ? lut_size * 2u : 1024u;
      orc_interpolate_lut64(lut64, lut_size, st->stops, st->n_stops);
      f->lut64 = lut64;
    }
    else {
      orc_interpolate_lut(lut_buf, lut_size, st->stops, st->n_stops);
    }
    f->lut = lut_buf;
    f->lut_size = lut_size;
    int r;
    if (st->g.type == 0) r = orc_init_linear(f, st->g.values, st->g.extend, &t);
    else if (st->g.type == 1) r = orc_init_radial(f, st->g.values, st->g.extend, &t);
    else r = orc_init_conic(f, st->g.values, &t);
    if (r != 0) return 1;
  }
  else if (st->type == 2) {
    OrcMat t = ft;
    if (st->pat.has_transform) {
      OrcMat pm = { st->pat.transform[0], st->pat.transform[1], st->pat.transform[2], st->pat.transform[3], st->pat.transform[4], st->pat.transform[5] };
      if (orc_mat_type(&pm) != ORC_TT_IDENTITY) orc_mat_multiply(&t, &pm, &ft);
    }
    const SceneImage* ii = &c->img_info[st->pat.image_index];
    f->img = c->img_data[st->pat.image_index];
    f->img_stride = (intptr_t)ii->stride;
    f->img_w = ii->width;
    f->img_h = ii->height;
    f->img_format = ii->format;
    if (st->has_view) {
      /* init_image_source (raster/renderfetchdata_p.h:146-157) */
      f->img += (size_t)st->view[1] * ii->stride + (size_t)st->view[0] * (ii->format == ORC_FMT_A8 ? 1u : 4u);
      f->img_w = (uint32_t)st->view[2];
      f->img_h = (uint32_t)st->view[3];
    }
    int r = orc_init_pattern(f, st->pat.extend, st->quality, &t);
    if (r < 0) return 1;
    if (r > 0) return ERR_NOT_IMPLEMENTED;
  }
  return 0;
}

static void fixed_transform(const Ctx* c, OrcMat* m, int* tt) {
  *m = c->user;
  m->m00 *= 256.0; m->m01 *= 256.0; m->m10 *= 256.0; m->m11 *= 256.0; m->m20 *= 256.0; m->m21 *= 256.0;
  int t = orc_mat_type(&c->user);
  *tt = t < ORC_TT_SCALE ? ORC_TT_SCALE : t;
}

static void clip_fixed(const Ctx* c, OrcBox* b) {
  b->x0 = 0.0; b->y0 = 0.0;
  b->x1 = (double)c->canvas.w * 256.0;
  b->y1 = (double)c->canvas.h * 256.0;
}

static const double kToleranceSq = (0.2 * 256.0) * (0.2 * 256.0);

static void fill_edges(Ctx* c, uint32_t fill_rule) {
  OrcFetch f;
  OrcPipe p;
  static uint32_t lut[2048];
  /* the pipeline (and its error) is looked up only for non-empty edges
   * (rastercontext.cpp:2734-2749) */
  if (c->edges.n == 0 || c->edges.bbox_y0 >= c->edges.bbox_y1) return;
  uint32_t r = resolve(c, &f, lut, &p);
  if (r > 1) { note(c, r); return; }
  if (r == 1) return;
  orc_fill_analytic(&c->canvas, &c->edges, &p, alpha_i(c), fill_rule);
}

/* A fill whose edges failed (INVALID_GEOMETRY): the reference resolves the
 * style and comp op first and returns success for a NOP call before it
 * builds any edge (BL_CONTEXT_RESOLVE_*, rastercontext.cpp:1358-1453); else
 * the error is the call's result and the edges are reverted
 * (rastercontextops.cpp:38-45). */
static void fail_edges(Ctx* c, uint32_t err) {
  OrcFetch f;
  OrcPipe p;
  static uint32_t lut[2048];
  orc_edges_clear(&c->edges);
  if (resolve(c, &f, lut, &p) != 1) note(c, err);
}

static void fill_path_origin(Ctx* c, const uint8_t* cmds, const OrcPt* vtx, size_t n, const double* origin) {
  OrcMat m; int tt; OrcBox clip;
  fixed_transform(c, &m, &tt);
  if (origin) {
    /* fill_unclipped_path_with_origin (rastercontext.cpp:2812-2821) */
    double ox = origin[0] * m.m00 + origin[1] * m.m10 + m.m20;
    double oy = origin[0] * m.m01 + origin[1] * m.m11 + m.m21;
    m.m20 = ox;
    m.m21 = oy;
  }
  clip_fixed(c, &clip);
  orc_edges_clear(&c->edges);
  int r = orc_edges_add_path(&c->edges, cmds, vtx, n, &m, tt, &clip, kToleranceSq);
  if (r) { fail_edges(c, (uint32_t)r); return; }
  fill_edges(c, c->fill_rule);
}

static void fill_path(Ctx* c, const uint8_t* cmds, const OrcPt* vtx, size_t n) { fill_path_origin(c, cmds, vtx, n, NULL); }

static void fill_poly(Ctx* c, const OrcPt* vtx, size_t n, uint32_t rule) {
  if (n < 3) return;
  OrcMat m; int tt; OrcBox clip;
  fixed_transform(c, &m, &tt);
  clip_fixed(c, &clip);
  orc_edges_clear(&c->edges);
  int r = orc_edges_add_poly(&c->edges, vtx, n, &m, tt, &clip, kToleranceSq);
  if (r) { fail_edges(c, (uint32_t)r); return; }
  fill_edges(c, rule);
}

static void fill_box_i(Ctx* c, int x0, int y0, int x1, int y1) {
  OrcFetch f;
  OrcPipe p;
  static uint32_t lut[2048];
  uint32_t r = resolve(c, &f, lut, &p);
  if (r > 1) { note(c, r); return; }
  if (r == 1) return;
  orc_fill_box_a(&c->canvas, &p, alpha_i(c), x0, y0, x1, y1);
}

/* fill_unclipped_box_d (rastercontext.cpp:2871-2891) */
static void fill_box_d(Ctx* c, double bx0, double by0, double bx1, double by1) {
  OrcMat m; int tt; OrcBox clip;
  fixed_transform(c, &m, &tt);
  clip_fixed(c, &clip);
  if (tt <= ORC_TT_SWAP) {
    double x0 = bx0 * m.m00 + by0 * m.m10 + m.m20;
    double y0 = bx0 * m.m01 + by0 * m.m11 + m.m21;
    double x1 = bx1 * m.m00 + by1 * m.m10 + m.m20;
    double y1 = bx1 * m.m01 + by1 * m.m11 + m.m21;
    OrcBox b = { orc_min(x0, x1), orc_min(y0, y1), orc_max(x0, x1), orc_max(y0, y1) };
    OrcBox d = { orc_max(b.x0, clip.x0), orc_max(b.y0, clip.y0), orc_min(b.x1, clip.x1), orc_min(b.y1, clip.y1) };
    if (!((d.x0 < d.x1) & (d.y0 < d.y1))) return;
    int ix0 = orc_trunc(d.x0), iy0 = orc_trunc(d.y0), ix1 = orc_trunc(d.x1), iy1 = orc_trunc(d.y1);
    if (ix0 >= ix1 || iy0 >= iy1) return;

    OrcFetch f;
    OrcPipe p;
    static uint32_t lut[2048];
    uint32_t r = resolve(c, &f, lut, &p);
    if (r > 1) { note(c, r); return; }
    if (r == 1) return;
    if (((ix0 | iy0 | ix1 | iy1) & 0xFF) == 0)
      orc_fill_box_a(&c->canvas, &p, alpha_i(c), ix0 >> 8, iy0 >> 8, ix1 >> 8, iy1 >> 8);
    else
      orc_fill_box_u(&c->canvas, &p, alpha_i(c), ix0, iy0, ix1, iy1);
  }
  else {
    OrcPt poly[4] = { { bx0, by0 }, { bx1, by0 }, { bx1, by1 }, { bx0, by1 } };
    fill_poly(c, poly, 4, 1 /* EVEN_ODD */);
  }
}

/* fillUnclippedRectI (rastercontext.cpp:2898-2917) */
static void fill_rect_i(Ctx* c, int x, int y, int w, int h) {
  int tt = orc_mat_type(&c->user);
  int integral = 0, tx = 0, ty = 0;
  if (tt <= ORC_TT_TRANSLATE) {
    double fx = c->user.m20 * 256.0, fy = c->user.m21 * 256.0;
    int64_t tx64 = orc_floor_to_int64(fx), ty64 = orc_floor_to_int64(fy);
    if (((tx64 | ty64) & 0xFF) == 0) { integral = 1; tx = (int)(tx64 >> 8); ty = (int)(ty64 >> 8); }
  }
  if (!integral) {
    if (w <= 0 || h <= 0) return;
    fill_box_d(c, (double)x, (double)y, (double)x + (double)w, (double)y + (double)h);
    return;
  }
  int64_t x0 = (int64_t)x + tx, y0 = (int64_t)y + ty;
  int64_t x1 = (int64_t)w + x0, y1 = (int64_t)h + y0;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 > (int64_t)c->canvas.w) x1 = c->canvas.w;
  if (y1 > (int64_t)c->canvas.h) y1 = c->canvas.h;
  if (x0 >= x1 || y0 >= y1) return;
  fill_box_i(c, (int)x0, (int)y0, (int)x1, (int)y1);
}

/* ---- blits (rastercontext.cpp:3938-4150) ----
 * A blit fills a box with the image area as a REFLECT pattern
 * (BL_RASTER_CONTEXT_PREFERRED_BLIT_EXTEND, raster/rastercontext_p.h:51)
 * under global alpha alone; comp-op simplification sees the image format. */

/* check_image_area (rastercontext.cpp:2305-2320) */
static int blit_area(const Ctx* c, uint32_t img, const int32_t* area, int32_t out[4]) {
  const SceneImage* ii = &c->img_info[img];
  out[0] = 0; out[1] = 0; out[2] = (int32_t)ii->width; out[3] = (int32_t)ii->height;
  if (!area) return 1;
  unsigned max_w = (unsigned)ii->width - (unsigned)area[0];
  unsigned max_h = (unsigned)ii->height - (unsigned)area[1];
  if ((max_w > ii->width) | ((unsigned)area[2] > max_w) | (max_h > ii->height) | ((unsigned)area[3] > max_h)) return 0;
  memcpy(out, area, 16);
  return 1;
}

/* The comp op replaces the image by a solid colour (solid_id of the
 * simplification with the image's format, prepare_overridden_fetch). */
static int blit_overridden(const Ctx* c, uint32_t img) {
  uint32_t pd, ps, sid;
  uint32_t op = orc_simplify(c->comp_op, c->canvas.format, c->img_info[img].format, &pd, &ps, &sid);
  return op != ORC_OP_NOP && sid != ORC_SOLID_NONE;
}

static void set_blit_style(Ctx* c, uint32_t img, const int32_t src[4], const OrcMat* t) {
  Style* st = &c->explicit_style;
  memset(st, 0, sizeof(*st));
  st->type = 2;
  st->xform = *t;
  st->quality = c->pattern_quality;
  st->pat.image_index = img;
  st->pat.extend = 2;   /* BL_EXTEND_MODE_REFLECT */
  st->has_view = 1;
  memcpy(st->view, src, 16);
  c->mode = 4;
}

/* blit_image_d_impl (rastercontext.cpp:3940-4031) */
static void blit_d(Ctx* c, double x, double y, uint32_t img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(c, img, area, src)) { note(c, ERR_INVALID_VALUE); return; }
  int tt = orc_mat_type(&c->user);
  if (tt == ORC_TT_INVALID) return;
  if (tt <= ORC_TT_TRANSLATE) {
    OrcMat m; int ftt;
    fixed_transform(c, &m, &ftt);
    double sx = x * m.m00 + m.m20, sy = y * m.m11 + m.m21;
    double dx0 = orc_max(sx, 0.0), dy0 = orc_max(sy, 0.0);
    double dx1 = orc_min(sx + (double)src[2] * m.m00, (double)c->canvas.w * 256.0);
    double dy1 = orc_min(sy + (double)src[3] * m.m11, (double)c->canvas.h * 256.0);
    int ix0 = 0, iy0 = 0, ix1 = 0, iy1 = 0;
    int empty = !((dx0 < dx1) & (dy0 < dy1));
    if (!empty) {
      ix0 = orc_trunc(dx0); iy0 = orc_trunc(dy0); ix1 = orc_trunc(dx1); iy1 = orc_trunc(dy1);
      empty = !((ix0 < ix1) & (iy0 < iy1));
    }
    if (empty && !blit_overridden(c, img)) return;
    OrcMat t = { 1.0, 0.0, 0.0, 1.0, sx / 256.0, sy / 256.0 };
    set_blit_style(c, img, src, &t);
    OrcFetch f;
    OrcPipe p;
    static uint32_t lut[2048];
    uint32_t r = resolve(c, &f, lut, &p);
    if (r == 1) return;
    if (r == 0 && f.type == ORC_FETCH_SOLID) {
      /* the comp op replaced the image by a solid colour: the unclipped box */
      fill_box_d(c, x, y, x + (double)src[2], y + (double)src[3]);
      return;
    }
    if (r > 1) { note(c, r); return; }
    /* fill_clipped_box_f (rastercontext.cpp:2687-2696) */
    if (((ix0 | iy0 | ix1 | iy1) & 0xFF) == 0)
      orc_fill_box_a(&c->canvas, &p, alpha_i(c), ix0 >> 8, iy0 >> 8, ix1 >> 8, iy1 >> 8);
    else
      orc_fill_box_u(&c->canvas, &p, alpha_i(c), ix0, iy0, ix1, iy1);
    return;
  }
  /* the area at the origin, translated to (x, y), through the user transform */
  OrcMat tr = { 1.0, 0.0, 0.0, 1.0, x, y }, t;
  orc_mat_multiply(&t, &tr, &c->user);
  set_blit_style(c, img, src, &t);
  fill_box_d(c, x, y, x + (double)src[2], y + (double)src[3]);
}

/* blit_image_i_impl (rastercontext.cpp:4033-4066) with
 * translate_and_clip_rect_to_blit_i (rastercontext.cpp:922-993) */
static void blit_i(Ctx* c, int x, int y, uint32_t img, const int32_t* area) {
  int tt = orc_mat_type(&c->user);
  int integral = 0, tx = 0, ty = 0;
  if (tt <= ORC_TT_TRANSLATE) {
    int64_t tx64 = orc_floor_to_int64(c->user.m20 * 256.0), ty64 = orc_floor_to_int64(c->user.m21 * 256.0);
    if (((tx64 | ty64) & 0xFF) == 0) { integral = 1; tx = (int)(tx64 >> 8); ty = (int)(ty64 >> 8); }
  }
  if (!integral) { blit_d(c, (double)x, (double)y, img, area); return; }
  int32_t src[4];
  if (!blit_area(c, img, area, src)) { note(c, ERR_INVALID_VALUE); return; }
  int64_t dx = (int64_t)x + tx, dy = (int64_t)y + ty;
  int64_t x0 = dx < 0 ? 0 : dx, y0 = dy < 0 ? 0 : dy;
  int64_t x1 = dx + src[2], y1 = dy + src[3];
  if (x1 > (int64_t)c->canvas.w) x1 = c->canvas.w;
  if (y1 > (int64_t)c->canvas.h) y1 = c->canvas.h;
  if (x0 >= x1 || y0 >= y1) return;
  OrcMat t = { 1.0, 0.0, 0.0, 1.0, (double)dx, (double)dy };
  set_blit_style(c, img, src, &t);
  fill_box_i(c, (int)x0, (int)y0, (int)x1, (int)y1);
}

/* blit_scaled_image_d_impl / _i_impl (rastercontext.cpp:4072-4150): the
 * area scaled onto the rectangle, through the user transform; a rectangle
 * of the area's size is a plain blit. */
static void blit_scaled(Ctx* c, int integer, double x, double y, double w, double h, uint32_t img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(c, img, area, src)) { note(c, ERR_INVALID_VALUE); return; }
  if (w == (double)src[2] && h == (double)src[3]) {
    if (integer) blit_i(c, (int)x, (int)y, img, area);
    else blit_d(c, x, y, img, area);
    return;
  }
  if (orc_mat_type(&c->user) == ORC_TT_INVALID) return;
  OrcMat sm = { w / (double)src[2], 0.0, 0.0, h / (double)src[3], x, y }, t;
  orc_mat_multiply(&t, &sm, &c->user);
  set_blit_style(c, img, src, &t);
  fill_box_d(c, x, y, x + w, y + h);
}

/* ---- mask fills (rastercontext.cpp:3590-3780) ----
 * Rendered only pixel-aligned under a translation
 * (fill_clipped_box_masked_a); any other placement is NOT_IMPLEMENTED in
 * the reference (fill_unclipped_mask_d). */

static void fill_masked_box(Ctx* c, int x0, int y0, int x1, int y1, uint32_t img, int mx, int my) {
  OrcFetch f;
  OrcPipe p;
  static uint32_t lut[2048];
  uint32_t r = resolve(c, &f, lut, &p);
  if (r > 1) { note(c, r); return; }
  if (r == 1) return;
  const SceneImage* ii = &c->img_info[img];
  const uint8_t* base = c->img_data[img] + (size_t)my * ii->stride + (size_t)mx * (ii->format == ORC_FMT_A8 ? 1u : 4u);
  orc_fill_box_masked(&c->canvas, &p, alpha_i(c), x0, y0, x1, y1, base, (intptr_t)ii->stride);
}

/* fill_mask_d_impl + fill_unclipped_mask_d (rastercontext.cpp:3594-3690) */
static void fill_mask_d(Ctx* c, double x, double y, uint32_t img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(c, img, area, src)) { note(c, ERR_INVALID_VALUE); return; }
  OrcFetch f;
  OrcPipe p;
  static uint32_t lut[2048];
  if (resolve(c, &f, lut, &p) == 1) return;   /* a NOP renders nothing and reports nothing */
  if (orc_mat_type(&c->user) > ORC_TT_TRANSLATE) { note(c, ERR_NOT_IMPLEMENTED); return; }
  OrcMat m; int ftt;
  fixed_transform(c, &m, &ftt);
  double sx = x * m.m00 + m.m20, sy = y * m.m11 + m.m21;
  double dx0 = orc_max(sx, 0.0), dy0 = orc_max(sy, 0.0);
  double dx1 = orc_min(sx + (double)src[2] * m.m00, (double)c->canvas.w * 256.0);
  double dy1 = orc_min(sy + (double)src[3] * m.m11, (double)c->canvas.h * 256.0);
  if (!((dx0 < dx1) & (dy0 < dy1))) return;
  int64_t fx = orc_floor_to_int64(sx), fy = orc_floor_to_int64(sy);
  int ix0 = orc_trunc(dx0), iy0 = orc_trunc(dy0), ix1 = orc_trunc(dx1), iy1 = orc_trunc(dy1);
  if (((fx | fy) & 0xFF) || ((ix0 | iy0 | ix1 | iy1) & 0xFF)) { note(c, ERR_NOT_IMPLEMENTED); return; }
  int x0 = ix0 >> 8, y0 = iy0 >> 8, x1 = (ix1 + 0xFF) >> 8, y1 = (iy1 + 0xFF) >> 8;
  int tx = (int)(fx >> 8), ty = (int)(fy >> 8);
  fill_masked_box(c, x0, y0, x1, y1, img, src[0] + x0 - tx, src[1] + y0 - ty);
}

/* fill_mask_i_impl (rastercontext.cpp:3723-3743) */
static void fill_mask_i(Ctx* c, int x, int y, uint32_t img, const int32_t* area) {
  int tt = orc_mat_type(&c->user);
  int integral = 0, tx = 0, ty = 0;
  if (tt <= ORC_TT_TRANSLATE) {
    int64_t tx64 = orc_floor_to_int64(c->user.m20 * 256.0), ty64 = orc_floor_to_int64(c->user.m21 * 256.0);
    if (((tx64 | ty64) & 0xFF) == 0) { integral = 1; tx = (int)(tx64 >> 8); ty = (int)(ty64 >> 8); }
  }
  if (!integral) { fill_mask_d(c, (double)x, (double)y, img, area); return; }
  int32_t src[4];
  if (!blit_area(c, img, area, src)) { note(c, ERR_INVALID_VALUE); return; }
  int64_t dx = (int64_t)x + tx, dy = (int64_t)y + ty;
  int64_t x0 = dx < 0 ? 0 : dx, y0 = dy < 0 ? 0 : dy;
  int64_t x1 = dx + src[2], y1 = dy + src[3];
  if (x1 > (int64_t)c->canvas.w) x1 = c->canvas.w;
  if (y1 > (int64_t)c->canvas.h) y1 = c->canvas.h;
  if (x0 >= x1 || y0 >= y1) return;
  fill_masked_box(c, (int)x0, (int)y0, (int)x1, (int)y1, img, src[0] + (int)(x0 - dx), src[1] + (int)(y0 - dy));
}

/* Path builder for BLPath::add_geometry (core/path.cpp:1396-1640),
 * clockwise direction. */
typedef struct { uint8_t* cmd; OrcPt* vtx; size_t n, cap; } TmpPath;

static void tp_push(TmpPath* p, uint8_t cmd, double x, double y) {
  if (p->n == p->cap) {
    p->cap = p->cap ? p->cap * 2 : 64;
    p->cmd = (uint8_t*)realloc(p->cmd, p->cap);
    p->vtx = (OrcPt*)realloc(p->vtx, p->cap * sizeof(OrcPt));
  }
  p->cmd[p->n] = cmd;
  p->vtx[p->n].x = x;
  p->vtx[p->n].y = y;
  p->n++;
}
static void tp_cubic(TmpPath* p, double x1, double y1, double x2, double y2, double x3, double y3) {
  tp_push(p, 4, x1, y1); tp_push(p, 4, x2, y2); tp_push(p, 1, x3, y3);
}
static void tp_close(TmpPath* p) { tp_push(p, 5, NAN, NAN); }
static void tp_box(TmpPath* p, double x0, double y0, double x1, double y1) {
  tp_push(p, 0, x0, y0); tp_push(p, 1, x1, y0); tp_push(p, 1, x1, y1); tp_push(p, 1, x0, y1); tp_close(p);
}

static const double kKappa = 0.55228474983079339840;

/* fill_unclipped_geometry (rastercontext.cpp:2925-2980) */
static void fill_geometry(Ctx* c, uint32_t type, const uint8_t* data, uint32_t count) {
  switch (type) {
    case 0: return;
    case 1: { int32_t b[4]; memcpy(b, data, 16); fill_box_d(c, b[0], b[1], b[2], b[3]); return; }
    case 2: { double b[4]; memcpy(b, data, 32); fill_box_d(c, b[0], b[1], b[2], b[3]); return; }
    case 3: { int32_t b[4]; memcpy(b, data, 16); fill_rect_i(c, b[0], b[1], b[2], b[3]); return; }
    case 4: { double b[4]; memcpy(b, data, 32); fill_box_d(c, b[0], b[1], b[0] + b[2], b[1] + b[3]); return; }
    case 13: case 14: case 15: case 16: {
      if (count < 3) return;
      OrcPt* v = (OrcPt*)malloc(count * sizeof(OrcPt));
      for (uint32_t i = 0; i < count; i++) {
        if (type == 13 || type == 15) { int32_t q[2]; memcpy(q, data + 8 * i, 8); v[i].x = q[0]; v[i].y = q[1]; }
        else memcpy(&v[i], data + 16 * i, 16);
      }
      fill_poly(c, v, count, c->fill_rule);
      free(v);
      return;
    }
    default: break;
  }
  TmpPath p;
  memset(&p, 0, sizeof(p));
  switch (type) {
    case 5: case 6: {
      double e[4];
      memcpy(e, data, type == 5 ? 24 : 32);
      double x0 = e[0], y0 = e[1], rx = e[2], ry = type == 5 ? fabs(e[2]) : e[3];
      double kx = rx * kKappa, ky = ry * kKappa;
      tp_push(&p, 0, x0 + rx, y0);
      tp_cubic(&p, x0 + rx, y0 + ky, x0 + kx, y0 + ry, x0, y0 + ry);
      tp_cubic(&p, x0 - kx, y0 + ry, x0 - rx, y0 + ky, x0 - rx, y0);
      tp_cubic(&p, x0 - rx, y0 - ky, x0 - kx, y0 - ry, x0, y0 - ry);
      tp_cubic(&p, x0 + kx, y0 - ry, x0 + rx, y0 - ky, x0 + rx, y0);
      tp_close(&p);
      break;
    }
    case 7: {
      double q[6];
      memcpy(q, data, 48);
      double x0 = q[0], y0 = q[1], x1 = q[0] + q[2], y1 = q[1] + q[3];
      double wh = q[2] * 0.5, hh = q[3] * 0.5;
      double rx = fabs(q[4]) < wh ? fabs(q[4]) : wh;
      double ry = fabs(q[5]) < hh ? fabs(q[5]) : hh;
      if (!(rx > 1e-14 && ry > 1e-14)) { tp_box(&p, x0, y0, x1, y1); break; }
      double kx = rx * (1.0 - kKappa), ky = ry * (1.0 - kKappa);
      tp_push(&p, 0, x0 + rx, y0);
      tp_push(&p, 1, x1 - rx, y0);
      tp_cubic(&p, x1 - kx, y0, x1, y0 + ky, x1, y0 + ry);
      tp_push(&p, 1, x1, y1 - ry);
      tp_cubic(&p, x1, y1 - ky, x1 - kx, y1, x1 - rx, y1);
      tp_push(&p, 1, x0 + rx, y1);
      tp_cubic(&p, x0 + kx, y1, x0, y1 - ky, x0, y1 - ry);
      tp_push(&p, 1, x0, y0 + ry);
      tp_cubic(&p, x0, y0 + ky, x0 + kx, y0, x0 + rx, y0);
      tp_close(&p);
      break;
    }
    case 11: {
      double l[4]; memcpy(l, data, 32);
      tp_push(&p, 0, l[0], l[1]); tp_push(&p, 1, l[2], l[3]);
      break;
    }
    case 12: {
      double t[6]; memcpy(t, data, 48);
      tp_push(&p, 0, t[0], t[1]); tp_push(&p, 1, t[2], t[3]); tp_push(&p, 1, t[4], t[5]); tp_close(&p);
      break;
    }
    case 17: case 18: case 19: case 20:
      /* invalid items are skipped (geometry/commons_p.h:42-56) */
      for (uint32_t i = 0; i < count; i++) {
        if (type == 17) {
          int32_t b[4]; memcpy(b, data + 16 * i, 16);
          if (b[0] < b[2] && b[1] < b[3]) tp_box(&p, b[0], b[1], b[2], b[3]);
        }
        else if (type == 18) {
          double b[4]; memcpy(b, data + 32 * i, 32);
          if (b[0] < b[2] && b[1] < b[3]) tp_box(&p, b[0], b[1], b[2], b[3]);
        }
        else if (type == 19) {
          int32_t b[4]; memcpy(b, data + 16 * i, 16);
          int64_t x1 = (int64_t)b[0] + b[2], y1 = (int64_t)b[1] + b[3];
          if (x1 != (int32_t)x1 || y1 != (int32_t)y1 || !(b[0] < x1 && b[1] < y1)) continue;
          double x0 = b[0], y0 = b[1];
          tp_box(&p, x0, y0, (double)b[2] + x0, (double)b[3] + y0);
        }
        else {
          double b[4]; memcpy(b, data + 32 * i, 32);
          double x1 = b[0] + b[2], y1 = b[1] + b[3];
          if (b[0] < x1 && b[1] < y1) tp_box(&p, b[0], b[1], b[2] + b[0], b[3] + b[1]);
        }
      }
      break;
    case 8: case 9: case 10: {
      /* arcs need the reference's arc approximation, which is not restated */
      OrcFetch f; OrcPipe pp; static uint32_t lut[2048];
      uint32_t r = resolve(c, &f, lut, &pp);
      if (r > 1) note(c, r);
      else if (r == 0) note(c, ERR_NOT_IMPLEMENTED);
      return;
    }
    default:
      note(c, ERR_INVALID_VALUE);
      return;
  }
  if (p.n) fill_path(c, p.cmd, p.vtx, p.n);
  free(p.cmd);
  free(p.vtx);
}

uint32_t orc_render(const uint8_t* scene, size_t size, uint8_t* out, size_t out_size) {
  Rd r = { scene, scene + size };
  SceneHeader hdr;
  if (!rd(&r, &hdr, sizeof(hdr)) || hdr.magic != SCENE_MAGIC || hdr.version != SCENE_VERSION) return 0x7FFFFFFFu;
  uint32_t bpp = hdr.format == SCENE_FMT_A8 ? 1u : 4u;
  if (out_size < (size_t)hdr.width * hdr.height * bpp) return 0x7FFFFFFFu;
  if (hdr.n_images > 16) return 0x7FFFFFFFu;

  Ctx c;
  memset(&c, 0, sizeof(c));
  c.canvas.pixels = out;
  c.canvas.stride = (intptr_t)hdr.width * bpp;
  c.canvas.w = hdr.width;
  c.canvas.h = hdr.height;
  c.canvas.format = hdr.format;
  c.comp_op = ORC_OP_SRC_OVER;
  c.fill_rule = 0;
  c.pattern_quality = 1;
  c.user.m00 = 1.0; c.user.m11 = 1.0;
  c.global_alpha = 1.0;
  c.fill_alpha = 1.0;
  c.style.type = 0;
  c.style.rgba32 = 0xFF000000u;
  orc_edges_init(&c.edges);

  for (uint32_t y = 0; y < hdr.height; y++) {
    uint8_t* row = out + (size_t)y * c.canvas.stride;
    if (bpp == 1) memset(row, (int)(hdr.clear_value & 0xFFu), hdr.width);
    else for (uint32_t x = 0; x < hdr.width; x++) memcpy(row + x * 4u, &hdr.clear_value, 4);
  }

  for (uint32_t i = 0; i < hdr.n_images; i++) {
    if (!rd(&r, &c.img_info[i], sizeof(SceneImage))) return 0x7FFFFFFFu;
    c.img_data[i] = r.p;
    r.p += pad8((size_t)c.img_info[i].height * c.img_info[i].stride);
  }
  c.n_images = hdr.n_images;

  for (uint32_t i = 0; i < hdr.n_cmds; i++) {
    SceneCmd cmd;
    if (!rd(&r, &cmd, sizeof(cmd)) || (size_t)(r.end - r.p) < cmd.size) { orc_edges_free(&c.edges); return 0x7FFFFFFFu; }
    const uint8_t* pl = r.p;
    r.p += cmd.size;
    switch (cmd.op) {
      case SCENE_OP_SET_COMP_OP: memcpy(&c.comp_op, pl, 4); break;
      case SCENE_OP_SET_FILL_RULE: memcpy(&c.fill_rule, pl, 4); break;
      case SCENE_OP_SET_PATTERN_QUALITY: memcpy(&c.pattern_quality, pl, 4); break;
      case SCENE_OP_SET_GRADIENT_QUALITY: {
        /* set_hint(BL_CONTEXT_HINT_GRADIENT_QUALITY) (rastercontext.cpp:1816-1821) */
        uint32_t q;
        memcpy(&q, pl, 4);
        if (q > 2u) { note(&c, ERR_INVALID_VALUE); break; }
        c.gradient_quality = q;
        break;
      }
      case SCENE_OP_SET_TRANSFORM: memcpy(&c.user, pl, 48); break;
      case SCENE_OP_SET_GLOBAL_ALPHA: {
        double a;
        memcpy(&a, pl, 8);
        if (isnan(a)) { note(&c, ERR_INVALID_VALUE); break; }
        c.global_alpha = orc_clamp(a, 0.0, 1.0);
        break;
      }
      case SCENE_OP_SET_STYLE_SOLID:
        c.style.type = 0;
        memcpy(&c.style.rgba32, pl, 4);
        break;
      case SCENE_OP_SET_STYLE_GRADIENT:
        parse_gradient_style(&c.style, pl, &c.user, c.gradient_quality);
        break;
      case SCENE_OP_SET_STYLE_PATTERN:
        parse_pattern_style(&c.style, pl, &c.user, c.pattern_quality);
        break;
      case SCENE_OP_FILL_EXPLICIT: {
        /* fill_*_rgba64 / fill_*_ext: resolve_explicit_style_op
         * (rastercontext.cpp:1277-1328) -- the style of this call only, with
         * the user transform of the moment */
        SceneExplicit ex;
        memcpy(&ex, pl, sizeof(ex));
        const uint8_t* p = pl + sizeof(ex);
        if (ex.style_kind == 0 || ex.style_kind == 1) {
          c.mode = 1;
          c.explicit_rgba32 = ex.style_kind == 0 ? (uint32_t)ex.rgba : rgba32_from_rgba64(ex.rgba);
        }
        else {
          c.mode = 3;
          p += ex.style_kind == 2 ? parse_gradient_style(&c.explicit_style, p, &c.user, c.gradient_quality)
                                  : parse_pattern_style(&c.explicit_style, p, &c.user, c.pattern_quality);
        }
        switch (ex.target) {
          case 0:
            c.call_all = 1;
            fill_box_i(&c, 0, 0, (int)c.canvas.w, (int)c.canvas.h);
            c.call_all = 0;
            break;
          case 1: { int32_t v[4]; memcpy(v, p, 16); fill_rect_i(&c, v[0], v[1], v[2], v[3]); break; }
          case 2: { double d[4]; memcpy(d, p, 32); fill_box_d(&c, d[0], d[1], d[0] + d[2], d[1] + d[3]); break; }
          case 3: {
            OrcPt* vtx = (OrcPt*)malloc((size_t)ex.count * sizeof(OrcPt) + 1);
            memcpy(vtx, p + pad8(ex.count), (size_t)ex.count * sizeof(OrcPt));
            if (ex.count) fill_path(&c, p, vtx, ex.count);
            free(vtx);
            break;
          }
          case 4: fill_geometry(&c, ex.geom_type, p, ex.count); break;
          case 5: {
            SceneMask mk;
            memcpy(&mk, p, sizeof(mk));
            if (mk.image_index >= c.n_images) { note(&c, ERR_INVALID_VALUE); break; }
            if (mk.kind == 0) fill_mask_i(&c, (int)mk.dst[0], (int)mk.dst[1], mk.image_index, mk.has_area ? mk.area : NULL);
            else fill_mask_d(&c, mk.dst[0], mk.dst[1], mk.image_index, mk.has_area ? mk.area : NULL);
            break;
          }
          default: break;
        }
        c.mode = 0;
        break;
      }
      case SCENE_OP_FILL_MASK: {
        SceneMask mk;
        memcpy(&mk, pl, sizeof(mk));
        if (mk.image_index >= c.n_images) { note(&c, ERR_INVALID_VALUE); break; }
        if (mk.kind == 0) fill_mask_i(&c, (int)mk.dst[0], (int)mk.dst[1], mk.image_index, mk.has_area ? mk.area : NULL);
        else fill_mask_d(&c, mk.dst[0], mk.dst[1], mk.image_index, mk.has_area ? mk.area : NULL);
        break;
      }
      case SCENE_OP_FILL_PATH: {
        uint32_t n;
        memcpy(&n, pl, 4);
        const uint8_t* cmds = pl + 8;
        OrcPt* vtx = (OrcPt*)malloc((size_t)n * sizeof(OrcPt) + 1);
        memcpy(vtx, cmds + pad8(n), (size_t)n * sizeof(OrcPt));
        if (n) fill_path(&c, cmds, vtx, n);
        free(vtx);
        break;
      }
      case SCENE_OP_FILL_POLYGON: {
        uint32_t n;
        memcpy(&n, pl, 4);
        OrcPt* vtx = (OrcPt*)malloc((size_t)n * sizeof(OrcPt) + 1);
        memcpy(vtx, pl + 8, (size_t)n * sizeof(OrcPt));
        fill_poly(&c, vtx, n, c.fill_rule);
        free(vtx);
        break;
      }
      case SCENE_OP_FILL_RECT_D: {
        double d[4];
        memcpy(d, pl, 32);
        fill_box_d(&c, d[0], d[1], d[0] + d[2], d[1] + d[3]);
        break;
      }
      case SCENE_OP_FILL_RECT_I: {
        int32_t v[4];
        memcpy(v, pl, 16);
        fill_rect_i(&c, v[0], v[1], v[2], v[3]);
        break;
      }
      case SCENE_OP_FILL_ALL:
        c.call_all = 1;
        fill_box_i(&c, 0, 0, (int)c.canvas.w, (int)c.canvas.h);
        c.call_all = 0;
        break;
      case SCENE_OP_SET_FILL_ALPHA: {
        double a;
        memcpy(&a, pl, 8);
        if (isnan(a)) { note(&c, ERR_INVALID_VALUE); break; }
        c.fill_alpha = orc_clamp(a, 0.0, 1.0);
        break;
      }
      case SCENE_OP_SET_STYLE_RGBA64: {
        uint64_t v;
        memcpy(&v, pl, 8);
        c.style.type = 0;
        c.style.rgba32 = rgba32_from_rgba64(v);
        break;
      }
      case SCENE_OP_SET_STYLE_RGBA: {
        /* set_style_rgba_impl (rastercontext.cpp:616-655) */
        float f[4];
        memcpy(f, pl, 16);
        if (isnan(f[0]) || isnan(f[1]) || isnan(f[2]) || isnan(f[3])) { c.style.type = 3; break; }
        for (int k = 0; k < 4; k++) f[k] = f[k] < 0.0f ? 0.0f : (f[k] > 1.0f ? 1.0f : f[k]);
        float as = f[3] * 255.0f;
        uint32_t rr = (uint32_t)lrintf(f[0] * as), gg = (uint32_t)lrintf(f[1] * as), bb = (uint32_t)lrintf(f[2] * as);
        uint32_t aa = (uint32_t)lrintf(as);
        c.style.type = 4;
        c.style.rgba32 = (aa << 24) | (rr << 16) | (gg << 8) | bb;
        break;
      }
      case SCENE_OP_CLEAR_ALL:
        c.mode = 2; c.call_all = 1;
        fill_box_i(&c, 0, 0, (int)c.canvas.w, (int)c.canvas.h);
        c.mode = 0; c.call_all = 0;
        break;
      case SCENE_OP_CLEAR_RECT_I: {
        int32_t v[4];
        memcpy(v, pl, 16);
        c.mode = 2;
        fill_rect_i(&c, v[0], v[1], v[2], v[3]);
        c.mode = 0;
        break;
      }
      case SCENE_OP_CLEAR_RECT_D: {
        double d[4];
        memcpy(d, pl, 32);
        c.mode = 2;
        fill_box_d(&c, d[0], d[1], d[0] + d[2], d[1] + d[3]);
        c.mode = 0;
        break;
      }
      case SCENE_OP_FILL_ALL_RGBA32:
        memcpy(&c.explicit_rgba32, pl, 4);
        c.mode = 1; c.call_all = 1;
        fill_box_i(&c, 0, 0, (int)c.canvas.w, (int)c.canvas.h);
        c.mode = 0; c.call_all = 0;
        break;
      case SCENE_OP_FILL_RECT_I_RGBA32: {
        int32_t v[4];
        memcpy(v, pl, 16);
        memcpy(&c.explicit_rgba32, pl + 16, 4);
        c.mode = 1;
        fill_rect_i(&c, v[0], v[1], v[2], v[3]);
        c.mode = 0;
        break;
      }
      case SCENE_OP_FILL_RECT_D_RGBA32: {
        double d[4];
        memcpy(d, pl, 32);
        memcpy(&c.explicit_rgba32, pl + 32, 4);
        c.mode = 1;
        fill_box_d(&c, d[0], d[1], d[0] + d[2], d[1] + d[3]);
        c.mode = 0;
        break;
      }
      case SCENE_OP_FILL_GEOMETRY: {
        SceneGeometry g;
        memcpy(&g, pl, sizeof(g));
        if (g.has_rgba32) { c.mode = 1; c.explicit_rgba32 = g.rgba32; }
        fill_geometry(&c, g.type, pl + sizeof(g), g.count);
        c.mode = 0;
        break;
      }
      case SCENE_OP_BLIT_IMAGE: {
        SceneBlit b;
        memcpy(&b, pl, sizeof(b));
        if (b.image_index >= c.n_images) { note(&c, ERR_INVALID_VALUE); break; }
        const int32_t* area = b.has_area ? b.area : NULL;
        switch (b.kind) {
          case 0: blit_i(&c, (int)b.dst[0], (int)b.dst[1], b.image_index, area); break;
          case 1: blit_d(&c, b.dst[0], b.dst[1], b.image_index, area); break;
          case 2: blit_scaled(&c, 1, (double)(int)b.dst[0], (double)(int)b.dst[1], (double)(int)b.dst[2],
                              (double)(int)b.dst[3], b.image_index, area); break;
          case 3: blit_scaled(&c, 0, b.dst[0], b.dst[1], b.dst[2], b.dst[3], b.image_index, area); break;
          default: break;
        }
        c.mode = 0;
        break;
      }
      case SCENE_OP_FILL_PATH_EX: {
        ScenePathEx h;
        memcpy(&h, pl, sizeof(h));
        const uint8_t* cmds = pl + sizeof(h);
        OrcPt* vtx = (OrcPt*)malloc((size_t)h.n * sizeof(OrcPt) + 1);
        memcpy(vtx, cmds + pad8(h.n), (size_t)h.n * sizeof(OrcPt));
        if (h.flags & 2u) { c.mode = 1; c.explicit_rgba32 = h.rgba32; }
        if (h.n) fill_path_origin(&c, cmds, vtx, h.n, (h.flags & 1u) ? h.origin : NULL);
        c.mode = 0;
        free(vtx);
        break;
      }
      default:
        break;
    }
  }
  orc_edges_free(&c.edges);
  return c.first_error;
}

/* ---- test hooks ---- */

size_t orc_path_segments(const uint8_t* cmds, const double* vtx, size_t n, const double* m6,
                         uint32_t width, uint32_t height, int32_t* out, size_t cap) {
  OrcEdges e;
  orc_edges_init(&e);
  OrcMat user = { m6[0], m6[1], m6[2], m6[3], m6[4], m6[5] };
  OrcMat m = user;
  m.m00 *= 256.0; m.m01 *= 256.0; m.m10 *= 256.0; m.m11 *= 256.0; m.m20 *= 256.0; m.m21 *= 256.0;
  int tt = orc_mat_type(&user);
  if (tt < ORC_TT_SCALE) tt = ORC_TT_SCALE;
  OrcBox clip = { 0.0, 0.0, (double)width * 256.0, (double)height * 256.0 };
  orc_edges_add_path(&e, cmds, (const OrcPt*)vtx, n, &m, tt, &clip, kToleranceSq);
  size_t cnt = e.n;
  for (size_t i = 0; i < cnt && i < cap; i++) {
    out[i * 5 + 0] = e.segs[i].x0; out[i * 5 + 1] = e.segs[i].y0;
    out[i * 5 + 2] = e.segs[i].x1; out[i * 5 + 3] = e.segs[i].y1;
    out[i * 5 + 4] = (int32_t)e.segs[i].sign;
  }
  orc_edges_free(&e);
  return cnt;
}

void orc_gradient_lut(const double* offsets, const uint32_t* rgba32, size_t n, uint32_t* out, uint32_t* lut_size_out) {
  Style s;
  memset(&s, 0, sizeof(s));
  for (size_t i = 0; i < n; i++) add_stop(&s, offsets[i], rgba32[i]);
  int solid;
  uint32_t fmt;
  uint32_t lut_size = orc_gradient_info(s.stops, s.n_stops, &solid, &fmt);
  *lut_size_out = lut_size;
  if (lut_size) orc_interpolate_lut(out, lut_size, s.stops, s.n_stops);
}

/* The 64-bit table of a dithered gradient: twice the 32-bit size, at most
 * 1024 entries (GradientInfo::lut_size, core/gradient_p.h:113-120). */
void orc_gradient_lut64(const double* offsets, const uint32_t* rgba32, size_t n, uint64_t* out, uint32_t* lut_size_out) {
  Style s;
  memset(&s, 0, sizeof(s));
  for (size_t i = 0; i < n; i++) add_stop(&s, offsets[i], rgba32[i]);
  int solid;
  uint32_t fmt;
  uint32_t lut_size = orc_gradient_info(s.stops, s.n_stops, &solid, &fmt);
  if (lut_size) lut_size = lut_size * 2u < 1024u ? lut_size * 2u : 1024u;
  *lut_size_out = lut_size;
  if (lut_size) orc_interpolate_lut64(out, lut_size, s.stops, s.n_stops);
}

uint32_t orc_composite_pixel(uint32_t op, uint32_t dst_format, uint32_t d, uint32_t s, uint32_t m) {
  OrcFetch f;
  memset(&f, 0, sizeof(f));
  f.type = ORC_FETCH_SOLID;
  f.solid_prgb32 = s;
  OrcPipe p;
  memset(&p, 0, sizeof(p));
  p.fetch = &f;
  p.op = op;
  uint8_t px[4];
  memcpy(px, &d, 4);
  OrcCanvas c = { px, 4, 1, 1, dst_format };
  orc_pipe_composite(&p, &c, 0, 0, m);
  uint32_t r = 0;
  memcpy(&r, px, dst_format == ORC_FMT_A8 ? 1 : 4);
  return r;
}

/* orc_composite_pixel over arrays: out[i] = op(d[i], s[i], m[i]). */
void orc_composite_array(uint32_t op, uint32_t dst_format, const uint32_t* d, const uint32_t* s, const uint32_t* m,
                         size_t n, uint32_t* out) {
  for (size_t i = 0; i < n; i++) out[i] = orc_composite_pixel(op, dst_format, d[i], s[i], m[i]);
}

uint32_t orc_simplify_op(uint32_t op, uint32_t dst_fmt, uint32_t src_fmt) {
  uint32_t d, s, id;
  uint32_t r = orc_simplify(op, dst_fmt, src_fmt, &d, &s, &id);
  return (r & 0xFFu) | (d << 8) | (s << 16) | (id << 24);
}

/* Coverage masks of a segment list: the segments rasterized as one fill
 * (SRC_COPY, opaque white, alpha 255) onto an A8 canvas cleared to zero, so
 * each pixel holds its mask.  `segs` holds (x0, y0, x1, y1, sign) records. */
void orc_segments_mask(const int32_t* segs, size_t n, uint32_t w, uint32_t h, uint32_t fill_rule, uint8_t* out) {
  OrcEdges e;
  orc_edges_init(&e);
  e.segs = (OrcSeg*)malloc((n ? n : 1) * sizeof(OrcSeg));
  e.cap = n;
  for (size_t i = 0; i < n; i++) {
    OrcSeg* s = &e.segs[e.n++];
    s->x0 = segs[i * 5 + 0]; s->y0 = segs[i * 5 + 1]; s->x1 = segs[i * 5 + 2]; s->y1 = segs[i * 5 + 3];
    s->sign = (uint32_t)segs[i * 5 + 4];
    if (s->y0 < e.bbox_y0) e.bbox_y0 = s->y0;
    if (s->y1 > e.bbox_y1) e.bbox_y1 = s->y1;
  }
  memset(out, 0, (size_t)w * h);
  OrcCanvas c = { out, (intptr_t)w, w, h, ORC_FMT_A8 };
  OrcFetch f;
  memset(&f, 0, sizeof(f));
  f.type = ORC_FETCH_SOLID;
  f.solid_prgb32 = 0xFFFFFFFFu;
  OrcPipe p;
  memset(&p, 0, sizeof(p));
  p.fetch = &f;
  p.op = ORC_OP_SRC_COPY;
  orc_fill_analytic(&c, &e, &p, 255, fill_rule);
  orc_edges_free(&e);
}
