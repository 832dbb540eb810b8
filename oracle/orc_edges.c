/* CPU oracle -- edge builder: transform, clip, monotonic split, flatten.
 *
 * TEST INFRASTRUCTURE (see orc_core.h).  Restates raster/edgebuilder_p.h and
 * geometry/bezier_p.h of the reference.  The reference links the emitted
 * points into monotonic chains (EdgeVector); the analytic rasterizer consumes
 * every consecutive point pair of a chain independently, so this restatement
 * emits those point pairs directly as OrcSeg records.  Border lines that the
 * reference merges in its accumulators (edgebuilder_p.h:2546-2622) are emitted
 * per accumulate call: the merged and split vertical lines rasterize to the
 * same cells because the cover/area a vertical line adds to each scanline is
 * additive over adjacent y intervals.  Lines and curves are processed one
 * path segment at a time; the reference's run-skipping loops accumulate the
 * same telescoping border ranges.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"

enum { CLIP_X0 = 1, CLIP_Y0 = 2, CLIP_X1 = 4, CLIP_Y1 = 8 };
enum { SHIFT_X0 = 0, SHIFT_X1 = 2 };

typedef struct {
  OrcEdges* out;
  OrcBox clip;          /* clip box in fixed point (double) */
  int32_t clip_ix0, clip_ix1;
  double tol_sq;
  /* Appender state (edgebuilder_p.h:835-894) */
  int32_t app_x, app_y;
  uint32_t app_sign;
  int app_open;
  /* Set when the line clipper meets a clip-flag combination no case handles
   * (NaN coordinates): the reference fails the path with
   * BL_ERROR_INVALID_GEOMETRY (edgebuilder_p.h:1494-1496, 1584-1586). */
  int invalid;
} EB;

/* ---- segment storage ---- */

void orc_edges_init(OrcEdges* e) { memset(e, 0, sizeof(*e)); e->bbox_y0 = INT32_MAX; e->bbox_y1 = INT32_MIN; }
void orc_edges_free(OrcEdges* e) { free(e->segs); orc_edges_init(e); }
void orc_edges_clear(OrcEdges* e) { e->n = 0; e->bbox_y0 = INT32_MAX; e->bbox_y1 = INT32_MIN; }

static void push_stored(EB* eb, int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign) {
  OrcEdges* e = eb->out;
  if (y0 == y1) return;   /* horizontal: the rasterizer's prepare() rejects it */
  if (e->n == e->cap) {
    e->cap = e->cap ? e->cap * 2 : 1024;
    e->segs = (OrcSeg*)realloc(e->segs, e->cap * sizeof(OrcSeg));
  }
  OrcSeg* s = &e->segs[e->n++];
  s->x0 = x0; s->y0 = y0; s->x1 = x1; s->y1 = y1; s->sign = sign;
  if (y0 < e->bbox_y0) e->bbox_y0 = y0;
  if (y1 > e->bbox_y1) e->bbox_y1 = y1;
}

/* Directed segment: stored top-down with sign 1 when the source went up
 * (descending/ascending chains, edgebuilder_p.h:2452-2530). */
static void emit_dir(EB* eb, int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
  if (y0 < y1) push_stored(eb, x0, y0, x1, y1, 0);
  else if (y0 > y1) push_stored(eb, x1, y1, x0, y0, 1);
}

/* add_line_segment (edgebuilder_p.h:1070-1090). */
static void add_line_segment(EB* eb, double x0, double y0, double x1, double y1) {
  emit_dir(eb, orc_trunc(x0), orc_trunc(y0), orc_trunc(x1), orc_trunc(y1));
}

/* accumulate_left/right_border + _emit_*_border (edgebuilder_p.h:2546-2622). */
static void acc_left(EB* eb, double y0, double y1) {
  emit_dir(eb, eb->clip_ix0, orc_trunc(y0), eb->clip_ix0, orc_trunc(y1));
}
static void acc_right(EB* eb, double y0, double y1) {
  emit_dir(eb, eb->clip_ix1, orc_trunc(y0), eb->clip_ix1, orc_trunc(y1));
}
static void acc_left_s(EB* eb, double y0, double y1, uint32_t sign) {
  if (sign) { double t = y0; y0 = y1; y1 = t; }
  acc_left(eb, y0, y1);
}
static void acc_right_s(EB* eb, double y0, double y1, uint32_t sign) {
  if (sign) { double t = y0; y0 = y1; y1 = t; }
  acc_right(eb, y0, y1);
}

/* Appender: points arrive top-down, stored with the chain sign. */
static void app_open_at(EB* eb, double x, double y) {
  eb->app_x = orc_trunc(x); eb->app_y = orc_trunc(y); eb->app_open = 1;
}
static void app_add_line(EB* eb, double x, double y) {
  int32_t fx = orc_trunc(x), fy = orc_trunc(y);
  push_stored(eb, eb->app_x, eb->app_y, fx, fy, eb->app_sign);
  eb->app_x = fx; eb->app_y = fy;
}
static void app_close(EB* eb) { eb->app_open = 0; }

/* ---- clip flags (edgebuilder_p.h:48-55) ---- */
static inline uint32_t fx0(const EB* eb, OrcPt p) { return (uint32_t)(!(p.x >= eb->clip.x0)) * CLIP_X0; }
static inline uint32_t fx1(const EB* eb, OrcPt p) { return (uint32_t)(!(p.x <= eb->clip.x1)) * CLIP_X1; }
static inline uint32_t fy0(const EB* eb, OrcPt p) { return (uint32_t)(!(p.y >= eb->clip.y0)) * CLIP_Y0; }
static inline uint32_t fy1(const EB* eb, OrcPt p) { return (uint32_t)(!(p.y <= eb->clip.y1)) * CLIP_Y1; }
static inline uint32_t fxf(const EB* eb, OrcPt p) { return fx0(eb, p) | fx1(eb, p); }
static inline uint32_t fyf(const EB* eb, OrcPt p) { return fy0(eb, p) | fy1(eb, p); }
static inline uint32_t fxy(const EB* eb, OrcPt p) { return fxf(eb, p) | fyf(eb, p); }

/* ---- lines (edgebuilder_p.h:1123-1605, one source line at a time) ---- */

static void clip_end(EB* eb, OrcPt p, OrcPt a, OrcPt b, uint32_t bf) {
  OrcPt d = { b.x - a.x, b.y - a.y };
  OrcPt q = { eb->clip.x1, eb->clip.y1 };
  switch (bf) {
    case CLIP_X0 | CLIP_Y0: q.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1 | CLIP_Y0:
      q.y = a.y + (q.x - a.x) * d.y / d.x;
      if (q.y >= eb->clip.y0) break;
      /* fallthrough */
    case CLIP_Y0:
      q.y = eb->clip.y0;
      q.x = a.x + (q.y - a.y) * d.x / d.y;
      break;
    case CLIP_X0 | CLIP_Y1: q.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1 | CLIP_Y1:
      q.y = a.y + (q.x - a.x) * d.y / d.x;
      if (q.y <= eb->clip.y1) break;
      /* fallthrough */
    case CLIP_Y1:
      q.y = eb->clip.y1;
      q.x = a.x + (q.y - a.y) * d.x / d.y;
      break;
    case CLIP_X0: q.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1:
      q.y = a.y + (q.x - a.x) * d.y / d.x;
      break;
    default:
      eb->invalid = 1; /* edgebuilder_p.h:1584-1586 */
      return;
  }
  add_line_segment(eb, p.x, p.y, q.x, q.y);
  double cby = orc_clamp(b.y, eb->clip.y0, eb->clip.y1);
  if (q.y != cby) {
    if (q.x == eb->clip.x0) acc_left(eb, q.y, cby);
    else acc_right(eb, q.y, cby);
  }
}

static void eb_line(EB* eb, OrcPt a, OrcPt b) {
  uint32_t af = fxy(eb, a);
  uint32_t bf = fxy(eb, b);

  if (!af) {
    if (!bf) {
      emit_dir(eb, orc_trunc(a.x), orc_trunc(a.y), orc_trunc(b.x), orc_trunc(b.y));
      return;
    }
    clip_end(eb, a, a, b, bf);
    return;
  }

  double bor_y0, bor_y1;
  if (af & CLIP_Y0) {
    if (!(eb->clip.y0 < b.y)) return;
    bor_y0 = eb->clip.y0;
    uint32_t common = af & bf;
    if (common) {
      bor_y1 = orc_min(eb->clip.y1, b.y);
      if (common & CLIP_X0) acc_left(eb, bor_y0, bor_y1);
      else acc_right(eb, bor_y0, bor_y1);
      return;
    }
  }
  else if (af & CLIP_Y1) {
    if (!(eb->clip.y1 > b.y)) return;
    bor_y0 = eb->clip.y1;
    uint32_t common = af & bf;
    if (common) {
      bor_y1 = orc_max(eb->clip.y0, b.y);
      if (common & CLIP_X0) acc_left(eb, bor_y0, bor_y1);
      else acc_right(eb, bor_y0, bor_y1);
      return;
    }
  }
  else if (af & CLIP_X0) {
    bor_y0 = orc_clamp(a.y, eb->clip.y0, eb->clip.y1);
    if (!(eb->clip.x0 < b.x)) {
      bor_y1 = orc_clamp(b.y, eb->clip.y0, eb->clip.y1);
      if (bor_y0 != bor_y1) acc_left(eb, bor_y0, bor_y1);
      return;
    }
  }
  else {
    bor_y0 = orc_clamp(a.y, eb->clip.y0, eb->clip.y1);
    if (!(eb->clip.x1 > b.x)) {
      bor_y1 = orc_clamp(b.y, eb->clip.y0, eb->clip.y1);
      if (bor_y0 != bor_y1) acc_right(eb, bor_y0, bor_y1);
      return;
    }
  }

  /* Clip the start point (edgebuilder_p.h:1436-1500). */
  OrcPt d = { b.x - a.x, b.y - a.y };
  OrcPt p = { eb->clip.x1, eb->clip.y1 };
  switch (af) {
    case CLIP_X0 | CLIP_Y0: p.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1 | CLIP_Y0:
      p.y = a.y + (p.x - a.x) * d.y / d.x;
      af = fyf(eb, p);
      if (p.y >= eb->clip.y0) break;
      /* fallthrough */
    case CLIP_Y0:
      p.y = eb->clip.y0;
      p.x = a.x + (p.y - a.y) * d.x / d.y;
      af = fxf(eb, p);
      break;
    case CLIP_X0 | CLIP_Y1: p.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1 | CLIP_Y1:
      p.y = a.y + (p.x - a.x) * d.y / d.x;
      af = fyf(eb, p);
      if (p.y <= eb->clip.y1) break;
      /* fallthrough */
    case CLIP_Y1:
      p.y = eb->clip.y1;
      p.x = a.x + (p.y - a.y) * d.x / d.y;
      af = fxf(eb, p);
      break;
    case CLIP_X0: p.x = eb->clip.x0; /* fallthrough */
    case CLIP_X1:
      p.y = a.y + (p.x - a.x) * d.y / d.x;
      af = fyf(eb, p);
      break;
    default:
      eb->invalid = 1; /* edgebuilder_p.h:1494-1496 */
      return;
  }

  if (af) {
    bor_y1 = orc_clamp(b.y, eb->clip.y0, eb->clip.y1);
    if (p.x <= eb->clip.x0) acc_left(eb, bor_y0, bor_y1);
    else if (p.x >= eb->clip.x1) acc_right(eb, bor_y0, bor_y1);
    return;
  }

  bor_y1 = orc_clamp(p.y, eb->clip.y0, eb->clip.y1);
  if (bor_y0 != bor_y1) {
    if (p.x <= eb->clip.x0) acc_left(eb, bor_y0, bor_y1);
    else acc_right(eb, bor_y0, bor_y1);
  }

  if (!bf) {
    emit_dir(eb, orc_trunc(p.x), orc_trunc(p.y), orc_trunc(b.x), orc_trunc(b.y));
    return;
  }
  clip_end(eb, p, a, b, bf);
}

/* ---- monotonic curve flattening (edgebuilder_p.h:418-768, 2029-2445) ---- */

enum { RECURSION_LIMIT = 32 };

typedef struct {
  int n;                      /* points per curve: 3 quad, 4 cubic */
  double tol_sq;
  OrcPt p[4];
  OrcPt stack[RECURSION_LIMIT * 4];
  int sp;                     /* number of stacked curves */
} Mono;

typedef struct {
  double value, limit;
  OrcPt p01, p12, p23, p012, p123, p0123;
} Step;

static inline OrcPt pmid(OrcPt a, OrcPt b) { OrcPt r = { (a.x + b.x) * 0.5, (a.y + b.y) * 0.5 }; return r; }
static inline double cross(OrcPt a, OrcPt b) { return a.x * b.y - a.y * b.x; }
static inline double mag_sq(OrcPt v) { return v.x * v.x + v.y * v.y; }
static inline OrcPt psub(OrcPt a, OrcPt b) { OrcPt r = { a.x - b.x, a.y - b.y }; return r; }

static void mono_begin(Mono* m, const OrcPt* src, uint32_t sign) {
  m->sp = 0;
  for (int i = 0; i < m->n; i++) m->p[i] = sign ? src[m->n - 1 - i] : src[i];
}
static inline OrcPt mfirst(const Mono* m) { return m->p[0]; }
static inline OrcPt mlast(const Mono* m) { return m->p[m->n - 1]; }
static inline int can_pop(const Mono* m) { return m->sp != 0; }
static inline int can_push(const Mono* m) { return m->sp != RECURSION_LIMIT; }
static inline int is_ltr(const Mono* m) { return mfirst(m).x < mlast(m).x; }

static void bound_ltr(Mono* m) {
  OrcPt f = mfirst(m), l = mlast(m);
  for (int i = 1; i < m->n - 1; i++) {
    m->p[i].x = orc_clamp(m->p[i].x, f.x, l.x);
    m->p[i].y = orc_clamp(m->p[i].y, f.y, l.y);
  }
}
static void bound_rtl(Mono* m) {
  OrcPt f = mfirst(m), l = mlast(m);
  for (int i = 1; i < m->n - 1; i++) {
    m->p[i].x = orc_clamp(m->p[i].x, l.x, f.x);
    m->p[i].y = orc_clamp(m->p[i].y, f.y, l.y);
  }
}

static int is_flat(const Mono* m, Step* s) {
  if (m->n == 3) {
    OrcPt v1 = psub(m->p[1], m->p[0]);
    OrcPt v2 = psub(m->p[2], m->p[0]);
    double d = cross(v2, v1);
    double len_sq = mag_sq(v2);
    s->value = d * d;
    s->limit = m->tol_sq * len_sq;
  }
  else {
    OrcPt v = psub(m->p[3], m->p[0]);
    double c1 = cross(v, psub(m->p[1], m->p[0]));
    double c2 = cross(v, psub(m->p[2], m->p[0]));
    double d1_sq = c1 * c1;
    double d2_sq = c2 * c2;
    double len_sq = mag_sq(v);
    s->value = orc_max(d1_sq, d2_sq);
    s->limit = m->tol_sq * len_sq;
  }
  return s->value <= s->limit;
}

static void msplit(const Mono* m, Step* s) {
  if (m->n == 3) {
    s->p01 = pmid(m->p[0], m->p[1]);
    s->p12 = pmid(m->p[1], m->p[2]);
    s->p012 = pmid(s->p01, s->p12);
  }
  else {
    s->p01 = pmid(m->p[0], m->p[1]);
    s->p12 = pmid(m->p[1], m->p[2]);
    s->p23 = pmid(m->p[2], m->p[3]);
    s->p012 = pmid(s->p01, s->p12);
    s->p123 = pmid(s->p12, s->p23);
    s->p0123 = pmid(s->p012, s->p123);
  }
}
static inline OrcPt step_mid(const Mono* m, const Step* s) { return m->n == 3 ? s->p012 : s->p0123; }

static void mpush(Mono* m, const Step* s) {
  OrcPt* st = &m->stack[m->sp * 4];
  if (m->n == 3) {
    st[0] = s->p012; st[1] = s->p12; st[2] = m->p[2];
    m->p[1] = s->p01; m->p[2] = s->p012;
  }
  else {
    st[0] = s->p0123; st[1] = s->p123; st[2] = s->p23; st[3] = m->p[3];
    m->p[1] = s->p01; m->p[2] = s->p012; m->p[3] = s->p0123;
  }
  m->sp++;
}
static void mdiscard(Mono* m, const Step* s) {
  if (m->n == 3) { m->p[0] = s->p012; m->p[1] = s->p12; }
  else { m->p[0] = s->p0123; m->p[1] = s->p123; m->p[2] = s->p23; }
}
static void mpop(Mono* m) {
  m->sp--;
  const OrcPt* st = &m->stack[m->sp * 4];
  for (int i = 0; i < m->n; i++) m->p[i] = st[i];
}

static void flatten_safe(EB* eb, Mono* m, const OrcPt* src, uint32_t sign) {
  mono_begin(m, src, sign);
  eb->app_sign = sign;
  if (is_ltr(m)) bound_ltr(m); else bound_rtl(m);
  app_open_at(eb, mfirst(m).x, mfirst(m).y);
  for (;;) {
    Step s;
    if (!is_flat(m, &s)) {
      if (can_push(m)) { msplit(m, &s); mpush(m, &s); continue; }
    }
    app_add_line(eb, mlast(m).x, mlast(m).y);
    if (!can_pop(m)) break;
    mpop(m);
  }
  app_close(eb);
}

static inline double dx_div_dy(OrcPt d) { return d.x / d.y; }
static inline double dy_div_dx(OrcPt d) { return d.y / d.x; }

static void flatten_unsafe(EB* eb, Mono* m, const OrcPt* src, uint32_t sign) {
  const OrcBox* c = &eb->clip;
  mono_begin(m, src, sign);
  eb->app_sign = sign;

  double y_start = mfirst(m).y;
  double y_end = orc_min(mlast(m).y, c->y1);
  if ((y_start >= y_end) | (y_end <= c->y0)) return;

  const double kDeltaLimit = 0.00390625;
  double x_delta = fabs(mfirst(m).x - mlast(m).x);
  uint32_t out = 0;
  Step s;

  if (x_delta <= kDeltaLimit) {
    y_start = orc_max(y_start, c->y0);
    double x_min = orc_min(mfirst(m).x, mlast(m).x);
    double x_max = orc_max(mfirst(m).x, mlast(m).x);
    if (x_max <= c->x0) acc_left_s(eb, y_start, y_end, sign);
    else if (x_min >= c->x1) acc_right_s(eb, y_start, y_end, sign);
    else {
      app_open_at(eb, mfirst(m).x, y_start);
      app_add_line(eb, mlast(m).x, y_end);
      app_close(eb);
    }
    return;
  }
  else if (is_ltr(m)) {
    bound_ltr(m);
    if (y_start < c->y0) {
      y_start = c->y0;
      for (;;) {
        out = (mfirst(m).x >= c->x1);
        if (out) break;
        if (!is_flat(m, &s)) {
          msplit(m, &s);
          if (step_mid(m, &s).y <= c->y0) { mdiscard(m, &s); continue; }
          if (can_push(m)) { mpush(m, &s); continue; }
        }
        if (mlast(m).y > c->y0) {
          out = mlast(m).x < c->x0;
          if (out) goto LTR_BeforeX0_Pop;
          double x_clipped = mfirst(m).x + (c->y0 - mfirst(m).y) * dx_div_dy(psub(mlast(m), mfirst(m)));
          if (x_clipped <= c->x0) goto LTR_BeforeX0_Clip;
          out = (x_clipped >= c->x1);
          if (out) break;
          app_open_at(eb, x_clipped, c->y0);
          goto LTR_AddLine;
        }
        if (!can_pop(m)) break;
        mpop(m);
      }
      out <<= SHIFT_X1;
    }
    else if (y_start < c->y1) {
      if (mfirst(m).x < c->x0) {
        for (;;) {
          out = (mfirst(m).y >= c->y1);
          if (out) break;
          if (!is_flat(m, &s)) {
            msplit(m, &s);
            if (step_mid(m, &s).x <= c->x0) { mdiscard(m, &s); continue; }
            if (can_push(m)) { mpush(m, &s); continue; }
          }
          if (mlast(m).x > c->x0) {
LTR_BeforeX0_Clip:;
            double y_clipped = mfirst(m).y + (c->x0 - mfirst(m).x) * dy_div_dx(psub(mlast(m), mfirst(m)));
            out = (y_clipped >= y_end);
            if (out) break;
            if (y_start < y_clipped) acc_left_s(eb, y_start, y_clipped, sign);
            app_open_at(eb, c->x0, y_clipped);
            goto LTR_AddLine;
          }
          out = (mlast(m).y >= y_end);
          if (out) break;
LTR_BeforeX0_Pop:
          if (!can_pop(m)) break;
          mpop(m);
        }
        out <<= SHIFT_X0;
      }
      else if (mfirst(m).x < c->x1) {
        app_open_at(eb, mfirst(m).x, mfirst(m).y);
        for (;;) {
          if (!is_flat(m, &s)) {
            msplit(m, &s);
            if (can_push(m)) { mpush(m, &s); continue; }
          }
LTR_AddLine:
          out = mlast(m).x > c->x1;
          if (out) {
            double y_clipped = mfirst(m).y + (c->x1 - mfirst(m).x) * dy_div_dx(psub(mlast(m), mfirst(m)));
            if (y_clipped <= y_end) {
              y_start = y_clipped;
              app_add_line(eb, c->x1, y_clipped);
              break;
            }
          }
          out = mlast(m).y >= c->y1;
          if (out) {
            double x_clipped = orc_min(mfirst(m).x + (c->y1 - mfirst(m).y) * dx_div_dy(psub(mlast(m), mfirst(m))), c->x1);
            app_add_line(eb, x_clipped, c->y1);
            out = 0;
            break;
          }
          app_add_line(eb, mlast(m).x, mlast(m).y);
          if (!can_pop(m)) break;
          mpop(m);
        }
        app_close(eb);
        out <<= SHIFT_X1;
      }
      else {
        out = CLIP_X1;
      }
    }
  }
  else {
    bound_rtl(m);
    if (y_start < c->y0) {
      y_start = c->y0;
      for (;;) {
        out = (mfirst(m).x <= c->x0);
        if (out) break;
        if (!is_flat(m, &s)) {
          msplit(m, &s);
          if (step_mid(m, &s).y <= c->y0) { mdiscard(m, &s); continue; }
          if (can_push(m)) { mpush(m, &s); continue; }
        }
        if (mlast(m).y > c->y0) {
          out = mlast(m).x > c->x1;
          if (out) goto RTL_AfterX1_Pop;
          double x_clipped = mfirst(m).x + (c->y0 - mfirst(m).y) * dx_div_dy(psub(mlast(m), mfirst(m)));
          if (x_clipped >= c->x1) goto RTL_AfterX1_Clip;
          out = (x_clipped <= c->x0);
          if (out) break;
          app_open_at(eb, x_clipped, c->y0);
          goto RTL_AddLine;
        }
        if (!can_pop(m)) break;
        mpop(m);
      }
      out <<= SHIFT_X0;
    }
    else if (y_start < c->y1) {
      if (mfirst(m).x > c->x1) {
        for (;;) {
          out = (mfirst(m).y >= c->y1);
          if (out) break;
          if (!is_flat(m, &s)) {
            msplit(m, &s);
            if (step_mid(m, &s).x >= c->x1) { mdiscard(m, &s); continue; }
            if (can_push(m)) { mpush(m, &s); continue; }
          }
          if (mlast(m).x < c->x1) {
RTL_AfterX1_Clip:;
            double y_clipped = mfirst(m).y + (c->x1 - mfirst(m).x) * dy_div_dx(psub(mlast(m), mfirst(m)));
            out = (y_clipped >= y_end);
            if (out) break;
            if (y_start < y_clipped) acc_right_s(eb, y_start, y_clipped, sign);
            app_open_at(eb, c->x1, y_clipped);
            goto RTL_AddLine;
          }
          out = (mlast(m).y >= y_end);
          if (out) break;
RTL_AfterX1_Pop:
          if (!can_pop(m)) break;
          mpop(m);
        }
        out <<= SHIFT_X1;
      }
      else if (mfirst(m).x > c->x0) {
        app_open_at(eb, mfirst(m).x, mfirst(m).y);
        for (;;) {
          if (!is_flat(m, &s)) {
            msplit(m, &s);
            if (can_push(m)) { mpush(m, &s); continue; }
          }
RTL_AddLine:
          out = mlast(m).x < c->x0;
          if (out) {
            double y_clipped = mfirst(m).y + (c->x0 - mfirst(m).x) * dy_div_dx(psub(mlast(m), mfirst(m)));
            if (y_clipped <= y_end) {
              y_start = y_clipped;
              app_add_line(eb, c->x0, y_clipped);
              break;
            }
          }
          out = mlast(m).y >= c->y1;
          if (out) {
            double x_clipped = orc_max(mfirst(m).x + (c->y1 - mfirst(m).y) * dx_div_dy(psub(mlast(m), mfirst(m))), c->x0);
            app_add_line(eb, x_clipped, c->y1);
            out = 0;
            break;
          }
          app_add_line(eb, mlast(m).x, mlast(m).y);
          if (!can_pop(m)) break;
          mpop(m);
        }
        app_close(eb);
        out <<= SHIFT_X0;
      }
      else {
        out = CLIP_X0;
      }
    }
  }

  if (out && y_start < y_end) {
    if (out & CLIP_X0) acc_left_s(eb, y_start, y_end, sign);
    else acc_right_s(eb, y_start, y_end, sign);
  }
}

/* ---- monotonic splitting (geometry/bezier_p.h:334-400, 835-920) ---- */

static const double kAfter0 = 1e-40;
static const double kBefore1 = 0.999999999999999889;

/* support/math_p.h:574-593 */
static size_t quad_roots(double* dst, double a, double b, double c, double t_min, double t_max) {
  double d = orc_max(b * b - 4.0 * a * c, 0.0);
  double s = sqrt(d);
  double q = -0.5 * (b + copysign(s, b));
  double t0 = q / a;
  double t1 = c / q;
  double x0 = orc_min(t0, t1);
  double x1 = orc_max(t1, t0);
  dst[0] = x0;
  size_t n = (size_t)((x0 >= t_min) & (x0 <= t_max));
  dst[n] = x1;
  n += (size_t)((x1 > x0) & (x1 >= t_min) & (x1 <= t_max));
  return n;
}

/* Returns the number of pieces written to `out` (points: 1 + 2 * pieces). */
static int split_quad(const OrcPt* c, OrcPt* out) {
  double ts[3];
  size_t n = 0;
  OrcPt num = { c[0].x - c[1].x, c[0].y - c[1].y };
  OrcPt den = { c[0].x - c[1].x * 2.0 + c[2].x, c[0].y - c[1].y * 2.0 + c[2].y };
  OrcPt ex = { num.x / den.x, num.y / den.y };
  double t0 = orc_min(ex.x, ex.y);
  double t1 = orc_max(ex.x, ex.y);
  ts[n] = t0; n += (size_t)((t0 > 0.0) & (t0 < 1.0));
  ts[n] = t1; n += (size_t)((t1 > orc_max(t0, 0.0)) & (t1 < 1.0));
  if (n == 0) {
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
    return 1;
  }
  ts[n++] = 1.0;

  OrcPt v1 = psub(c[1], c[0]);
  OrcPt v2 = psub(c[2], c[1]);
  OrcPt qa = psub(v2, v1);
  OrcPt qb = { v1.x + v1.x, v1.y + v1.y };
  OrcPt qc = c[0];
  OrcPt last = c[2];

  out[0] = c[0];
  size_t i = 0;
  double t_cut = 0.0;
  int pieces = 0;
  do {
    double t = ts[i];
    double dt = (t - t_cut) * 0.5;
    OrcPt cp = { (qa.x * (t * 2.0) + qb.x) * dt, (qa.y * (t * 2.0) + qb.y) * dt };
    OrcPt tp = { (qa.x * t + qb.x) * t + qc.x, (qa.y * t + qb.y) * t + qc.y };
    if (++i == n) tp = last;
    out[1].x = tp.x - cp.x; out[1].y = tp.y - cp.y;
    out[2] = tp;
    out += 2;
    pieces++;
    t_cut = t;
  } while (i != n);
  return pieces;
}

static int split_cubic(const OrcPt* c, OrcPt* out) {
  double ts[8];
  size_t n = 0;

  OrcPt v1 = psub(c[1], c[0]);
  OrcPt v2 = psub(c[2], c[1]);
  OrcPt v3 = psub(c[3], c[2]);

  OrcPt ca = { v3.x - v2.x - v2.x + v1.x, v3.y - v2.y - v2.y + v1.y };
  OrcPt cb = { 3.0 * (v2.x - v1.x), 3.0 * (v2.y - v1.y) };
  OrcPt cc = { 3.0 * v1.x, 3.0 * v1.y };
  OrcPt cd = c[0];

  double q0 = cross(cb, ca);
  double q1 = cross(cc, ca);
  double q2 = cross(cc, cb);

  double t_cusp = (q1 / q0) * -0.5;
  ts[n] = t_cusp; n += (size_t)((t_cusp > 0.0) & (t_cusp < 1.0));
  n += quad_roots(ts + n, q0 * 6.0, q1 * 6.0, q2 * 2.0, kAfter0, kBefore1);

  OrcPt da = { 3.0 * (v3.x - v2.x - v2.x + v1.x), 3.0 * (v3.y - v2.y - v2.y + v1.y) };
  OrcPt db = { 6.0 * (v2.x - v1.x), 6.0 * (v2.y - v1.y) };
  OrcPt dc = { 3.0 * v1.x, 3.0 * v1.y };
  n += quad_roots(ts + n, da.x, db.x, dc.x, kAfter0, kBefore1);
  n += quad_roots(ts + n, da.y, db.y, dc.y, kAfter0, kBefore1);

  if (n == 0) {
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
    return 1;
  }

  /* insertion sort, ascending (support/algorithm_p.h:133-137) */
  for (size_t m = 1; m < n; m++)
    for (size_t l = m; l > 0 && ts[l - 1] > ts[l]; l--) { double t = ts[l - 1]; ts[l - 1] = ts[l]; ts[l] = t; }
  ts[n++] = 1.0;

  out[0] = c[0];
  OrcPt last = c[3];
  size_t i = 0;
  double t_cut = 0.0;
  int pieces = 0;
  do {
    double tv = ts[i++];
    if (tv == t_cut) continue;
    const double k1Div3 = 1.0 / 3.0;
    double dt = (tv - t_cut) * k1Div3;
    OrcPt tp = { ((ca.x * tv + cb.x) * tv + cc.x) * tv + cd.x, ((ca.y * tv + cb.y) * tv + cc.y) * tv + cd.y };
    if (i == n) tp = last;
    OrcPt cp1 = { ((ca.x * (t_cut * 3.0) + cb.x * 2.0) * t_cut + cc.x) * dt, ((ca.y * (t_cut * 3.0) + cb.y * 2.0) * t_cut + cc.y) * dt };
    OrcPt cp2 = { ((ca.x * (tv * 3.0) + cb.x * 2.0) * tv + cc.x) * dt, ((ca.y * (tv * 3.0) + cb.y * 2.0) * tv + cc.y) * dt };
    out[1].x = out[0].x + cp1.x; out[1].y = out[0].y + cp1.y;
    out[2].x = tp.x - cp2.x; out[2].y = tp.y - cp2.y;
    out[3] = tp;
    out += 3;
    pieces++;
    t_cut = tv;
  } while (i != n);
  return pieces;
}

/* quad_to / cubic_to for a single source curve (edgebuilder_p.h:1618-1890). */
static void eb_curve(EB* eb, const OrcPt* c, int npts) {
  /* A NaN or infinite control point has no defined flattening in the
   * reference (it subdivides to its recursion limit or trips its own
   * assertions, edgebuilder_p.h:2506); like the product, fail the fill. */
  for (int i = 0; i < npts; i++) {
    if (!isfinite(c[i].x) || !isfinite(c[i].y)) { eb->invalid = 1; return; }
  }
  uint32_t common = 0xF, any = 0;
  for (int i = 0; i < npts; i++) { uint32_t f = fxy(eb, c[i]); common &= f; any |= f; }
  if (common) {
    if (common & CLIP_Y0) return;
    if (common & CLIP_Y1) return;
    double y0 = orc_clamp(c[0].y, eb->clip.y0, eb->clip.y1);
    double y1 = orc_clamp(c[npts - 1].y, eb->clip.y0, eb->clip.y1);
    if (common & CLIP_X0) acc_left(eb, y0, y1);
    else acc_right(eb, y0, y1);
    return;
  }

  OrcPt spline[8 * 3 + 1];
  int pieces = npts == 3 ? split_quad(c, spline) : split_cubic(c, spline);
  Mono m;
  m.n = npts;
  m.tol_sq = eb->tol_sq;
  const OrcPt* sp = spline;
  for (int k = 0; k < pieces; k++, sp += npts - 1) {
    uint32_t sign = sp[0].y > sp[npts - 1].y;
    if (any) flatten_unsafe(eb, &m, sp, sign);
    else flatten_safe(eb, &m, sp, sign);
  }
}

/* ---- sources (edgebuilder_p.h:57-286, 998-1070) ---- */

static inline OrcPt xform(const OrcMat* m, int tt, OrcPt p) {
  OrcPt r;
  if (tt <= ORC_TT_SCALE) { r.x = p.x * m->m00 + m->m20; r.y = p.y * m->m11 + m->m21; }
  else { r = orc_mat_map(m, p.x, p.y); }
  return r;
}

static void eb_init(EB* eb, OrcEdges* e, const OrcBox* clip, double tol_sq) {
  memset(eb, 0, sizeof(*eb));
  eb->out = e;
  eb->clip = *clip;
  eb->clip_ix0 = orc_trunc(clip->x0);
  eb->clip_ix1 = orc_trunc(clip->x1);
  eb->tol_sq = tol_sq;
}

enum { CMD_MOVE = 0, CMD_ON = 1, CMD_QUAD = 2, CMD_CONIC = 3, CMD_CUBIC = 4, CMD_CLOSE = 5 };

int orc_edges_add_path(OrcEdges* e, const uint8_t* cmds, const OrcPt* vtx, size_t n,
                       const OrcMat* m, int tt, const OrcBox* clip, double tol_sq) {
  EB eb;
  eb_init(&eb, e, clip, tol_sq);
  size_t i = 0;
  for (;;) {
    /* begin(): find the next move-to */
    while (i < n && cmds[i] != CMD_MOVE) i++;
    if (i >= n) break;
    OrcPt start = xform(m, tt, vtx[i]);
    OrcPt a = start;
    i++;
    for (;;) {
      if (i < n && cmds[i] == CMD_ON) {
        OrcPt b = xform(m, tt, vtx[i]);
        i++;
        eb_line(&eb, a, b);
        if (eb.invalid) return ORC_ERR_INVALID_GEOMETRY;
        a = b;
      }
      else if (i + 2 <= n && cmds[i] == CMD_QUAD) {
        OrcPt c[3] = { a, xform(m, tt, vtx[i]), xform(m, tt, vtx[i + 1]) };
        i += 2;
        eb_curve(&eb, c, 3);
        if (eb.invalid) return ORC_ERR_INVALID_GEOMETRY;
        a = c[2];
      }
      else if (i + 3 <= n && cmds[i] == CMD_CUBIC) {
        OrcPt c[4] = { a, xform(m, tt, vtx[i]), xform(m, tt, vtx[i + 1]), xform(m, tt, vtx[i + 2]) };
        i += 3;
        eb_curve(&eb, c, 4);
        if (eb.invalid) return ORC_ERR_INVALID_GEOMETRY;
        a = c[3];
      }
      else {
        /* close the figure: fills are always closed */
        eb_line(&eb, a, start);
        if (eb.invalid) return ORC_ERR_INVALID_GEOMETRY;
        break;
      }
    }
  }
  return 0;
}

int orc_edges_add_poly(OrcEdges* e, const OrcPt* vtx, size_t n,
                       const OrcMat* m, int tt, const OrcBox* clip, double tol_sq) {
  EB eb;
  eb_init(&eb, e, clip, tol_sq);
  if (n == 0) return 0;
  OrcPt start = xform(m, tt, vtx[0]);
  OrcPt a = start;
  for (size_t i = 1; i < n; i++) {
    OrcPt b = xform(m, tt, vtx[i]);
    eb_line(&eb, a, b);
    if (eb.invalid) return ORC_ERR_INVALID_GEOMETRY;
    a = b;
  }
  eb_line(&eb, a, start);
  return eb.invalid ? ORC_ERR_INVALID_GEOMETRY : 0;
}
