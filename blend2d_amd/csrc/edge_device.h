// Device edge builder: transform, clip, monotonic split and flattening of one
// source segment (line, quad or cubic) into 24.8 fixed-point edge segments.
//
// Follows the reference edge builder (raster/edgebuilder_p.h, geometry/
// bezier_p.h) operation by operation in double precision; kernels using it
// are compiled with -ffp-contract=off so no multiply-add is fused, as the
// reference's Math::madd is unfused (support/math_p.h:134-135).  The
// reference builds monotonic chains per path; the rasterizer consumes each
// consecutive point pair of a chain independently, so every thread here
// handles one source segment and emits point pairs directly.  Border lines
// (edgebuilder_p.h:2546-2622) are emitted per accumulate call; vertical lines
// at one x rasterize additively over adjacent y intervals, so this equals the
// reference's merged accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"

namespace b2amd {
namespace edge {

struct P { double x, y; };

__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return dmin(hi, dmax(lo, v)); }
__device__ __forceinline__ int trunc_i(double x) { return (int)x; }
__device__ __forceinline__ P psub(P a, P b) { return P{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ P pmid(P a, P b) { return P{(a.x + b.x) * 0.5, (a.y + b.y) * 0.5}; }
__device__ __forceinline__ double cross(P a, P b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ double mag_sq(P v) { return v.x * v.x + v.y * v.y; }

enum : uint32_t { CX0 = 1, CY0 = 2, CX1 = 4, CY1 = 8 };

struct Clip {
  double x0, y0, x1, y1;
  int32_t ix0, ix1;
};

__device__ __forceinline__ uint32_t fx0(const Clip& c, P p) { return (uint32_t)(!(p.x >= c.x0)) * CX0; }
__device__ __forceinline__ uint32_t fx1(const Clip& c, P p) { return (uint32_t)(!(p.x <= c.x1)) * CX1; }
__device__ __forceinline__ uint32_t fy0(const Clip& c, P p) { return (uint32_t)(!(p.y >= c.y0)) * CY0; }
__device__ __forceinline__ uint32_t fy1(const Clip& c, P p) { return (uint32_t)(!(p.y <= c.y1)) * CY1; }
__device__ __forceinline__ uint32_t fxf(const Clip& c, P p) { return fx0(c, p) | fx1(c, p); }
__device__ __forceinline__ uint32_t fyf(const Clip& c, P p) { return fy0(c, p) | fy1(c, p); }
__device__ __forceinline__ uint32_t fxy(const Clip& c, P p) { return fxf(c, p) | fyf(c, p); }


// Emitter receives stored segments: y0 < y1, sign = 1 when the source went up.
template<typename Emit>
struct Builder {
  Emit& out;
  Clip c;
  double tol_sq;
  int32_t ax, ay;       // appender: last point
  uint32_t asign;
  uint32_t invalid;     // CmdFail bits
  uint32_t n_out;       // segments emitted so far

  __device__ Builder(Emit& e, const Clip& clip, double t)
      : out(e), c(clip), tol_sq(t), ax(0), ay(0), asign(0), invalid(0), n_out(0) {}

  // Past the segment budget nothing more is emitted and is_flat() holds for
  // every piece, so the flattening loops only unwind their stacks.  Both the
  // counting and the writing pass stop at the same segment.
  __device__ __forceinline__ void emit(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign) {
    if (n_out >= kMaxItemSegments) { invalid |= kFailTooManySegments; return; }
    n_out++;
    out.push(x0, y0, x1, y1, sign);
  }
  __device__ __forceinline__ void stored(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign) {
    if (y0 != y1) emit(x0, y0, x1, y1, sign);
  }
  __device__ __forceinline__ void dir(int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    if (y0 < y1) emit(x0, y0, x1, y1, 0u);
    else if (y0 > y1) emit(x1, y1, x0, y0, 1u);
  }
  __device__ __forceinline__ void line_seg(double x0, double y0, double x1, double y1) {
    dir(trunc_i(x0), trunc_i(y0), trunc_i(x1), trunc_i(y1));
  }
  __device__ __forceinline__ void acc_left(double y0, double y1) { dir(c.ix0, trunc_i(y0), c.ix0, trunc_i(y1)); }
  __device__ __forceinline__ void acc_right(double y0, double y1) { dir(c.ix1, trunc_i(y0), c.ix1, trunc_i(y1)); }
  __device__ __forceinline__ void acc_left_s(double y0, double y1, uint32_t s) { if (s) acc_left(y1, y0); else acc_left(y0, y1); }
  __device__ __forceinline__ void acc_right_s(double y0, double y1, uint32_t s) { if (s) acc_right(y1, y0); else acc_right(y0, y1); }
  __device__ __forceinline__ void open_at(double x, double y) { ax = trunc_i(x); ay = trunc_i(y); }
  __device__ __forceinline__ void add_line(double x, double y) {
    int32_t nx = trunc_i(x), ny = trunc_i(y);
    stored(ax, ay, nx, ny, asign);
    ax = nx; ay = ny;
  }

  // -- lines (edgebuilder_p.h:1123-1605) --------------------------------
  __device__ void clip_end(P p, P a, P b, uint32_t bf) {
    P d = psub(b, a);
    P q{c.x1, c.y1};
    switch (bf) {
      case CX0 | CY0: q.x = c.x0; [[fallthrough]];
      case CX1 | CY0:
        q.y = a.y + (q.x - a.x) * d.y / d.x;
        if (q.y >= c.y0) break;
        [[fallthrough]];
      case CY0:
        q.y = c.y0;
        q.x = a.x + (q.y - a.y) * d.x / d.y;
        break;
      case CX0 | CY1: q.x = c.x0; [[fallthrough]];
      case CX1 | CY1:
        q.y = a.y + (q.x - a.x) * d.y / d.x;
        if (q.y <= c.y1) break;
        [[fallthrough]];
      case CY1:
        q.y = c.y1;
        q.x = a.x + (q.y - a.y) * d.x / d.y;
        break;
      case CX0: q.x = c.x0; [[fallthrough]];
      case CX1:
        q.y = a.y + (q.x - a.x) * d.y / d.x;
        break;
      default:
        invalid |= kFailInvalidGeometry;   // edgebuilder_p.h:1584-1586
        return;
    }
    line_seg(p.x, p.y, q.x, q.y);
    double cby = dclamp(b.y, c.y0, c.y1);
    if (q.y != cby) {
      if (q.x == c.x0) acc_left(q.y, cby);
      else acc_right(q.y, cby);
    }
  }

  __device__ void line(P a, P b) {
    uint32_t af = fxy(c, a);
    uint32_t bf = fxy(c, b);
    if (!af) {
      if (!bf) { dir(trunc_i(a.x), trunc_i(a.y), trunc_i(b.x), trunc_i(b.y)); return; }
      clip_end(a, a, b, bf);
      return;
    }
    double by0, by1;
    if (af & CY0) {
      if (!(c.y0 < b.y)) return;
      by0 = c.y0;
      uint32_t cm = af & bf;
      if (cm) {
        by1 = dmin(c.y1, b.y);
        if (cm & CX0) acc_left(by0, by1); else acc_right(by0, by1);
        return;
      }
    }
    else if (af & CY1) {
      if (!(c.y1 > b.y)) return;
      by0 = c.y1;
      uint32_t cm = af & bf;
      if (cm) {
        by1 = dmax(c.y0, b.y);
        if (cm & CX0) acc_left(by0, by1); else acc_right(by0, by1);
        return;
      }
    }
    else if (af & CX0) {
      by0 = dclamp(a.y, c.y0, c.y1);
      if (!(c.x0 < b.x)) {
        by1 = dclamp(b.y, c.y0, c.y1);
        if (by0 != by1) acc_left(by0, by1);
        return;
      }
    }
    else {
      by0 = dclamp(a.y, c.y0, c.y1);
      if (!(c.x1 > b.x)) {
        by1 = dclamp(b.y, c.y0, c.y1);
        if (by0 != by1) acc_right(by0, by1);
        return;
      }
    }

    P d = psub(b, a);
    P p{c.x1, c.y1};
    switch (af) {
      case CX0 | CY0: p.x = c.x0; [[fallthrough]];
      case CX1 | CY0:
        p.y = a.y + (p.x - a.x) * d.y / d.x;
        af = fyf(c, p);
        if (p.y >= c.y0) break;
        [[fallthrough]];
      case CY0:
        p.y = c.y0;
        p.x = a.x + (p.y - a.y) * d.x / d.y;
        af = fxf(c, p);
        break;
      case CX0 | CY1: p.x = c.x0; [[fallthrough]];
      case CX1 | CY1:
        p.y = a.y + (p.x - a.x) * d.y / d.x;
        af = fyf(c, p);
        if (p.y <= c.y1) break;
        [[fallthrough]];
      case CY1:
        p.y = c.y1;
        p.x = a.x + (p.y - a.y) * d.x / d.y;
        af = fxf(c, p);
        break;
      case CX0: p.x = c.x0; [[fallthrough]];
      case CX1:
        p.y = a.y + (p.x - a.x) * d.y / d.x;
        af = fyf(c, p);
        break;
      default:
        invalid |= kFailInvalidGeometry;   // edgebuilder_p.h:1494-1496
        return;
    }
    if (af) {
      by1 = dclamp(b.y, c.y0, c.y1);
      if (p.x <= c.x0) acc_left(by0, by1);
      else if (p.x >= c.x1) acc_right(by0, by1);
      return;
    }
    by1 = dclamp(p.y, c.y0, c.y1);
    if (by0 != by1) {
      if (p.x <= c.x0) acc_left(by0, by1); else acc_right(by0, by1);
    }
    if (!bf) { dir(trunc_i(p.x), trunc_i(p.y), trunc_i(b.x), trunc_i(b.y)); return; }
    clip_end(p, a, b, bf);
  }

  // -- monotonic curves (edgebuilder_p.h:418-768, 2029-2445) --------------
  static constexpr int kLimit = 32;

  struct Mono {
    int n;
    P p[4];
    int sp;
  };
  struct Step { P p01, p12, p23, p012, p123, p0123; };

  __device__ __forceinline__ static void mbegin(Mono& m, const P* src, uint32_t sign) {
    m.sp = 0;
    for (int i = 0; i < m.n; i++) m.p[i] = sign ? src[m.n - 1 - i] : src[i];
  }
  __device__ __forceinline__ static P mfirst(const Mono& m) { return m.p[0]; }
  __device__ __forceinline__ static P mlast(const Mono& m) { return m.n == 3 ? m.p[2] : m.p[3]; }
  __device__ __forceinline__ static bool is_ltr(const Mono& m) { return mfirst(m).x < mlast(m).x; }

  __device__ __forceinline__ static void bound_ltr(Mono& m) {
    P f = mfirst(m), l = mlast(m);
    for (int i = 1; i < m.n - 1; i++) { m.p[i].x = dclamp(m.p[i].x, f.x, l.x); m.p[i].y = dclamp(m.p[i].y, f.y, l.y); }
  }
  __device__ __forceinline__ static void bound_rtl(Mono& m) {
    P f = mfirst(m), l = mlast(m);
    for (int i = 1; i < m.n - 1; i++) { m.p[i].x = dclamp(m.p[i].x, l.x, f.x); m.p[i].y = dclamp(m.p[i].y, f.y, l.y); }
  }
  __device__ __forceinline__ bool is_flat(const Mono& m) const {
    if (invalid & kFailTooManySegments) return true;
    double value, limit;
    if (m.n == 3) {
      P v1 = psub(m.p[1], m.p[0]);
      P v2 = psub(m.p[2], m.p[0]);
      double d = cross(v2, v1);
      value = d * d;
      limit = tol_sq * mag_sq(v2);
    }
    else {
      P v = psub(m.p[3], m.p[0]);
      double c1 = cross(v, psub(m.p[1], m.p[0]));
      double c2 = cross(v, psub(m.p[2], m.p[0]));
      value = dmax(c1 * c1, c2 * c2);
      limit = tol_sq * mag_sq(v);
    }
    return value <= limit;
  }
  __device__ __forceinline__ static void msplit(const Mono& m, Step& s) {
    if (m.n == 3) {
      s.p01 = pmid(m.p[0], m.p[1]);
      s.p12 = pmid(m.p[1], m.p[2]);
      s.p012 = pmid(s.p01, s.p12);
    }
    else {
      s.p01 = pmid(m.p[0], m.p[1]);
      s.p12 = pmid(m.p[1], m.p[2]);
      s.p23 = pmid(m.p[2], m.p[3]);
      s.p012 = pmid(s.p01, s.p12);
      s.p123 = pmid(s.p12, s.p23);
      s.p0123 = pmid(s.p012, s.p123);
    }
  }
  __device__ __forceinline__ static P step_mid(const Mono& m, const Step& s) { return m.n == 3 ? s.p012 : s.p0123; }
  __device__ __forceinline__ static void mpush(Mono& m, const Step& s, P* stack) {
    P* st = stack + m.sp * 4;
    if (m.n == 3) {
      st[0] = s.p012; st[1] = s.p12; st[2] = m.p[2];
      m.p[1] = s.p01; m.p[2] = s.p012;
    }
    else {
      st[0] = s.p0123; st[1] = s.p123; st[2] = s.p23; st[3] = m.p[3];
      m.p[1] = s.p01; m.p[2] = s.p012; m.p[3] = s.p0123;
    }
    m.sp++;
  }
  __device__ __forceinline__ static void mdiscard(Mono& m, const Step& s) {
    if (m.n == 3) { m.p[0] = s.p012; m.p[1] = s.p12; }
    else { m.p[0] = s.p0123; m.p[1] = s.p123; m.p[2] = s.p23; }
  }
  __device__ __forceinline__ static void mpop(Mono& m, const P* stack) {
    m.sp--;
    const P* st = stack + m.sp * 4;
    for (int i = 0; i < m.n; i++) m.p[i] = st[i];
  }

  __device__ void flatten_safe(Mono& m, const P* src, uint32_t sign, P* stack) {
    mbegin(m, src, sign);
    asign = sign;
    if (is_ltr(m)) bound_ltr(m); else bound_rtl(m);
    open_at(mfirst(m).x, mfirst(m).y);
    for (;;) {
      if (!is_flat(m)) {
        if (m.sp != kLimit) { Step s; msplit(m, s); mpush(m, s, stack); continue; }
      }
      add_line(mlast(m).x, mlast(m).y);
      if (m.sp == 0) break;
      mpop(m, stack);
    }
  }

  __device__ void flatten_unsafe(Mono& m, const P* src, uint32_t sign, P* stack) {
    mbegin(m, src, sign);
    asign = sign;
    double y_start = mfirst(m).y;
    double y_end = dmin(mlast(m).y, c.y1);
    if ((y_start >= y_end) | (y_end <= c.y0)) return;

    const double kDeltaLimit = 0.00390625;
    double x_delta = fabs(mfirst(m).x - mlast(m).x);
    uint32_t out = 0;
    Step s;

    if (x_delta <= kDeltaLimit) {
      y_start = dmax(y_start, c.y0);
      double x_min = dmin(mfirst(m).x, mlast(m).x);
      double x_max = dmax(mfirst(m).x, mlast(m).x);
      if (x_max <= c.x0) acc_left_s(y_start, y_end, sign);
      else if (x_min >= c.x1) acc_right_s(y_start, y_end, sign);
      else { open_at(mfirst(m).x, y_start); add_line(mlast(m).x, y_end); }
      return;
    }
    else if (is_ltr(m)) {
      bound_ltr(m);
      if (y_start < c.y0) {
        y_start = c.y0;
        for (;;) {
          out = (mfirst(m).x >= c.x1);
          if (out) break;
          if (!is_flat(m)) {
            msplit(m, s);
            if (step_mid(m, s).y <= c.y0) { mdiscard(m, s); continue; }
            if (m.sp != kLimit) { mpush(m, s, stack); continue; }
          }
          if (mlast(m).y > c.y0) {
            out = mlast(m).x < c.x0;
            if (out) goto LTR_BeforeX0_Pop;
            double x_clipped = mfirst(m).x + (c.y0 - mfirst(m).y) * (psub(mlast(m), mfirst(m)).x / psub(mlast(m), mfirst(m)).y);
            if (x_clipped <= c.x0) goto LTR_BeforeX0_Clip;
            out = (x_clipped >= c.x1);
            if (out) break;
            open_at(x_clipped, c.y0);
            goto LTR_AddLine;
          }
          if (m.sp == 0) break;
          mpop(m, stack);
        }
        out <<= 2;
      }
      else if (y_start < c.y1) {
        if (mfirst(m).x < c.x0) {
          for (;;) {
            out = (mfirst(m).y >= c.y1);
            if (out) break;
            if (!is_flat(m)) {
              msplit(m, s);
              if (step_mid(m, s).x <= c.x0) { mdiscard(m, s); continue; }
              if (m.sp != kLimit) { mpush(m, s, stack); continue; }
            }
            if (mlast(m).x > c.x0) {
LTR_BeforeX0_Clip:
              {
                double y_clipped = mfirst(m).y + (c.x0 - mfirst(m).x) * (psub(mlast(m), mfirst(m)).y / psub(mlast(m), mfirst(m)).x);
                out = (y_clipped >= y_end);
                if (out) break;
                if (y_start < y_clipped) acc_left_s(y_start, y_clipped, sign);
                open_at(c.x0, y_clipped);
              }
              goto LTR_AddLine;
            }
            out = (mlast(m).y >= y_end);
            if (out) break;
LTR_BeforeX0_Pop:
            if (m.sp == 0) break;
            mpop(m, stack);
          }
          out <<= 0;
        }
        else if (mfirst(m).x < c.x1) {
          open_at(mfirst(m).x, mfirst(m).y);
          for (;;) {
            if (!is_flat(m)) {
              msplit(m, s);
              if (m.sp != kLimit) { mpush(m, s, stack); continue; }
            }
LTR_AddLine:
            out = mlast(m).x > c.x1;
            if (out) {
              double y_clipped = mfirst(m).y + (c.x1 - mfirst(m).x) * (psub(mlast(m), mfirst(m)).y / psub(mlast(m), mfirst(m)).x);
              if (y_clipped <= y_end) {
                y_start = y_clipped;
                add_line(c.x1, y_clipped);
                break;
              }
            }
            out = mlast(m).y >= c.y1;
            if (out) {
              double x_clipped = dmin(mfirst(m).x + (c.y1 - mfirst(m).y) * (psub(mlast(m), mfirst(m)).x / psub(mlast(m), mfirst(m)).y), c.x1);
              add_line(x_clipped, c.y1);
              out = 0;
              break;
            }
            add_line(mlast(m).x, mlast(m).y);
            if (m.sp == 0) break;
            mpop(m, stack);
          }
          out <<= 2;
        }
        else {
          out = CX1;
        }
      }
    }
    else {
      bound_rtl(m);
      if (y_start < c.y0) {
        y_start = c.y0;
        for (;;) {
          out = (mfirst(m).x <= c.x0);
          if (out) break;
          if (!is_flat(m)) {
            msplit(m, s);
            if (step_mid(m, s).y <= c.y0) { mdiscard(m, s); continue; }
            if (m.sp != kLimit) { mpush(m, s, stack); continue; }
          }
          if (mlast(m).y > c.y0) {
            out = mlast(m).x > c.x1;
            if (out) goto RTL_AfterX1_Pop;
            double x_clipped = mfirst(m).x + (c.y0 - mfirst(m).y) * (psub(mlast(m), mfirst(m)).x / psub(mlast(m), mfirst(m)).y);
            if (x_clipped >= c.x1) goto RTL_AfterX1_Clip;
            out = (x_clipped <= c.x0);
            if (out) break;
            open_at(x_clipped, c.y0);
            goto RTL_AddLine;
          }
          if (m.sp == 0) break;
          mpop(m, stack);
        }
        out <<= 0;
      }
      else if (y_start < c.y1) {
        if (mfirst(m).x > c.x1) {
          for (;;) {
            out = (mfirst(m).y >= c.y1);
            if (out) break;
            if (!is_flat(m)) {
              msplit(m, s);
              if (step_mid(m, s).x >= c.x1) { mdiscard(m, s); continue; }
              if (m.sp != kLimit) { mpush(m, s, stack); continue; }
            }
            if (mlast(m).x < c.x1) {
RTL_AfterX1_Clip:
              {
                double y_clipped = mfirst(m).y + (c.x1 - mfirst(m).x) * (psub(mlast(m), mfirst(m)).y / psub(mlast(m), mfirst(m)).x);
                out = (y_clipped >= y_end);
                if (out) break;
                if (y_start < y_clipped) acc_right_s(y_start, y_clipped, sign);
                open_at(c.x1, y_clipped);
              }
              goto RTL_AddLine;
            }
            out = (mlast(m).y >= y_end);
            if (out) break;
RTL_AfterX1_Pop:
            if (m.sp == 0) break;
            mpop(m, stack);
          }
          out <<= 2;
        }
        else if (mfirst(m).x > c.x0) {
          open_at(mfirst(m).x, mfirst(m).y);
          for (;;) {
            if (!is_flat(m)) {
              msplit(m, s);
              if (m.sp != kLimit) { mpush(m, s, stack); continue; }
            }
RTL_AddLine:
            out = mlast(m).x < c.x0;
            if (out) {
              double y_clipped = mfirst(m).y + (c.x0 - mfirst(m).x) * (psub(mlast(m), mfirst(m)).y / psub(mlast(m), mfirst(m)).x);
              if (y_clipped <= y_end) {
                y_start = y_clipped;
                add_line(c.x0, y_clipped);
                break;
              }
            }
            out = mlast(m).y >= c.y1;
            if (out) {
              double x_clipped = dmax(mfirst(m).x + (c.y1 - mfirst(m).y) * (psub(mlast(m), mfirst(m)).x / psub(mlast(m), mfirst(m)).y), c.x0);
              add_line(x_clipped, c.y1);
              out = 0;
              break;
            }
            add_line(mlast(m).x, mlast(m).y);
            if (m.sp == 0) break;
            mpop(m, stack);
          }
          out <<= 0;
        }
        else {
          out = CX0;
        }
      }
    }

    if (out && y_start < y_end) {
      if (out & CX0) acc_left_s(y_start, y_end, sign);
      else acc_right_s(y_start, y_end, sign);
    }
  }

  // -- monotonic splitting (geometry/bezier_p.h:334-400, 835-920) ---------
  __device__ __forceinline__ static int quad_roots(double* dst, double a, double b, double cc, double t_min, double t_max) {
    double d = dmax(b * b - 4.0 * a * cc, 0.0);
    double s = sqrt(d);
    double q = -0.5 * (b + copysign(s, b));
    double t0 = q / a;
    double t1 = cc / q;
    double x0 = dmin(t0, t1);
    double x1 = dmax(t1, t0);
    dst[0] = x0;
    int n = (int)((x0 >= t_min) & (x0 <= t_max));
    dst[n] = x1;
    n += (int)((x1 > x0) & (x1 >= t_min) & (x1 <= t_max));
    return n;
  }

  __device__ int split_quad(const P* cv, P* out) {
    double ts[3];
    int n = 0;
    P num{cv[0].x - cv[1].x, cv[0].y - cv[1].y};
    P den{cv[0].x - cv[1].x * 2.0 + cv[2].x, cv[0].y - cv[1].y * 2.0 + cv[2].y};
    P ex{num.x / den.x, num.y / den.y};
    double t0 = dmin(ex.x, ex.y);
    double t1 = dmax(ex.x, ex.y);
    ts[n] = t0; n += (int)((t0 > 0.0) & (t0 < 1.0));
    ts[n] = t1; n += (int)((t1 > dmax(t0, 0.0)) & (t1 < 1.0));
    if (n == 0) { out[0] = cv[0]; out[1] = cv[1]; out[2] = cv[2]; return 1; }
    ts[n++] = 1.0;
    P v1 = psub(cv[1], cv[0]);
    P v2 = psub(cv[2], cv[1]);
    P qa = psub(v2, v1);
    P qb{v1.x + v1.x, v1.y + v1.y};
    P qc = cv[0];
    P last = cv[2];
    out[0] = cv[0];
    int i = 0, pieces = 0;
    double t_cut = 0.0;
    do {
      double t = ts[i];
      double dt = (t - t_cut) * 0.5;
      P cp{(qa.x * (t * 2.0) + qb.x) * dt, (qa.y * (t * 2.0) + qb.y) * dt};
      P tp{(qa.x * t + qb.x) * t + qc.x, (qa.y * t + qb.y) * t + qc.y};
      if (++i == n) tp = last;
      out[1] = P{tp.x - cp.x, tp.y - cp.y};
      out[2] = tp;
      out += 2;
      pieces++;
      t_cut = t;
    } while (i != n);
    return pieces;
  }

  __device__ int split_cubic(const P* cv, P* out) {
    double ts[8];
    int n = 0;
    P v1 = psub(cv[1], cv[0]);
    P v2 = psub(cv[2], cv[1]);
    P v3 = psub(cv[3], cv[2]);
    P ca{v3.x - v2.x - v2.x + v1.x, v3.y - v2.y - v2.y + v1.y};
    P cb{3.0 * (v2.x - v1.x), 3.0 * (v2.y - v1.y)};
    P cc{3.0 * v1.x, 3.0 * v1.y};
    P cd = cv[0];
    double q0 = cross(cb, ca);
    double q1 = cross(cc, ca);
    double q2 = cross(cc, cb);
    double t_cusp = (q1 / q0) * -0.5;
    ts[n] = t_cusp; n += (int)((t_cusp > 0.0) & (t_cusp < 1.0));
    const double kAfter0 = 1e-40;
    const double kBefore1 = 0.999999999999999889;
    n += quad_roots(ts + n, q0 * 6.0, q1 * 6.0, q2 * 2.0, kAfter0, kBefore1);
    P da{3.0 * (v3.x - v2.x - v2.x + v1.x), 3.0 * (v3.y - v2.y - v2.y + v1.y)};
    P db{6.0 * (v2.x - v1.x), 6.0 * (v2.y - v1.y)};
    P dc{3.0 * v1.x, 3.0 * v1.y};
    n += quad_roots(ts + n, da.x, db.x, dc.x, kAfter0, kBefore1);
    n += quad_roots(ts + n, da.y, db.y, dc.y, kAfter0, kBefore1);
    if (n == 0) { out[0] = cv[0]; out[1] = cv[1]; out[2] = cv[2]; out[3] = cv[3]; return 1; }
    for (int m = 1; m < n; m++)
      for (int l = m; l > 0 && ts[l - 1] > ts[l]; l--) { double t = ts[l - 1]; ts[l - 1] = ts[l]; ts[l] = t; }
    ts[n++] = 1.0;
    out[0] = cv[0];
    P last = cv[3];
    int i = 0, pieces = 0;
    double t_cut = 0.0;
    do {
      double tv = ts[i++];
      if (tv == t_cut) continue;
      const double k1Div3 = 1.0 / 3.0;
      double dt = (tv - t_cut) * k1Div3;
      P tp{((ca.x * tv + cb.x) * tv + cc.x) * tv + cd.x, ((ca.y * tv + cb.y) * tv + cc.y) * tv + cd.y};
      if (i == n) tp = last;
      P cp1{((ca.x * (t_cut * 3.0) + cb.x * 2.0) * t_cut + cc.x) * dt, ((ca.y * (t_cut * 3.0) + cb.y * 2.0) * t_cut + cc.y) * dt};
      P cp2{((ca.x * (tv * 3.0) + cb.x * 2.0) * tv + cc.x) * dt, ((ca.y * (tv * 3.0) + cb.y * 2.0) * tv + cc.y) * dt};
      out[1] = P{out[0].x + cp1.x, out[0].y + cp1.y};
      out[2] = P{tp.x - cp2.x, tp.y - cp2.y};
      out[3] = tp;
      out += 3;
      pieces++;
      t_cut = tv;
    } while (i != n);
    return pieces;
  }

  // quad_to / cubic_to for one source curve (edgebuilder_p.h:1618-1890),
  // in two parts.  curve_split: a curve wholly on one side of the clip box
  // emits its border line here and gives 0; otherwise the monotonic pieces
  // are written to `spline` (npts - 1 points apart, end point shared) and
  // their number returned, with `any` set when a control point lies outside
  // the clip box.  curve_piece flattens one piece; pieces are independent.
  __device__ int curve_split(const P* cv, int npts, P* spline, uint32_t* any_out) {
    uint32_t common = 0xFu, any = 0u;
    for (int i = 0; i < npts; i++) { uint32_t f = fxy(c, cv[i]); common &= f; any |= f; }
    *any_out = any;
    if (common) {
      if (common & CY0) return 0;
      if (common & CY1) return 0;
      double y0 = dclamp(cv[0].y, c.y0, c.y1);
      double y1 = dclamp(cv[npts - 1].y, c.y0, c.y1);
      if (common & CX0) acc_left(y0, y1); else acc_right(y0, y1);
      return 0;
    }
    return npts == 3 ? split_quad(cv, spline) : split_cubic(cv, spline);
  }

  __device__ void curve_piece(const P* sp, int npts, uint32_t any, P* stack) {
    Mono m;
    m.n = npts;
    const uint32_t sign = sp[0].y > sp[npts - 1].y;
    if (any) flatten_unsafe(m, sp, sign, stack);
    else flatten_safe(m, sp, sign, stack);
  }

  __device__ void curve(const P* cv, int npts, P* stack) {
    P spline[8 * 3 + 1];
    uint32_t any;
    const int pieces = curve_split(cv, npts, spline, &any);
    const P* sp = spline;
    for (int k = 0; k < pieces; k++, sp += npts - 1) curve_piece(sp, npts, any, stack);
  }
};

}  // namespace edge
}  // namespace b2amd
