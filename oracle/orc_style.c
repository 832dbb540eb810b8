/* CPU oracle -- numeric helpers, matrices, style setup, comp-op simplification.
 *
 * TEST INFRASTRUCTURE (see orc_core.h).  Restates
 *   support/math_p.h:329-414         (nearby/round/floor conversions)
 *   core/matrix.cpp:133-210,292-320,442-462, core/matrix_p.h:48-64
 *   pixelops/scalar_p.h:122-156      (premultiply)
 *   pixelops/interpolation.cpp:19-156 (gradient LUT)
 *   core/gradient.cpp:202-287        (LUT size, solid/opaque detection)
 *   pipeline/pipedefs.cpp:162-671    (linear/radial/conic/affine-pattern setup)
 *   core/compopsimplifyimpl_p.h      (comp-op simplification, 5 public ops and
 *                                     the ops they reduce to)
 */
#include <math.h>
#include <string.h>

#include "orc_core.h"

int orc_nearby_to_int(double x) {
  if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
  return (int)rint(x);
}

int orc_round_to_int(double x) {
  int y = orc_nearby_to_int(x);
  return y + ((double)y - x == -0.5);
}

int64_t orc_floor_to_int64(double x) {
  int64_t y = (int64_t)x;
  return y - (int64_t)((double)y > x);
}

/* ---- matrices ---- */

int orc_mat_type(const OrcMat* m) {
  uint32_t msk = ((uint32_t)(m->m00 != 0.0) << 3) | ((uint32_t)(m->m01 != 0.0) << 2) |
                 ((uint32_t)(m->m10 != 0.0) << 1) | ((uint32_t)(m->m11 != 0.0) << 0);
  /* valid combinations: see core/matrix.cpp:158-174 */
  const uint32_t valid = (1u << 3) | (1u << 6) | (1u << 7) | (1u << 9) | (1u << 11) |
                         (1u << 12) | (1u << 13) | (1u << 14) | (1u << 15);
  double d = m->m00 * m->m11 - m->m01 * m->m10;
  if (!((1u << msk) & valid) || !isfinite(d) || !isfinite(m->m20) || !isfinite(m->m21))
    return ORC_TT_INVALID;
  if (msk != 9u) return msk == 6u ? ORC_TT_SWAP : ORC_TT_AFFINE;
  if (!((m->m00 == 1.0) & (m->m11 == 1.0))) return ORC_TT_SCALE;
  if (!((m->m20 == 0.0) & (m->m21 == 0.0))) return ORC_TT_TRANSLATE;
  return ORC_TT_IDENTITY;
}

int orc_mat_invert(OrcMat* dst, const OrcMat* s) {
  double d = s->m00 * s->m11 - s->m01 * s->m10;
  if (d == 0.0 || !isfinite(d)) return -1;
  double t00 = s->m11, t01 = -s->m01, t10 = -s->m10, t11 = s->m00;
  t00 /= d; t01 /= d; t10 /= d; t11 /= d;
  double t20 = -(s->m20 * t00 + s->m21 * t10);
  double t21 = -(s->m20 * t01 + s->m21 * t11);
  dst->m00 = t00; dst->m01 = t01; dst->m10 = t10; dst->m11 = t11; dst->m20 = t20; dst->m21 = t21;
  return 0;
}

void orc_mat_multiply(OrcMat* dst, const OrcMat* a, const OrcMat* b) {
  OrcMat r;
  r.m00 = a->m00 * b->m00 + a->m01 * b->m10;
  r.m01 = a->m00 * b->m01 + a->m01 * b->m11;
  r.m10 = a->m10 * b->m00 + a->m11 * b->m10;
  r.m11 = a->m10 * b->m01 + a->m11 * b->m11;
  r.m20 = a->m20 * b->m00 + a->m21 * b->m10 + b->m20;
  r.m21 = a->m20 * b->m01 + a->m21 * b->m11 + b->m21;
  *dst = r;
}

OrcPt orc_mat_map(const OrcMat* m, double x, double y) {
  OrcPt r = { x * m->m00 + y * m->m10 + m->m20, x * m->m01 + y * m->m11 + m->m21 };
  return r;
}

/* core/matrix.cpp:292-320 (BL_TRANSFORM_OP_ROTATE_PT) */
static void mat_rotate_pt(OrcMat* a, double angle, double px, double py) {
  double as = sin(angle), ac = cos(angle);
  double t00 = as * a->m10 + ac * a->m00;
  double t01 = as * a->m11 + ac * a->m01;
  double t10 = ac * a->m10 - as * a->m00;
  double t11 = ac * a->m11 - as * a->m01;
  double tx = px - ac * px + as * py;
  double ty = py - as * px - ac * py;
  double t20 = tx * a->m00 + ty * a->m10 + a->m20;
  double t21 = tx * a->m01 + ty * a->m11 + a->m21;
  a->m20 = t20; a->m21 = t21;
  a->m00 = t00; a->m01 = t01; a->m10 = t10; a->m11 = t11;
}

/* ---- pixels ---- */

static inline uint32_t udiv255(uint32_t x) { return ((x + 0x80u) * 0x101u) >> 16; }

uint32_t orc_premultiply(uint32_t v) {
  uint32_t a = v >> 24;
  uint32_t r = udiv255(((v >> 16) & 0xFFu) * a);
  uint32_t g = udiv255(((v >> 8) & 0xFFu) * a);
  uint32_t b = udiv255((v & 0xFFu) * a);
  return (a << 24) | (r << 16) | (g << 8) | b;
}

static inline uint32_t rgba32_from_rgba64(uint64_t v) {
  return (uint32_t)(((v >> 56) & 0xFFu) << 24 | ((v >> 40) & 0xFFu) << 16 | ((v >> 24) & 0xFFu) << 8 | ((v >> 8) & 0xFFu));
}

/* Gradient LUT as the reference computes it on x86 hosts with AVX2, where
 * PixelOps::funcs.interpolate_prgb32 resolves to interpolate_prgb32_avx2
 * (pixelops/funcs.cpp:31-35, pixelops/interpolation_avx2.cpp:18-176).  Each
 * channel is the high byte of the 16-bit stop colour; a span of i >= 2 cells
 * steps a 9.23 fixed-point value by trunc(delta * (2^23 / i)) computed in
 * double precision, and every entry is premultiplied with div255. */
static inline uint32_t premul_channels(const uint32_t c[4]) {
  /* c = {b, g, r, a} */
  uint32_t a = c[3];
  uint32_t b = udiv255(c[0] * a), g = udiv255(c[1] * a), r = udiv255(c[2] * a);
  uint32_t aa = udiv255(255u * a);
  return (aa << 24) | (r << 16) | (g << 8) | b;
}

static inline void hi_bytes(uint64_t rgba64, uint32_t out[4]) {
  out[0] = (uint32_t)(rgba64 >> 8) & 0xFFu;   /* b */
  out[1] = (uint32_t)(rgba64 >> 24) & 0xFFu;  /* g */
  out[2] = (uint32_t)(rgba64 >> 40) & 0xFFu;  /* r */
  out[3] = (uint32_t)(rgba64 >> 56) & 0xFFu;  /* a */
}

/* udiv65535 (pixelops/scalar_p.h:37) */
static uint32_t div65535(uint32_t x) { return ((x + 0x8000u) + ((x + 0x8000u) >> 16)) >> 16; }

/* cvt_prgb64_8888_from_argb64_8888 (pixelops/scalar_p.h:179-222) */
static uint64_t premul64(uint64_t c) {
  uint32_t a = (uint32_t)(c >> 48);
  uint64_t r = div65535((uint32_t)((c >> 32) & 0xFFFFu) * a);
  uint64_t g = div65535((uint32_t)((c >> 16) & 0xFFFFu) * a);
  uint64_t b = div65535((uint32_t)(c & 0xFFFFu) * a);
  return ((uint64_t)a << 48) | (r << 32) | (g << 16) | b;
}

/* interpolate_prgb64 (pixelops/interpolation.cpp:161-275): the 64-bit table
 * of dithered gradients, span by span. */
void orc_interpolate_lut64(uint64_t* d_ptr, uint32_t d_size, const OrcStop* s_ptr, size_t s_size) {
  const uint32_t shift = 15, mask = 0xFFFFu << 15;
  uint64_t first = premul64(s_ptr[0].rgba64);
  uint64_t c0 = s_ptr[0].rgba64;
  double f_width = (double)((int32_t)(d_size - 1u) << 8);
  uint32_t p0 = 0;
  uint32_t written = 0;   /* entries 0 .. written - 1 hold their span's values */
  if (s_size == 1) {
    for (uint32_t e = 0; e < d_size; e++) d_ptr[e] = first;
    return;
  }
  for (size_t k = 0; k < s_size; k++) {
    uint64_t c1 = s_ptr[k].rgba64;
    uint32_t p1 = (uint32_t)orc_round_to_int(s_ptr[k].offset * f_width);
    uint32_t at = p0 >> 8, count = (p1 >> 8) - (p0 >> 8);
    if (count == 0) c0 = c1;
    p0 = p1;
    d_ptr[at] = premul64(c0);
    if (c0 == c1) {
      for (uint32_t j = 1; j <= count; j++) d_ptr[at + j] = d_ptr[at];
    }
    else if (count) {
      uint32_t pos[4], inc[4];
      for (int ch = 0; ch < 4; ch++) {          /* b, g, r, a */
        uint32_t a0 = (uint32_t)(((c0 >> (16 * ch)) & 0xFFFFu) << shift) & mask;
        uint32_t a1 = (uint32_t)(((c1 >> (16 * ch)) & 0xFFFFu) << shift) & mask;
        inc[ch] = (uint32_t)((int32_t)(a1 - a0) / (int32_t)count);
        pos[ch] = a0 + (1u << (shift - 1));
      }
      int opaque = ((c0 & c1) >> 48) == 0xFFFFu;
      for (uint32_t j = 1; j <= count; j++) {
        for (int ch = 0; ch < 4; ch++) pos[ch] += inc[ch];
        uint64_t v;
        if (opaque) {
          v = 0xFFFF000000000000ull;
          for (int ch = 0; ch < 3; ch++) v |= (uint64_t)((pos[ch] & mask) >> shift) << (16 * ch);
        }
        else {
          uint32_t ca = pos[3] >> shift;
          v = (uint64_t)ca << 48;
          for (int ch = 0; ch < 3; ch++) v |= (uint64_t)div65535((pos[ch] >> shift) * ca) << (16 * ch);
        }
        d_ptr[at + j] = v;
      }
    }
    written = at + count + 1u;
    c0 = c1;
  }
  for (uint32_t e = written; e < d_size; e++) d_ptr[e] = premul64(c0);
  d_ptr[0] = first;
}

void orc_interpolate_lut(uint32_t* d_ptr, uint32_t d_size, const OrcStop* s_ptr, size_t s_size) {
  uint32_t* span = d_ptr;
  uint32_t i;
  uint64_t c0 = s_ptr[0].rgba64, c1;
  uint32_t u0 = 0, u1;
  size_t s_index = (size_t)(s_ptr[0].offset == 0.0 && s_size > 1);
  double f_width = (double)((int32_t)(--d_size) << 8);

  do {
    c1 = s_ptr[s_index].rgba64;
    u1 = (uint32_t)orc_round_to_int(s_ptr[s_index].offset * f_width);
    span = d_ptr + (u0 >> 8);
    i = (u1 >> 8) - (u0 >> 8);
    u0 = u1;

    if (i <= 1) {
      uint32_t a[4], b[4];
      hi_bytes(c0, a);
      hi_bytes(c1, b);
      c0 = c1;
      *span++ = premul_channels(a);
      if (i == 0) continue;
      *span++ = premul_channels(b);
    }
    else {
      uint32_t n = i + 1;
      double scale = (double)(1 << 23) / (double)(int)i;
      uint32_t a[4], b[4];
      hi_bytes(c0, a);
      hi_bytes(c1, b);
      int32_t cx[4], dx[4];
      for (int k = 0; k < 4; k++) {
        int32_t d32 = (int32_t)b[k] - (int32_t)a[k];
        dx[k] = (int32_t)((double)d32 * scale);
        cx[k] = (int32_t)((uint32_t)a[k] << 23) + (1 << 22);
      }
      for (uint32_t j = 0; j < n; j++) {
        uint32_t px[4];
        for (int k = 0; k < 4; k++) {
          uint32_t v = (uint32_t)cx[k] + (uint32_t)dx[k] * j;
          px[k] = v >> 23;
          if (px[k] > 255u) px[k] = 255u;
        }
        *span++ = premul_channels(px);
      }
      c0 = c1;
    }
  } while (++s_index < s_size);

  i = (uint32_t)((size_t)((d_ptr + d_size + 1) - span));
  uint32_t last[4], first[4];
  hi_bytes(c0, last);
  hi_bytes(s_ptr[0].rgba64, first);
  uint32_t last_px = premul_channels(last);
  while (i--) *span++ = last_px;
  d_ptr[0] = premul_channels(first);
}

/* core/gradient.cpp:202-287 */
uint32_t orc_gradient_info(const OrcStop* stops, size_t n, int* solid, uint32_t* format) {
  const uint32_t F_NOT_ONE = 1, F_NOT_ZERO = 2, F_TRANSITION = 4;
  uint32_t flags = 0, lut_size = 0;
  *solid = 1;
  *format = ORC_FMT_NONE;
  if (n == 0) return 0;
  uint64_t prev = stops[0].rgba64 & 0xFF00FF00FF00FF00ull;
  if (prev < 0xFF00000000000000ull) flags |= F_NOT_ONE;
  if (prev > 0x00FFFFFFFFFFFFFFull) flags |= F_NOT_ZERO;
  for (size_t i = 1; i < n; i++) {
    uint64_t v = stops[i].rgba64 & 0xFF00FF00FF00FF00ull;
    if (v == prev) continue;
    flags |= F_TRANSITION;
    if (v < 0xFF00000000000000ull) flags |= F_NOT_ONE;
    if (v > 0x00FFFFFFFFFFFFFFull) flags |= F_NOT_ZERO;
    prev = v;
  }
  if (!(flags & F_NOT_ZERO)) flags &= ~F_TRANSITION;
  if (!(flags & F_TRANSITION)) {
    lut_size = 256;
  }
  else {
    switch (n) {
      case 1: lut_size = 256; break;
      case 2: lut_size = (stops[1].offset - stops[0].offset >= 0.998) ? 256 : 512; break;
      case 3: lut_size = (stops[0].offset <= 0.002 && stops[1].offset == 0.5 && stops[2].offset >= 0.998) ? 512 : 1024; break;
      default: lut_size = 1024; break;
    }
  }
  *solid = (flags & F_TRANSITION) ? 0 : 1;
  *format = (flags & F_NOT_ONE) ? ORC_FMT_PRGB32 : ORC_FMT_FRGB32;
  return lut_size;
}

/* pipedefs.cpp:386-434 */
int orc_init_linear(OrcFetch* f, const double v[4], uint32_t extend, const OrcMat* t) {
  OrcMat inv;
  if (orc_mat_invert(&inv, t) != 0) return -1;
  uint32_t lut_size = f->lut_size;
  f->type = ORC_FETCH_LINEAR;
  f->extend = extend;
  f->maxi = extend == 2 ? lut_size * 2u - 1u : lut_size - 1u;
  f->rori = extend == 2 ? f->maxi : 0u;

  double ax = v[2] - v[0], ay = v[3] - v[1];
  double dist = ax * ax + ay * ay;
  OrcPt mp = orc_mat_map(t, v[0], v[1]);
  OrcPt o = { 0.5 - mp.x, 0.5 - mp.y };
  double dt = ax * inv.m00 + ay * inv.m01;
  double dy = ax * inv.m10 + ay * inv.m11;
  double scale = (double)(int64_t)((uint64_t)lut_size << 32) / dist;
  double offset = o.x * dt + o.y * dy;
  dt *= scale;
  dy *= scale;
  offset *= scale;
  f->lin_dy = (uint64_t)orc_floor_to_int64(dy);
  f->lin_dt = (uint64_t)orc_floor_to_int64(dt);
  f->lin_pt0 = (uint64_t)orc_floor_to_int64(offset);
  return 0;
}

/* pipedefs.cpp:480-576 */
int orc_init_radial(OrcFetch* f, const double v[6], uint32_t extend, const OrcMat* t) {
  OrcMat inv;
  if (orc_mat_invert(&inv, t) != 0) return -1;
  uint32_t lut_size = f->lut_size;
  f->type = ORC_FETCH_RADIAL;
  f->extend = extend;
  f->maxi = extend == 2 ? lut_size * 2u - 1u : lut_size - 1u;
  f->rori = extend == 2 ? f->maxi : 0u;

  OrcPt cp = { v[0], v[1] };
  OrcPt fp = { v[2], v[3] };
  double cr = v[4], fr = v[5];
  OrcPt dp = { cp.x - fp.x, cp.y - fp.y };
  double dr = cr - fr;
  double sq_dx_plus_dy = dp.x * dp.x + dp.y * dp.y;
  double dx_plus_dy = sqrt(sq_dx_plus_dy);
  double dist_from_border = fabs(dx_plus_dy - dr);
  const double dist_limit = 0.5;
  if (dist_from_border < dist_limit) {
    OrcPt dp0 = { (dp.x * (dr - dist_limit)) / dx_plus_dy, (dp.y * (dr - dist_limit)) / dx_plus_dy };
    OrcPt dp1 = { (dp.x * (dr + dist_limit)) / dx_plus_dy, (dp.y * (dr + dist_limit)) / dx_plus_dy };
    double dp0_dist = fabs(dp0.x * dp0.x + dp0.y * dp0.y - sq_dx_plus_dy);
    double dp1_dist = fabs(dp1.x * dp1.x + dp1.y * dp1.y - sq_dx_plus_dy);
    dp = (dp0_dist < dp1_dist) ? dp0 : dp1;
    fp.x = cp.x - dp.x; fp.y = cp.y - dp.y;
    sq_dx_plus_dy = dp.x * dp.x + dp.y * dp.y;
  }

  double a = dr * dr - sq_dx_plus_dy;
  double sq_fr = fr * fr;
  double scale = (double)lut_size;
  double xx = inv.m00, xy = inv.m01, yx = inv.m10, yy = inv.m11;
  OrcPt tp = { inv.m20 + (xx + xy) * 0.5 - fp.x, inv.m21 + (yx + yy) * 0.5 - fp.y };

  f->rad_tx = tp.x; f->rad_ty = tp.y; f->rad_yx = yx; f->rad_yy = yy;
  double a_mul_4 = a * 4.0;
  double inv2a = (scale * 0.5) / a;
  double sq_inv2a = inv2a * inv2a;
  f->rad_amul4 = a_mul_4; f->rad_inv2a = inv2a; f->rad_sq_inv2a = sq_inv2a; f->rad_sq_fr = sq_fr;

  double sq_xx_plus_sq_yx = xx * xx + xy * xy;
  double b0 = 2.0 * (dr * fr + tp.x * dp.x + tp.y * dp.y);
  double bx = 2.0 * (dp.x * xx + dp.y * xy);
  double by = 2.0 * (dp.x * yx + dp.y * yy);
  f->rad_b0 = -b0;
  f->rad_by = -by;
  double bx_mul_2 = bx * 2.0;
  double sq_bx = bx * bx;
  double dd0 = sq_bx + bx_mul_2 * b0 + a_mul_4 * (sq_xx_plus_sq_yx + 2.0 * (tp.x * xx + tp.y * xy));
  double ddy = bx_mul_2 * by + a_mul_4 * (2.0 * (xx * yx + yy * xy));
  double ddd_half = (sq_bx + a_mul_4 * sq_xx_plus_sq_yx);
  double ddd_half_inv = ddd_half * sq_inv2a;
  f->rad_dd0 = dd0 - ddd_half;
  f->rad_ddy = ddy;
  f->rad_bd = (float)(-bx * inv2a);
  f->rad_ddd = (float)ddd_half_inv;
  return 0;
}

/* pipedefs.cpp:610-671 */
int orc_init_conic(OrcFetch* f, const double v[4], const OrcMat* t) {
  static const double q256[4] = { 4.071421038552e+1, -1.311160794048e+1, 6.017670215625, -1.623253505085 };
  const double kPI_MUL_2 = 6.28318530717958647692;
  uint32_t lut_size = f->lut_size;
  f->type = ORC_FETCH_CONIC;

  OrcPt mc = orc_mat_map(t, v[0], v[1]);
  OrcPt c = { 0.5 - mc.x, 0.5 - mc.y };
  double angle = v[2];
  double repeat = v[3];
  OrcPt vv = { 1.0 * t->m00 + 0.0 * t->m10, 1.0 * t->m01 + 0.0 * t->m11 };
  double matrix_angle = atan2(vv.y, vv.x);

  OrcMat upd = *t;
  mat_rotate_pt(&upd, -matrix_angle, c.x, c.y);

  angle += matrix_angle;
  double q = angle / -kPI_MUL_2;
  double off = q - floor(q);
  if (off != 0.0) off = -1.0 + off;

  OrcMat inv;
  if (orc_mat_invert(&inv, &upd) != 0) return -1;

  f->con_tx = c.x * inv.m00 + c.y * inv.m10;
  f->con_ty = c.x * inv.m01 + c.y * inv.m11;
  f->con_yx = inv.m10;
  f->con_yy = inv.m11;

  double lut_d = (double)(int)lut_size;
  double rep = lut_d * repeat;
  double qs = rep / 256.0;
  for (int k = 0; k < 4; k++) f->con_q[k] = (float)(q256[k] * qs);
  f->con_n1 = (float)rep;
  f->con_n2 = (float)(rep * 0.5);
  f->con_n4 = (float)(rep * 0.25);
  f->con_offset = (float)(off * rep - 0.5);
  f->con_xx = (float)inv.m00;
  f->con_maxi = INT32_MAX;
  f->con_rori = (int32_t)(lut_size - 1u);
  return 0;
}

/* pipedefs.cpp:60-111 (init_pattern_tx_ty) and :119-160 (init_pattern_fx_fy):
 * a pattern whose inverse transform is a pure translation samples texels at
 * integer offsets, or (bilinear, fractional offset) blends 2x2 texels with
 * fixed weights.  Simple extend modes apply to both axes. */
static void init_pattern_fx_fy(OrcFetch* f, uint32_t extend, uint32_t quality, int64_t tx64, int64_t ty64) {
  uint32_t wx = (uint32_t)(tx64 & 0xFF);
  uint32_t wy = (uint32_t)(ty64 & 0xFF);
  int tx = -(int)(tx64 >> 8);
  int ty = -(int)(ty64 >> 8);
  int fractional = (wx | wy) != 0;
  f->sp_wa = f->sp_wb = f->sp_wc = f->sp_wd = 0;
  if (fractional) {
    if (quality == 0) {
      tx -= (wx >= 128);
      ty -= (wy >= 128);
      fractional = 0;
    }
    else {
      f->sp_wa = ((wy) * (wx)) >> 8;
      f->sp_wb = ((wy) * (256 - wx) + 255) >> 8;
      f->sp_wc = ((256 - wy) * (wx)) >> 8;
      f->sp_wd = ((256 - wy) * (256 - wx) + 255) >> 8;
      tx--;
      ty--;
    }
  }
  f->type = fractional ? ORC_FETCH_PATTERN_FXFY : ORC_FETCH_PATTERN_ALIGNED;
  uint32_t ex = extend, ey = extend;
  if (f->img_w <= 1) ex = 0;
  if (f->img_h <= 1) ey = 0;
  int rx = 0, ry = 0;
  if (ex >= 1) {
    rx = (int)f->img_w << (ex == 2);
    if ((unsigned)tx >= (unsigned)rx) tx %= rx;
    if (tx < 0) tx += rx;
  }
  if (ey >= 1) {
    ry = (int)f->img_h << (ey == 2);
    if ((unsigned)ty >= (unsigned)ry) ty %= ry;
    if (ty < 0) ty += ry;
  }
  f->sp_tx = tx; f->sp_ty = ty;
  f->sp_rx = rx; f->sp_ry = ry;
}

/* pipedefs.cpp:162-381 */
int orc_init_pattern(OrcFetch* f, uint32_t extend, uint32_t quality, const OrcMat* t) {
  OrcMat inv;
  if (orc_mat_invert(&inv, t) != 0) return -1;
  int tw = (int)f->img_w, th = (int)f->img_h;
  /* pipedefs.cpp:170-175 tests the width only; a zero-height blit area makes
   * the reference's pattern fetch divide by zero, so an empty height is a
   * NOP here (no reference output to follow). */
  if (tw == 0 || th == 0) return -1;

  double xx = inv.m00, xy = inv.m01, yx = inv.m10, yy = inv.m11;
  const double eps = 1e-14;   /* Math::epsilon<double>, support/math_p.h:42 */
  if (fabs(xx - 1.0) <= eps && fabs(xy) <= eps && fabs(yx) <= eps && fabs(yy - 1.0) <= eps) {
    init_pattern_fx_fy(f, extend, quality, orc_floor_to_int64(-inv.m20 * 256.0), orc_floor_to_int64(-inv.m21 * 256.0));
    return 0;
  }

  f->type = quality == 0 ? ORC_FETCH_PATTERN_AFFINE_NN : ORC_FETCH_PATTERN_AFFINE_BI;
  uint32_t extend_x = extend, extend_y = extend;   /* simple extend modes only */

  double tx = inv.m20, ty = inv.m21;
  tx += 0.5 * (xx + yx);
  ty += 0.5 * (xy + yy);
  const double fp_scale = 4294967296.0;

  int ox = INT32_MAX, oy = INT32_MAX, rx = 0, ry = 0;
  f->pa_minx = 0; f->pa_miny = 0;
  f->pa_maxx = tw - 1; f->pa_maxy = th - 1;
  f->pa_corx = tw - 1; f->pa_cory = th - 1;

  if (extend_x != 0) {
    f->pa_minx = INT32_MIN;
    if (extend_x == 1) f->pa_corx = 0;
    ox = tw;
    if (extend_x == 2) tw *= 2;
    if (xx < 0.0) {
      xx = -xx; yx = -yx; tx = (double)tw - tx;
      if (extend_x == 1) { ox = 0; f->pa_corx = f->pa_maxx; f->pa_maxx = -1; }
    }
    ox--;
  }
  if (extend_y != 0) {
    f->pa_miny = INT32_MIN;
    if (extend_y == 1) f->pa_cory = 0;
    oy = th;
    if (extend_y == 2) th *= 2;
    if (xy < 0.0) {
      xy = -xy; yy = -yy; ty = (double)th - ty;
      if (extend_y == 1) { oy = 0; f->pa_cory = f->pa_maxy; f->pa_maxy = -1; }
    }
    oy--;
  }

  if (quality != 0) { tx -= 0.5; ty -= 0.5; }

  double tw_d = (double)tw, th_d = (double)th;
  if (extend_x == 0) tw_d = 2147483647.0;
  else { tx = fmod(tx, tw_d); rx = tw; if (xx >= tw_d) xx = fmod(xx, tw_d); }
  if (extend_y == 0) th_d = 2147483647.0;
  else { ty = fmod(ty, th_d); ry = th; if (xy >= th_d) xy = fmod(xy, th_d); }

  xx *= fp_scale; xy *= fp_scale; yx *= fp_scale; yy *= fp_scale; tx *= fp_scale; ty *= fp_scale;

  double all_min = orc_min(orc_min(yy, yx), orc_min(xy, xx));
  double all_max = orc_max(orc_max(xx, xy), orc_max(yx, yy));
  all_min = orc_min(all_min, orc_min(ty, tx));
  all_max = orc_max(orc_max(tx, ty), all_max);
  if (all_min >= (double)(INT64_MIN + 1) && all_max <= (double)INT64_MAX) {
    f->pa_xx = (uint64_t)orc_floor_to_int64(xx);
    f->pa_xy = (uint64_t)orc_floor_to_int64(xy);
    f->pa_yx = (uint64_t)orc_floor_to_int64(yx);
    f->pa_yy = (uint64_t)orc_floor_to_int64(yy);
    f->pa_tx = (uint64_t)orc_floor_to_int64(tx);
    f->pa_ty = (uint64_t)orc_floor_to_int64(ty);
  }
  else {
    f->pa_xx = f->pa_xy = f->pa_yx = f->pa_yy = f->pa_tx = f->pa_ty = 0;
  }
  f->pa_rx = rx; f->pa_ry = ry;
  f->pa_ox = ox; f->pa_oy = oy;
  /* The fetcher normalizes with the integer pattern period (tw/th as int). */
  f->pa_tw = (int32_t)tw_d == 2147483647 ? 2147483647 : (int32_t)tw;
  f->pa_th = (int32_t)th_d == 2147483647 ? 2147483647 : (int32_t)th;
  return 0;
}

/* ---- comp-op simplification (core/compopsimplifyimpl_p.h) ---- */

typedef struct { uint32_t op, d, s, solid; } Simp;

static Simp mk(uint32_t op, uint32_t d, uint32_t s) { Simp r = { op, d, s, ORC_SOLID_NONE }; return r; }
static Simp mk_solid(uint32_t op, uint32_t d, uint32_t s, uint32_t id) { Simp r = { op, d, s, id }; return r; }
static Simp nop(void) { Simp r = { ORC_OP_NOP, 0, 0, ORC_SOLID_NONE }; return r; }

enum { P = ORC_FMT_PRGB32, X = ORC_FMT_XRGB32, A = ORC_FMT_A8, F = ORC_FMT_FRGB32, Z = ORC_FMT_ZERO32 };

static Simp s_clear(uint32_t d, uint32_t s) {
  if (d == P) return mk_solid(ORC_OP_SRC_COPY, P, P, ORC_SOLID_TRANSPARENT);
  if (d == X) return mk_solid(ORC_OP_SRC_COPY, P, P, ORC_SOLID_OPAQUE_BLACK);
  if (d == A) return mk_solid(ORC_OP_SRC_COPY, A, P, ORC_SOLID_TRANSPARENT);
  return mk(ORC_OP_CLEAR, d, s);
}
static Simp s_src_copy(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return mk(ORC_OP_SRC_COPY, P, P);
  if (d == P && s == F) return mk(ORC_OP_SRC_COPY, P, P);
  if (d == X && s == P) return mk(ORC_OP_SRC_COPY, P, X);
  if (d == X && s == Z) return mk(ORC_OP_SRC_COPY, P, X);
  if (d == X && s == X) return mk(ORC_OP_SRC_COPY, P, X);
  if (d == X && s == F) return mk(ORC_OP_SRC_COPY, P, P);
  if (d == A && s == Z) return s_clear(A, Z);
  if (d == A && s == X) return mk_solid(ORC_OP_SRC_COPY, d, P, ORC_SOLID_OPAQUE_WHITE);
  if (d == A && s == F) return mk_solid(ORC_OP_SRC_COPY, d, P, ORC_SOLID_OPAQUE_WHITE);
  return mk(ORC_OP_SRC_COPY, d, s);
}
static Simp s_src_over(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == X) return s_src_copy(P, X);
  if (d == P && s == F) return s_src_copy(P, F);
  if (d == X && s == P) return s_src_over(P, P);
  if (d == X && s == Z) return nop();
  if (d == X && s == X) return s_src_copy(P, X);
  if (d == X && s == F) return s_src_copy(P, F);
  if (d == A && s == Z) return nop();
  if (d == A && s == X) return s_src_copy(A, X);
  if (d == A && s == F) return s_src_copy(A, F);
  return mk(ORC_OP_SRC_OVER, d, s);
}
static Simp s_dst_over(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_dst_over(P, P);
  if (d == X) return nop();
  if (d == A) return s_src_over(d, s);
  return mk(ORC_OP_DST_OVER, d, s);
}
static Simp s_src_in(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return s_src_in(P, P);
  if (d == P && s == F) return s_src_in(P, P);
  if (d == X) return s_src_copy(d, s);
  if (d == A && s == Z) return s_clear(d, s);
  if (d == A && s == X) return nop();
  if (d == A && s == F) return nop();
  return mk(ORC_OP_SRC_IN, d, s);
}
static Simp s_dst_in(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return s_src_copy(d, s);
  if (d == P && s == X) return nop();
  if (d == P && s == F) return nop();
  if (d == X && s == P) return s_dst_in(P, P);
  if (d == X && s == Z) return s_dst_in(P, F);
  if (d == X && s == X) return nop();
  if (d == X && s == F) return nop();
  if (d == A) return s_src_in(d, s);
  return mk(ORC_OP_DST_IN, d, s);
}
static Simp s_plus(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_plus(P, P);
  if (d == X && s == P) return s_plus(P, P);
  if (d == X && s == Z) return nop();
  if (d == X && s == X) return s_plus(P, P);
  if (d == X && s == F) return s_plus(P, P);
  if (d == A && s == Z) return nop();
  if (d == A && s == X) return mk_solid(ORC_OP_PLUS, d, P, ORC_SOLID_OPAQUE_WHITE);
  if (d == A && s == F) return mk_solid(ORC_OP_PLUS, d, P, ORC_SOLID_OPAQUE_WHITE);
  return mk(ORC_OP_PLUS, d, s);
}
static Simp s_modulate(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return mk_solid(ORC_OP_SRC_COPY, P, P, ORC_SOLID_TRANSPARENT);
  if (d == P && s == F) return s_modulate(P, P);
  if (d == X && s == Z) return mk_solid(ORC_OP_SRC_COPY, P, P, ORC_SOLID_OPAQUE_BLACK);
  if (d == X && s == F) return s_modulate(X, P);
  if (d == X && s == X) return s_modulate(X, P);
  if (d == A || s == A) return s_dst_in(d, s);
  return mk(ORC_OP_MODULATE, d, s);
}
static Simp s_multiply(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_multiply(P, X);
  if (d == X && s == Z) return nop();
  if (d == X && s == F) return s_modulate(X, X);
  if (d == X && s == X) return s_modulate(X, X);
  if (d == A || s == A) return s_dst_over(d, s);
  return mk(ORC_OP_MULTIPLY, d, s);
}
static Simp s_screen(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_screen(P, P);
  if (d == X && s == P) return s_screen(P, P);
  if (d == X && s == Z) return nop();
  if (d == X && s == F) return s_screen(P, P);
  if (d == X && s == X) return s_screen(P, X);
  if (d == A || s == A) return s_src_over(d, s);
  return mk(ORC_OP_SCREEN, d, s);
}

uint32_t orc_simplify(uint32_t op, uint32_t d, uint32_t s, uint32_t* out_d, uint32_t* out_s, uint32_t* solid_id) {
  Simp r;
  switch (op) {
    case ORC_OP_SRC_OVER: r = s_src_over(d, s); break;
    case ORC_OP_SRC_COPY: r = s_src_copy(d, s); break;
    case ORC_OP_PLUS: r = s_plus(d, s); break;
    case ORC_OP_MULTIPLY: r = s_multiply(d, s); break;
    case ORC_OP_SCREEN: r = s_screen(d, s); break;
    case ORC_OP_CLEAR: r = s_clear(d, s); break;
    default: r = mk(op, d, s); break;
  }
  *out_d = r.d;
  *out_s = r.s;
  *solid_id = r.solid;
  return r.op;
}
