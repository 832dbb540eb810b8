// Host-side style setup.  Each function follows the reference routine cited
// above it; operation order is kept so the double-precision results match.
// Compiled without FMA contraction (-ffp-contract=off).
#include "style.h"

#include <math.h>
#include <string.h>

namespace b2amd {

static inline double dmin(double a, double b) { return b < a ? b : a; }
static inline double dmax(double a, double b) { return a < b ? b : a; }

// support/math_p.h:337-343, 369-370, 414
int nearby_to_int(double x) {
  if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
  return (int)rint(x);
}
int round_to_int(double x) {
  int y = nearby_to_int(x);
  return y + ((double)y - x == -0.5);
}
int64_t floor_to_int64(double x) {
  int64_t y = (int64_t)x;
  return y - (int64_t)((double)y > x);
}

// core/matrix.cpp:133-210
int transform_type(const Mat& m) {
  uint32_t msk = ((uint32_t)(m.m00 != 0.0) << 3) | ((uint32_t)(m.m01 != 0.0) << 2) |
                 ((uint32_t)(m.m10 != 0.0) << 1) | ((uint32_t)(m.m11 != 0.0) << 0);
  const uint32_t valid = (1u << 3) | (1u << 6) | (1u << 7) | (1u << 9) | (1u << 11) |
                         (1u << 12) | (1u << 13) | (1u << 14) | (1u << 15);
  double d = m.m00 * m.m11 - m.m01 * m.m10;
  if (!((1u << msk) & valid) || !isfinite(d) || !isfinite(m.m20) || !isfinite(m.m21)) return kTTInvalid;
  if (msk != 9u) return msk == 6u ? kTTSwap : kTTAffine;
  if (!((m.m00 == 1.0) & (m.m11 == 1.0))) return kTTScale;
  if (!((m.m20 == 0.0) & (m.m21 == 0.0))) return kTTTranslate;
  return kTTIdentity;
}

// core/matrix.cpp:442-462
bool invert(Mat& dst, const Mat& s) {
  double d = s.m00 * s.m11 - s.m01 * s.m10;
  if (d == 0.0 || !isfinite(d)) return false;
  double t00 = s.m11, t01 = -s.m01, t10 = -s.m10, t11 = s.m00;
  t00 /= d; t01 /= d; t10 /= d; t11 /= d;
  double t20 = -(s.m20 * t00 + s.m21 * t10);
  double t21 = -(s.m20 * t01 + s.m21 * t11);
  dst = Mat{t00, t01, t10, t11, t20, t21};
  return true;
}

// core/matrix_p.h:57-64
Mat multiply(const Mat& a, const Mat& b) {
  return Mat{a.m00 * b.m00 + a.m01 * b.m10,
             a.m00 * b.m01 + a.m01 * b.m11,
             a.m10 * b.m00 + a.m11 * b.m10,
             a.m10 * b.m01 + a.m11 * b.m11,
             a.m20 * b.m00 + a.m21 * b.m10 + b.m20,
             a.m20 * b.m01 + a.m21 * b.m11 + b.m21};
}

static inline void map_point(const Mat& m, double x, double y, double& ox, double& oy) {
  ox = x * m.m00 + y * m.m10 + m.m20;
  oy = x * m.m01 + y * m.m11 + m.m21;
}

// core/matrix.cpp:292-320 (rotate about a point)
static void rotate_about(Mat& a, double angle, double px, double py) {
  double as = sin(angle), ac = cos(angle);
  double t00 = as * a.m10 + ac * a.m00;
  double t01 = as * a.m11 + ac * a.m01;
  double t10 = ac * a.m10 - as * a.m00;
  double t11 = ac * a.m11 - as * a.m01;
  double tx = px - ac * px + as * py;
  double ty = py - as * px - ac * py;
  double t20 = tx * a.m00 + ty * a.m10 + a.m20;
  double t21 = tx * a.m01 + ty * a.m11 + a.m21;
  a.m20 = t20; a.m21 = t21;
  a.m00 = t00; a.m01 = t01; a.m10 = t10; a.m11 = t11;
}

static inline uint32_t udiv255(uint32_t x) { return ((x + 0x80u) * 0x101u) >> 16; }

// pixelops/scalar_p.h:146-156
uint32_t premultiply(uint32_t v) {
  uint32_t a = v >> 24;
  uint32_t r = udiv255(((v >> 16) & 0xFFu) * a);
  uint32_t g = udiv255(((v >> 8) & 0xFFu) * a);
  uint32_t b = udiv255((v & 0xFFu) * a);
  return (a << 24) | (r << 16) | (g << 8) | b;
}

// core/gradient.cpp:751-800 (stops stay sorted by offset; a third stop at an
// existing offset replaces the second one); colours widen to 16 bits per
// channel by byte replication (core/rgba_p.h:57-65).
void add_stop(std::vector<Stop>& stops, double offset, uint32_t rgba32) {
  if (!(offset >= 0.0 && offset <= 1.0)) return;
  uint64_t v = 0;
  for (int i = 0; i < 4; i++) v |= ((uint64_t)((rgba32 >> (i * 8)) & 0xFFu) * 0x101u) << (i * 16);
  size_t n = stops.size(), i = 0;
  if (n && offset >= stops[0].offset) {
    size_t k = 0;
    while (k + 1 < n && stops[k + 1].offset <= offset) k++;
    i = k;
    if (i > 0 && stops[i - 1].offset == offset) { stops[i] = Stop{offset, v}; return; }
    i++;
  }
  stops.insert(stops.begin() + (ptrdiff_t)i, Stop{offset, v});
}

// core/gradient.cpp:202-287
uint32_t gradient_info(const std::vector<Stop>& stops, bool& solid, uint32_t& format) {
  const uint32_t kNotOne = 1, kNotZero = 2, kTransition = 4;
  uint32_t flags = 0, lut = 0;
  solid = true;
  format = kFxNone;
  size_t n = stops.size();
  if (!n) return 0;
  uint64_t prev = stops[0].rgba64 & 0xFF00FF00FF00FF00ull;
  if (prev < 0xFF00000000000000ull) flags |= kNotOne;
  if (prev > 0x00FFFFFFFFFFFFFFull) flags |= kNotZero;
  for (size_t i = 1; i < n; i++) {
    uint64_t v = stops[i].rgba64 & 0xFF00FF00FF00FF00ull;
    if (v == prev) continue;
    flags |= kTransition;
    if (v < 0xFF00000000000000ull) flags |= kNotOne;
    if (v > 0x00FFFFFFFFFFFFFFull) flags |= kNotZero;
    prev = v;
  }
  if (!(flags & kNotZero)) flags &= ~kTransition;
  if (!(flags & kTransition)) lut = 256;
  else if (n == 1) lut = 256;
  else if (n == 2) lut = (stops[1].offset - stops[0].offset >= 0.998) ? 256 : 512;
  else if (n == 3) lut = (stops[0].offset <= 0.002 && stops[1].offset == 0.5 && stops[2].offset >= 0.998) ? 512 : 1024;
  else lut = 1024;
  solid = !(flags & kTransition);
  format = (flags & kNotOne) ? kFxPRGB32 : kFxFRGB32;
  return lut;
}

// pipeline/pipedefs.cpp:386-434
bool init_linear(DevStyle& f, uint32_t lut_size, const double v[4], uint32_t extend, const Mat& t) {
  Mat inv;
  if (!invert(inv, t)) return false;
  f.kind = kFetchLinear;
  f.extend = extend;
  f.maxi = extend == 2 ? lut_size * 2u - 1u : lut_size - 1u;
  f.rori = extend == 2 ? f.maxi : 0u;
  double ax = v[2] - v[0], ay = v[3] - v[1];
  double dist = ax * ax + ay * ay;
  double mx, my;
  map_point(t, v[0], v[1], mx, my);
  double ox = 0.5 - mx, oy = 0.5 - my;
  double dt = ax * inv.m00 + ay * inv.m01;
  double dy = ax * inv.m10 + ay * inv.m11;
  double scale = (double)(int64_t)((uint64_t)lut_size << 32) / dist;
  double offset = ox * dt + oy * dy;
  dt *= scale;
  dy *= scale;
  offset *= scale;
  f.lin_dy = (uint64_t)floor_to_int64(dy);
  f.lin_dt = (uint64_t)floor_to_int64(dt);
  f.lin_pt0 = (uint64_t)floor_to_int64(offset);
  return true;
}

// pipeline/pipedefs.cpp:480-576
bool init_radial(DevStyle& f, uint32_t lut_size, const double v[6], uint32_t extend, const Mat& t) {
  Mat inv;
  if (!invert(inv, t)) return false;
  f.kind = kFetchRadial;
  f.extend = extend;
  f.maxi = extend == 2 ? lut_size * 2u - 1u : lut_size - 1u;
  f.rori = extend == 2 ? f.maxi : 0u;

  double cpx = v[0], cpy = v[1], fpx = v[2], fpy = v[3], cr = v[4], fr = v[5];
  double dpx = cpx - fpx, dpy = cpy - fpy;
  double dr = cr - fr;
  double sq_dx_plus_dy = dpx * dpx + dpy * dpy;
  double dx_plus_dy = sqrt(sq_dx_plus_dy);
  double dist_from_border = fabs(dx_plus_dy - dr);
  const double dist_limit = 0.5;
  if (dist_from_border < dist_limit) {
    double dp0x = (dpx * (dr - dist_limit)) / dx_plus_dy, dp0y = (dpy * (dr - dist_limit)) / dx_plus_dy;
    double dp1x = (dpx * (dr + dist_limit)) / dx_plus_dy, dp1y = (dpy * (dr + dist_limit)) / dx_plus_dy;
    double d0 = fabs(dp0x * dp0x + dp0y * dp0y - sq_dx_plus_dy);
    double d1 = fabs(dp1x * dp1x + dp1y * dp1y - sq_dx_plus_dy);
    if (d0 < d1) { dpx = dp0x; dpy = dp0y; } else { dpx = dp1x; dpy = dp1y; }
    fpx = cpx - dpx; fpy = cpy - dpy;
    sq_dx_plus_dy = dpx * dpx + dpy * dpy;
  }
  double a = dr * dr - sq_dx_plus_dy;
  double sq_fr = fr * fr;
  double scale = (double)lut_size;
  double xx = inv.m00, xy = inv.m01, yx = inv.m10, yy = inv.m11;
  double tpx = inv.m20 + (xx + xy) * 0.5 - fpx;
  double tpy = inv.m21 + (yx + yy) * 0.5 - fpy;
  f.rad_tx = tpx; f.rad_ty = tpy; f.rad_yx = yx; f.rad_yy = yy;
  double a_mul_4 = a * 4.0;
  double inv2a = (scale * 0.5) / a;
  double sq_inv2a = inv2a * inv2a;
  f.rad_amul4 = a_mul_4; f.rad_inv2a = inv2a; f.rad_sq_inv2a = sq_inv2a; f.rad_sq_fr = sq_fr;
  double sq_xx_plus_sq_yx = xx * xx + xy * xy;
  double b0 = 2.0 * (dr * fr + tpx * dpx + tpy * dpy);
  double bx = 2.0 * (dpx * xx + dpy * xy);
  double by = 2.0 * (dpx * yx + dpy * yy);
  f.rad_b0 = -b0;
  f.rad_by = -by;
  double bx_mul_2 = bx * 2.0;
  double sq_bx = bx * bx;
  double dd0 = sq_bx + bx_mul_2 * b0 + a_mul_4 * (sq_xx_plus_sq_yx + 2.0 * (tpx * xx + tpy * xy));
  double ddy = bx_mul_2 * by + a_mul_4 * (2.0 * (xx * yx + yy * xy));
  double ddd_half = (sq_bx + a_mul_4 * sq_xx_plus_sq_yx);
  double ddd_half_inv = ddd_half * sq_inv2a;
  f.rad_dd0 = dd0 - ddd_half;
  f.rad_ddy = ddy;
  f.rad_bd = (float)(-bx * inv2a);
  f.rad_ddd = (float)ddd_half_inv;
  return true;
}

// pipeline/pipedefs.cpp:610-671 (minimax atan coefficients for a 256-entry table)
bool init_conic(DevStyle& f, uint32_t lut_size, const double v[4], const Mat& t) {
  static const double q256[4] = { 4.071421038552e+1, -1.311160794048e+1, 6.017670215625, -1.623253505085 };
  const double kPI_MUL_2 = 6.28318530717958647692;
  f.kind = kFetchConic;
  double mx, my;
  map_point(t, v[0], v[1], mx, my);
  double cx = 0.5 - mx, cy = 0.5 - my;
  double angle = v[2];
  double repeat = v[3];
  double vx = 1.0 * t.m00 + 0.0 * t.m10;
  double vy = 1.0 * t.m01 + 0.0 * t.m11;
  double matrix_angle = atan2(vy, vx);
  Mat upd = t;
  rotate_about(upd, -matrix_angle, cx, cy);
  angle += matrix_angle;
  double q = angle / -kPI_MUL_2;
  double off = q - floor(q);
  if (off != 0.0) off = -1.0 + off;
  Mat inv;
  if (!invert(inv, upd)) return false;
  f.con_tx = cx * inv.m00 + cy * inv.m10;
  f.con_ty = cx * inv.m01 + cy * inv.m11;
  f.con_yx = inv.m10;
  f.con_yy = inv.m11;
  double lut_d = (double)(int)lut_size;
  double rep = lut_d * repeat;
  double qs = rep / 256.0;
  for (int k = 0; k < 4; k++) f.con_q[k] = (float)(q256[k] * qs);
  f.con_n1 = (float)rep;
  f.con_n2 = (float)(rep * 0.5);
  f.con_n4 = (float)(rep * 0.25);
  f.con_offset = (float)(off * rep - 0.5);
  f.con_xx = (float)inv.m00;
  f.con_maxi = INT32_MAX;
  f.con_rori = (int32_t)(lut_size - 1u);
  return true;
}

// pipeline/pipedefs.cpp:119-160 (init_pattern_fx_fy) and :52-111
// (init_pattern_tx_ty): the 24.8 translation splits into a texel offset and,
// for bilinear quality, four weights of the 2x2 neighbourhood; nearest
// quality rounds the offset instead.  Repeat / reflect keep the offset in
// [0, period); the fractional fetcher always uses the repeat-or-reflect
// column mapping (reflect period 2w, repeat period w).
static void init_pattern_translated(DevStyle& f, uint32_t img_w, uint32_t img_h, uint32_t extend, uint32_t quality,
                                    int64_t tx64, int64_t ty64) {
  const uint32_t wx = (uint32_t)(tx64 & 0xFF), wy = (uint32_t)(ty64 & 0xFF);
  int tx = -(int)(tx64 >> 8), ty = -(int)(ty64 >> 8);
  bool fractional = (wx | wy) != 0;
  f.sp_wa = f.sp_wb = f.sp_wc = f.sp_wd = 0;
  if (fractional) {
    if (quality == 0) {
      tx -= (wx >= 128);
      ty -= (wy >= 128);
      fractional = false;
    }
    else {
      f.sp_wa = (wy * wx) >> 8;                        // texel (x0, y0)
      f.sp_wb = (wy * (256 - wx) + 255) >> 8;          // texel (x1, y0)
      f.sp_wc = ((256 - wy) * wx) >> 8;                // texel (x0, y1)
      f.sp_wd = ((256 - wy) * (256 - wx) + 255) >> 8;  // texel (x1, y1)
      tx--;
      ty--;
    }
  }
  f.kind = fractional ? kFetchPatternFxFy : kFetchPatternAligned;
  const uint32_t ex = img_w <= 1 ? 0u : extend, ey = img_h <= 1 ? 0u : extend;
  int rx = 0, ry = 0;
  if (ex != 0) {
    rx = (int)img_w << (ex == 2);
    if ((unsigned)tx >= (unsigned)rx) tx %= rx;
    if (tx < 0) tx += rx;
  }
  if (ey != 0) {
    ry = (int)img_h << (ey == 2);
    if ((unsigned)ty >= (unsigned)ry) ty %= ry;
    if (ty < 0) ty += ry;
  }
  f.sp_tx = tx; f.sp_ty = ty;
  f.sp_rx = rx; f.sp_ry = ry;
  f.sp_w = (int32_t)img_w; f.sp_h = (int32_t)img_h;
}

// pipeline/pipedefs.cpp:162-381
int init_pattern(DevStyle& f, uint32_t img_w, uint32_t img_h, uint32_t extend, uint32_t quality, const Mat& t) {
  Mat inv;
  if (!invert(inv, t)) return 0;
  int tw = (int)img_w, th = (int)img_h;
  // An empty source renders nothing.  The reference tests only the width;
  // a zero-height blit area reaches its pattern fetch, which divides by the
  // height (SIGFPE in the portable pipeline).
  if (tw == 0 || th == 0) return 0;
  double xx = inv.m00, xy = inv.m01, yx = inv.m10, yy = inv.m11;
  const double eps = 1e-14;   // Math::epsilon<double>, support/math_p.h:42
  if (fabs(xx - 1.0) <= eps && fabs(xy) <= eps && fabs(yx) <= eps && fabs(yy - 1.0) <= eps) {
    init_pattern_translated(f, img_w, img_h, extend, quality, floor_to_int64(-inv.m20 * 256.0),
                            floor_to_int64(-inv.m21 * 256.0));
    return 1;
  }

  f.kind = quality == 0 ? kFetchPatternNN : kFetchPatternBI;
  uint32_t ex = extend, ey = extend;
  double tx = inv.m20, ty = inv.m21;
  tx += 0.5 * (xx + yx);
  ty += 0.5 * (xy + yy);
  const double fp_scale = 4294967296.0;
  int ox = INT32_MAX, oy = INT32_MAX, rx = 0, ry = 0;
  f.pa_minx = 0; f.pa_miny = 0;
  f.pa_maxx = tw - 1; f.pa_maxy = th - 1;
  f.pa_corx = tw - 1; f.pa_cory = th - 1;
  if (ex != 0) {
    f.pa_minx = INT32_MIN;
    if (ex == 1) f.pa_corx = 0;
    ox = tw;
    if (ex == 2) tw *= 2;
    if (xx < 0.0) {
      xx = -xx; yx = -yx; tx = (double)tw - tx;
      if (ex == 1) { ox = 0; f.pa_corx = f.pa_maxx; f.pa_maxx = -1; }
    }
    ox--;
  }
  if (ey != 0) {
    f.pa_miny = INT32_MIN;
    if (ey == 1) f.pa_cory = 0;
    oy = th;
    if (ey == 2) th *= 2;
    if (xy < 0.0) {
      xy = -xy; yy = -yy; ty = (double)th - ty;
      if (ey == 1) { oy = 0; f.pa_cory = f.pa_maxy; f.pa_maxy = -1; }
    }
    oy--;
  }
  if (quality != 0) { tx -= 0.5; ty -= 0.5; }
  double tw_d = (double)tw, th_d = (double)th;
  if (ex == 0) tw_d = 2147483647.0;
  else { tx = fmod(tx, tw_d); rx = tw; if (xx >= tw_d) xx = fmod(xx, tw_d); }
  if (ey == 0) th_d = 2147483647.0;
  else { ty = fmod(ty, th_d); ry = th; if (xy >= th_d) xy = fmod(xy, th_d); }
  xx *= fp_scale; xy *= fp_scale; yx *= fp_scale; yy *= fp_scale; tx *= fp_scale; ty *= fp_scale;
  double all_min = dmin(dmin(yy, yx), dmin(xy, xx));
  double all_max = dmax(dmax(xx, xy), dmax(yx, yy));
  all_min = dmin(all_min, dmin(ty, tx));
  all_max = dmax(dmax(tx, ty), all_max);
  if (all_min >= (double)(INT64_MIN + 1) && all_max <= (double)INT64_MAX) {
    f.pa_xx = (uint64_t)floor_to_int64(xx);
    f.pa_xy = (uint64_t)floor_to_int64(xy);
    f.pa_yx = (uint64_t)floor_to_int64(yx);
    f.pa_yy = (uint64_t)floor_to_int64(yy);
    f.pa_tx = (uint64_t)floor_to_int64(tx);
    f.pa_ty = (uint64_t)floor_to_int64(ty);
  }
  else {
    f.pa_xx = f.pa_xy = f.pa_yx = f.pa_yy = f.pa_tx = f.pa_ty = 0;
  }
  f.pa_rx = rx; f.pa_ry = ry;
  f.pa_ox = ox; f.pa_oy = oy;
  f.pa_tw = (int32_t)tw_d;
  f.pa_th = (int32_t)th_d;
  return 1;
}

// ---- comp-op simplification (core/compopsimplifyimpl_p.h) ----------------

namespace {
constexpr uint32_t P = kFxPRGB32, X = kFxXRGB32, A = kFxA8, F = kFxFRGB32, Z = kFxZERO32;
Simplified mk(uint32_t op, uint32_t d, uint32_t s, uint32_t id = kSolidNone) { return Simplified{op, d, s, id}; }
Simplified nop() { return Simplified{kCompNop, 0, 0, kSolidNone}; }

Simplified s_src_over(uint32_t d, uint32_t s);
Simplified s_src_copy(uint32_t d, uint32_t s);

Simplified s_clear(uint32_t d, uint32_t s) {
  if (d == P) return mk(kCompSrcCopy, P, P, kSolidTransparent);
  if (d == X) return mk(kCompSrcCopy, P, P, kSolidOpaqueBlack);
  if (d == A) return mk(kCompSrcCopy, A, P, kSolidTransparent);
  return mk(kCompClear, d, s);
}
Simplified s_src_copy(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return mk(kCompSrcCopy, P, P);
  if (d == P && s == F) return mk(kCompSrcCopy, P, P);
  if (d == X && s == P) return mk(kCompSrcCopy, P, X);
  if (d == X && s == Z) return mk(kCompSrcCopy, P, X);
  if (d == X && s == X) return mk(kCompSrcCopy, P, X);
  if (d == X && s == F) return mk(kCompSrcCopy, P, P);
  if (d == A && s == Z) return s_clear(A, Z);
  if (d == A && s == X) return mk(kCompSrcCopy, d, P, kSolidOpaqueWhite);
  if (d == A && s == F) return mk(kCompSrcCopy, d, P, kSolidOpaqueWhite);
  return mk(kCompSrcCopy, d, s);
}
Simplified s_src_over(uint32_t d, uint32_t s) {
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
  return mk(kCompSrcOver, d, s);
}
Simplified s_dst_over(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_dst_over(P, P);
  if (d == X) return nop();
  if (d == A) return s_src_over(d, s);
  return mk(kCompDstOver, d, s);
}
Simplified s_src_in(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return s_src_in(P, P);
  if (d == P && s == F) return s_src_in(P, P);
  if (d == X) return s_src_copy(d, s);
  if (d == A && s == Z) return s_clear(d, s);
  if (d == A && s == X) return nop();
  if (d == A && s == F) return nop();
  return mk(kCompSrcIn, d, s);
}
Simplified s_dst_in(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return s_src_copy(d, s);
  if (d == P && s == X) return nop();
  if (d == P && s == F) return nop();
  if (d == X && s == P) return s_dst_in(P, P);
  if (d == X && s == Z) return s_dst_in(P, F);
  if (d == X && s == X) return nop();
  if (d == X && s == F) return nop();
  if (d == A) return s_src_in(d, s);
  return mk(kCompDstIn, d, s);
}
Simplified s_plus(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_plus(P, P);
  if (d == X && s == P) return s_plus(P, P);
  if (d == X && s == Z) return nop();
  if (d == X && s == X) return s_plus(P, P);
  if (d == X && s == F) return s_plus(P, P);
  if (d == A && s == Z) return nop();
  if (d == A && s == X) return mk(kCompPlus, d, P, kSolidOpaqueWhite);
  if (d == A && s == F) return mk(kCompPlus, d, P, kSolidOpaqueWhite);
  return mk(kCompPlus, d, s);
}
Simplified s_modulate(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return mk(kCompSrcCopy, P, P, kSolidTransparent);
  if (d == P && s == F) return s_modulate(P, P);
  if (d == X && s == Z) return mk(kCompSrcCopy, P, P, kSolidOpaqueBlack);
  if (d == X && s == F) return s_modulate(X, P);
  if (d == X && s == X) return s_modulate(X, P);
  if (d == A || s == A) return s_dst_in(d, s);
  return mk(kCompModulate, d, s);
}
Simplified s_multiply(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_multiply(P, X);
  if (d == X && s == Z) return nop();
  if (d == X && s == F) return s_modulate(X, X);
  if (d == X && s == X) return s_modulate(X, X);
  if (d == A || s == A) return s_dst_over(d, s);
  return mk(kCompMultiply, d, s);
}
Simplified s_screen(uint32_t d, uint32_t s) {
  if (d == P && s == Z) return nop();
  if (d == P && s == F) return s_screen(P, P);
  if (d == X && s == P) return s_screen(P, P);
  if (d == X && s == Z) return nop();
  if (d == X && s == F) return s_screen(P, P);
  if (d == X && s == X) return s_screen(P, X);
  if (d == A || s == A) return s_src_over(d, s);
  return mk(kCompScreen, d, s);
}
}  // namespace

Simplified simplify(uint32_t op, uint32_t d, uint32_t s) {
  switch (op) {
    case kCompSrcOver: return s_src_over(d, s);
    case kCompSrcCopy: return s_src_copy(d, s);
    case kCompPlus: return s_plus(d, s);
    case kCompMultiply: return s_multiply(d, s);
    case kCompScreen: return s_screen(d, s);
    case kCompClear: return s_clear(d, s);
    default: return mk(op, d, s);
  }
}

}  // namespace b2amd
