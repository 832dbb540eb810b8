// Device fetchers and composition operators.
//
// Fetch: solid, linear/radial/conic gradient LUT lookup, affine pattern with
// nearest or bilinear filtering, translated pattern at an integer or
// fractional offset (pipeline/reference/fetchgeneric_p.h).  The
// reference steps each fetcher per pixel and per row; all of them are pure
// functions of the pixel position except the conic gradient, whose per-row
// origin accumulates in double precision from the first row of the fill call
// (fetchgeneric_p.h:1172-1173, 1227-1230).  That accumulation is reproduced.
//
// Composite: SRC_COPY / SRC_OVER / PLUS (pipeline/reference/compopgeneric_p.h)
// and MULTIPLY / SCREEN (pipeline/jit/compoppart.cpp:4291-4460, 4592-4628)
// with Blend2D's 8-bit div255 rounding (pixelgeneric_p.h:386-389).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"

namespace b2amd {
namespace pipe {

__device__ __forceinline__ uint32_t div255(uint32_t x) {
  uint32_t u = x + 0x80u;
  return ((u + ((u >> 8) & 0xFFu)) >> 8) & 0xFFu;
}

__device__ __forceinline__ uint32_t ch(uint32_t p, int i) { return (p >> (i * 8)) & 0xFFu; }

// div255 of two 16-bit lanes at once (bits 0-15 and 16-31, each <= 255 * 255):
// every intermediate of a lane stays below 65536, so no carry crosses lanes.
__device__ __forceinline__ uint32_t div255x2(uint32_t x) {
  const uint32_t u = x + 0x00800080u;
  return ((u + ((u >> 8) & 0x00FF00FFu)) >> 8) & 0x00FF00FFu;
}

// Per-lane min(v, 255) of two lanes holding values <= 510.
__device__ __forceinline__ uint32_t sat255x2(uint32_t v) {
  const uint32_t o = v & 0x01000100u;
  return (v | (o - (o >> 8))) & 0x00FF00FFu;
}

// x86 cvttss2si / cvtss2si: out-of-range and NaN produce INT32_MIN.
__device__ __forceinline__ int32_t trunc_f32_to_i32(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
  return (int32_t)v;
}
__device__ __forceinline__ int32_t rint_f32_to_i32(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
  return (int32_t)__builtin_rintf(v);
}

// Translated pattern, texel row of pattern row v (fetchgeneric_p.h:80-165):
// pad clamps to [0, h-1]; repeat wraps by h; reflect wraps by 2h and mirrors
// the second half.
__device__ __forceinline__ uint32_t sp_row(const DevStyle& s, int64_t v) {
  const int64_t h = (int64_t)s.sp_h;
  if (s.sp_ry == 0) return (uint32_t)(v < 0 ? 0 : v > h - 1 ? h - 1 : v);
  int64_t r = v % s.sp_ry;
  if (r < 0) r += s.sp_ry;
  if (r >= h) r = s.sp_ry - 1 - r;
  return (uint32_t)r;
}

// Translated pattern, texel column of pattern column u
// (fetchgeneric_p.h:187-352): pad clamps; a period rx (repeat: w, reflect:
// 2w) wraps, and a wrapped position v >= w maps to rx - 1 - v.  For repeat
// and reflect the offset is kept in [0, rx), so u >= 0 there.
__device__ __forceinline__ uint32_t sp_col(const DevStyle& s, int64_t u) {
  const int64_t w = (int64_t)s.sp_w;
  if (s.sp_rx == 0) return (uint32_t)(u < 0 ? 0 : u > w - 1 ? w - 1 : u);
  uint32_t v = (uint32_t)u % (uint32_t)s.sp_rx;
  if (v >= (uint32_t)w) v = (uint32_t)s.sp_rx - 1u - v;
  return v;
}

// Per (command, row) fetch state.
struct Row {
  uint64_t a, b;       // linear: pt at x = 0; pattern: tx, ty of the row;
                       // gradients: the row in b (dither thresholds)
  float f0, f1, f2;    // radial: b, d, dd;  conic: tx, ay, by
};

__device__ __forceinline__ Row row_setup(const DevStyle& s, int y, int fill_row0, uint32_t ref_band_h) {
  Row r;
  r.a = 0; r.b = (uint32_t)y; r.f0 = 0.f; r.f1 = 0.f; r.f2 = 0.f;
  switch (s.kind) {
    case kFetchLinear:
      r.a = s.lin_pt0 + (uint64_t)(uint32_t)y * s.lin_dy;
      break;
    case kFetchRadial: {
      double yd = (double)y;
      double ptx = s.rad_yx * yd + s.rad_tx;
      double pty = s.rad_yy * yd + s.rad_ty;
      double b = yd * s.rad_by + s.rad_b0;
      double sq_dist = ptx * ptx + pty * pty;
      r.f0 = (float)(b * s.rad_inv2a);
      r.f1 = (float)((s.rad_amul4 * (sq_dist - s.rad_sq_fr) + b * b) * s.rad_sq_inv2a);
      r.f2 = (float)((yd * s.rad_ddy + s.rad_dd0) * s.rad_sq_inv2a);
      break;
    }
    case kFetchConic: {
      // First row of the reference fill call that covers `y`: analytic fills
      // run per band (first band starts at the path's first row), box fills
      // start at the box's first row.
      int ys = fill_row0;
      if (ref_band_h) {
        int band_start = (y / (int)ref_band_h) * (int)ref_band_h;
        if (band_start > ys) ys = band_start;
      }
      double tpx = s.con_tx + s.con_yx * (double)ys;
      double tpy = s.con_ty + s.con_yy * (double)ys;
      for (int k = ys; k < y; k++) { tpx += s.con_yx; tpy += s.con_yy; }
      float ay = (float)tpy;
      r.f0 = (float)tpx;
      r.f2 = __builtin_signbit(ay) ? s.con_n1 : 0.0f;
      r.f1 = __builtin_fabsf(ay);
      break;
    }
    case kFetchPatternNN:
    case kFetchPatternBI:
      r.a = s.pa_tx + s.pa_yx * (uint64_t)(uint32_t)y;
      r.b = s.pa_ty + s.pa_yy * (uint64_t)(uint32_t)y;
      break;
    case kFetchPatternAligned:
    case kFetchPatternFxFy:
      // texel rows of this row and of the next one (FxFy)
      r.a = sp_row(s, (int64_t)y + s.sp_ty);
      r.b = sp_row(s, (int64_t)y + 1 + s.sp_ty);
      break;
    default:
      break;
  }
  return r;
}

// bayer_matrix_16x16 (tables/tables_p.h:448-469): dither thresholds of
// pixel (x & 15, y & 15), row by row.
__constant__ uint8_t kBayer16[256] = {
    0, 191,  48, 239,  12, 203,  60, 251,   3, 194,  51, 242,  15, 206,  63, 254,
  127,  64, 175, 112, 139,  76, 187, 124, 130,  67, 178, 115, 142,  79, 190, 127,
   32, 223,  16, 207,  44, 235,  28, 219,  35, 226,  19, 210,  47, 238,  31, 222,
  159,  96, 143,  80, 171, 108, 155,  92, 162,  99, 146,  83, 174, 111, 158,  95,
    8, 199,  56, 247,   4, 195,  52, 243,  11, 202,  59, 250,   7, 198,  55, 246,
  135,  72, 183, 120, 131,  68, 179, 116, 138,  75, 186, 123, 134,  71, 182, 119,
   40, 231,  24, 215,  36, 227,  20, 211,  43, 234,  27, 218,  39, 230,  23, 214,
  167, 104, 151,  88, 163, 100, 147,  84, 170, 107, 154,  91, 166, 103, 150,  87,
    2, 193,  50, 241,  14, 205,  62, 253,   1, 192,  49, 240,  13, 204,  61, 252,
  129,  66, 177, 114, 141,  78, 189, 126, 128,  65, 176, 113, 140,  77, 188, 125,
   34, 225,  18, 209,  46, 237,  30, 221,  33, 224,  17, 208,  45, 236,  29, 220,
  161,  98, 145,  82, 173, 110, 157,  94, 160,  97, 144,  81, 172, 109, 156,  93,
   10, 201,  58, 249,   6, 197,  54, 245,   9, 200,  57, 248,   5, 196,  53, 244,
  137,  74, 185, 122, 133,  70, 181, 118, 136,  73, 184, 121, 132,  69, 180, 117,
   42, 233,  26, 217,  38, 229,  22, 213,  41, 232,  25, 216,  37, 228,  21, 212,
  169, 106, 153,  90, 165, 102, 149,  86, 168, 105, 152,  89, 164, 101, 148,  85};

// Gradient table entry `i` at pixel (x, row y): the 32-bit entry, or for a
// dithered gradient the 64-bit one with the pixel's Bayer threshold added to
// each colour channel before it is cut to 8 bits, limited by alpha
// (FetchGradientBase<DITHER>::fetch_pixel, fetchgeneric_p.h:899-934).
// kDither = false (frames without dithered gradients) leaves the dither path
// out of the code, and its registers out of the render kernels.
template<bool kDither>
__device__ __forceinline__ uint32_t lut_px(const DevStyle& s, const uint32_t* __restrict__ luts, uint32_t i, int x,
                                           uint32_t y) {
  if (!kDither || !s.dither) return luts[s.lut + i];
  const uint32_t lo = luts[s.lut + 2u * i], hi = luts[s.lut + 2u * i + 1u];
  const uint32_t dd = kBayer16[((y & 15u) << 4) | ((uint32_t)x & 15u)];
  const uint32_t a = hi >> 24;
  uint32_t r = ((hi & 0xFFFFu) + dd) >> 8;
  uint32_t g = ((lo >> 16) + dd) >> 8;
  uint32_t b = ((lo & 0xFFFFu) + dd) >> 8;
  r = r < a ? r : a;
  g = g < a ? g : a;
  b = b < a ? b : a;
  return (a << 24) | (r << 16) | (g << 8) | b;
}

__device__ __forceinline__ uint32_t grad_index(uint32_t idx, const DevStyle& s) {
  if (s.extend == 0) {
    int32_t i = (int32_t)idx;
    if (i < 0) i = 0;
    if (i > (int32_t)s.maxi) i = (int32_t)s.maxi;
    return (uint32_t)i;
  }
  uint32_t a = idx & s.maxi, b = a ^ s.rori;
  return b < a ? b : a;
}

__device__ __forceinline__ uint64_t pa_norm(int32_t tw, int32_t rx, int32_t ox, uint64_t v) {
  uint32_t x = (uint32_t)((int32_t)(v >> 32) % tw);
  if ((int32_t)x < 0) x += (uint32_t)rx;
  if ((int32_t)x > ox) x -= (uint32_t)rx;
  return ((uint64_t)x << 32) | (v & 0xFFFFFFFFull);
}

__device__ __forceinline__ int32_t pa_index(int32_t v, int32_t mn, int32_t mx, int32_t cor) {
  if (v < mn) v = mn;
  if (v > mx) v = cor;
  return v ^ (v >> 31);
}

__device__ __forceinline__ uint32_t texel(const DevImage& im, uint32_t fmt, int32_t x, int32_t y) {
  const uint8_t* row = (const uint8_t*)im.ptr + (int64_t)y * im.stride;
  if (fmt == kFmtA8) return (uint32_t)row[x] * 0x01010101u;
  uint32_t v = ((const uint32_t*)row)[x];
  if (fmt == kFmtXRGB32) v |= 0xFF000000u;
  return v;
}

template<bool kDither = false>
__device__ __forceinline__ uint32_t fetch(const DevStyle& s, const Row& r, int x, const uint32_t* __restrict__ luts,
                                          const DevImage* __restrict__ images, uint32_t src_fmt) {
  switch (s.kind) {
    case kFetchSolid:
      return s.solid;
    case kFetchLinear: {
      uint64_t pt = r.a + (uint64_t)(uint32_t)x * s.lin_dt;
      return lut_px<kDither>(s, luts, grad_index((uint32_t)(pt >> 32), s), x, (uint32_t)r.b);
    }
    case kFetchRadial: {
      float fx = (float)x;
      float sq_x = fx * fx;
      float a = __builtin_sqrtf(__builtin_fabsf(sq_x * s.rad_ddd + (fx * r.f2 + r.f1)));
      float v = (fx * s.rad_bd + r.f0) + a;
      return lut_px<kDither>(s, luts, grad_index((uint32_t)trunc_f32_to_i32(v), s), x, (uint32_t)r.b);
    }
    case kFetchConic: {
      float fx = (float)x;
      float xv = fx * s.con_xx + r.f0;
      float ax = __builtin_fabsf(xv);
      float ay = r.f1;
      float xy_min = ay < ax ? ay : ax;
      float xy_max = ax < ay ? ay : ax;
      float sv = (ax == xy_min) ? s.con_n4 : 0.0f;
      float p = xy_min / xy_max;
      float p_sq = p * p;
      float v = p_sq * s.con_q[3] + s.con_q[2];
      v = v * p_sq + s.con_q[1];
      v = v * p_sq + s.con_q[0];
      v = __builtin_fabsf(v * p - sv);
      v = __builtin_fabsf(v - (__builtin_signbit(xv) ? s.con_n2 : 0.0f));
      v = __builtin_fabsf(v - r.f2) + s.con_offset;
      int32_t i = rint_f32_to_i32(v);
      if (i > s.con_maxi) i = s.con_maxi;
      return lut_px<kDither>(s, luts, (uint32_t)i & (uint32_t)s.con_rori, x, (uint32_t)r.b);
    }
    case kFetchPatternNN:
    case kFetchPatternBI: {
      const DevImage& im = images[s.img];
      uint64_t px = pa_norm(s.pa_tw, s.pa_rx, s.pa_ox, r.a + s.pa_xx * (uint64_t)(uint32_t)x);
      uint64_t py = pa_norm(s.pa_th, s.pa_ry, s.pa_oy, r.b + s.pa_xy * (uint64_t)(uint32_t)x);
      int32_t ix = (int32_t)(px >> 32), iy = (int32_t)(py >> 32);
      if (s.kind == kFetchPatternNN)
        return texel(im, src_fmt, pa_index(ix, s.pa_minx, s.pa_maxx, s.pa_corx), pa_index(iy, s.pa_miny, s.pa_maxy, s.pa_cory));
      int32_t x0 = pa_index(ix, s.pa_minx, s.pa_maxx, s.pa_corx);
      int32_t y0 = pa_index(iy, s.pa_miny, s.pa_maxy, s.pa_cory);
      int32_t x1 = pa_index(ix + 1, s.pa_minx, s.pa_maxx, s.pa_corx);
      int32_t y1 = pa_index(iy + 1, s.pa_miny, s.pa_maxy, s.pa_cory);
      uint32_t wx = (uint32_t)(px & 0xFFFFFFFFull) >> 24;
      uint32_t wy = (uint32_t)(py & 0xFFFFFFFFull) >> 24;
      uint32_t iw = 256 - wx, ih = 256 - wy;
      uint32_t t00 = texel(im, src_fmt, x0, y0), t01 = texel(im, src_fmt, x0, y1);
      uint32_t t10 = texel(im, src_fmt, x1, y0), t11 = texel(im, src_fmt, x1, y1);
      uint32_t out = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        uint32_t p0 = ((ch(t00, c) * ih + ch(t01, c) * wy) >> 8) & 0xFFu;
        uint32_t p1 = ((ch(t10, c) * ih + ch(t11, c) * wy) >> 8) & 0xFFu;
        out |= (((p0 * iw + p1 * wx) >> 8) & 0xFFu) << (c * 8);
      }
      return out;
    }
    case kFetchPatternAligned:
      return texel(images[s.img], src_fmt, (int32_t)sp_col(s, (int64_t)x + s.sp_tx), (int32_t)r.a);
    case kFetchPatternFxFy: {
      // fetchgeneric_p.h:601-687: columns x0 = x + tx and x1 = x0 + 1 of rows
      // y0, y1, weighted wa (x0 y0), wb (x1 y0), wc (x0 y1), wd (x1 y1).  The
      // weights sum to at most 256, so each 16-bit lane of the R/B and A/G
      // pairs stays below 65536.
      const DevImage& im = images[s.img];
      const int32_t x0 = (int32_t)sp_col(s, (int64_t)x + s.sp_tx), x1 = (int32_t)sp_col(s, (int64_t)x + 1 + s.sp_tx);
      const uint32_t t00 = texel(im, src_fmt, x0, (int32_t)r.a), t01 = texel(im, src_fmt, x0, (int32_t)r.b);
      const uint32_t t10 = texel(im, src_fmt, x1, (int32_t)r.a), t11 = texel(im, src_fmt, x1, (int32_t)r.b);
      const uint32_t m = 0x00FF00FFu;
      const uint32_t rb = (t00 & m) * s.sp_wa + (t01 & m) * s.sp_wc + (t10 & m) * s.sp_wb + (t11 & m) * s.sp_wd;
      const uint32_t ag = ((t00 >> 8) & m) * s.sp_wa + ((t01 >> 8) & m) * s.sp_wc +
                          ((t10 >> 8) & m) * s.sp_wb + ((t11 >> 8) & m) * s.sp_wd;
      return ((rb >> 8) & m) | (ag & ~m);
    }
    default:
      return 0;
  }
}

// One PRGB32 destination pixel.
__device__ __forceinline__ uint32_t composite32(uint32_t op, uint32_t d, uint32_t s, uint32_t m) {
  uint32_t r = 0;
  switch (op) {
    // SRC_COPY, SRC_OVER and PLUS work on the R/B and A/G channel pairs
    // (two 16-bit lanes per register); per channel the arithmetic is the
    // scalar div255 form.
    case kOpSrcCopy: {
      const uint32_t im = m ^ 0xFFu;
      const uint32_t rb = div255x2((d & 0x00FF00FFu) * im + (s & 0x00FF00FFu) * m);
      const uint32_t ag = div255x2(((d >> 8) & 0x00FF00FFu) * im + ((s >> 8) & 0x00FF00FFu) * m);
      return rb | (ag << 8);
    }
    case kOpSrcOver: {
      const uint32_t s1 = div255x2((s & 0x00FF00FFu) * m) | (div255x2(((s >> 8) & 0x00FF00FFu) * m) << 8);
      const uint32_t ia = (s1 >> 24) ^ 0xFFu;
      const uint32_t t = div255x2((d & 0x00FF00FFu) * ia) | (div255x2(((d >> 8) & 0x00FF00FFu) * ia) << 8);
      return s1 + t;
    }
    case kOpPlus: {
      const uint32_t rb = sat255x2((d & 0x00FF00FFu) + div255x2((s & 0x00FF00FFu) * m));
      const uint32_t ag = sat255x2(((d >> 8) & 0x00FF00FFu) + div255x2(((s >> 8) & 0x00FF00FFu) * m));
      return rb | (ag << 8);
    }
    case kOpMultiply: {
      uint32_t s1[4];
#pragma unroll
      for (int i = 0; i < 4; i++) s1[i] = div255(ch(s, i) * m);
      uint32_t da = ch(d, 3);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint32_t v = ch(d, i) * ((s1[i] + (s1[3] ^ 0xFFu)) & 0xFFFFu) + s1[i] * (da ^ 0xFFu);
        r |= div255(v & 0xFFFFu) << (i * 8);
      }
      return r;
    }
    case kOpScreen: {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint32_t s1 = div255(ch(s, i) * m);
        r |= ((div255(ch(d, i) * (s1 ^ 0xFFu)) + s1) & 0xFFu) << (i * 8);
      }
      return r;
    }
    default:
      return d;
  }
}

// One A8 destination pixel; `sa` is the source alpha.
__device__ __forceinline__ uint32_t composite8(uint32_t op, uint32_t d, uint32_t sa, uint32_t m) {
  switch (op) {
    case kOpSrcCopy: return div255(d * (m ^ 0xFFu) + sa * m);
    case kOpSrcOver: {
      uint32_t s1 = div255(sa * m);
      return (s1 + div255(d * (s1 ^ 0xFFu))) & 0xFFu;
    }
    case kOpPlus: {
      uint32_t v = d + div255(sa * m);
      return v > 255u ? 255u : v;
    }
    default:
      return d;
  }
}

}  // namespace pipe
}  // namespace b2amd
