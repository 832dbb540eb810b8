/* CPU oracle -- fetchers and composition operators.
 *
 * TEST INFRASTRUCTURE (see orc_core.h).  Restates
 *   pipeline/reference/fetchgeneric_p.h  (solid, linear/radial/conic gradient,
 *                                         affine NN/bilinear pattern)
 *   pipeline/reference/compopgeneric_p.h (SRC_COPY, SRC_OVER, PLUS)
 *   pipeline/reference/pixelgeneric_p.h  (div255, addus8, PixelIO)
 *   pipeline/jit/compoppart.cpp:4150-4176, 4291-4460, 4592-4628
 *                                        (PLUS, MULTIPLY, SCREEN; the non-JIT
 *                                          reference has no MULTIPLY/SCREEN)
 * The reference fetchers step their state per pixel and per row; every one of
 * them is a function of the pixel position except the conic gradient, whose
 * per-row origin accumulates in double precision from the row the fill call
 * started at (fetchgeneric_p.h:1172-1173, 1227-1230).  That accumulation is
 * kept here.
 */
#include <math.h>
#include <string.h>

#include "orc_core.h"

/* pixelgeneric_p.h:386-389 (x <= 65025 + 128) */
uint32_t orc_div255(uint32_t x) {
  uint32_t u = x + 0x80u;
  return ((u + ((u >> 8) & 0xFFu)) >> 8) & 0xFFu;
}

static inline uint32_t ch(uint32_t p, int i) { return (p >> (i * 8)) & 0xFFu; }

/* x86 cvttss2si / cvtss2si behaviour: out-of-range and NaN give INT32_MIN. */
static inline int32_t trunc_f32_to_i32(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
  return (int32_t)v;
}
static inline int32_t rint_f32_to_i32(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
  return (int32_t)rintf(v);
}

/* ---- source texel fetch (PixelIO, pixelgeneric_p.h:591-676) ---- */
static inline uint32_t tex_fetch(const OrcFetch* f, uint32_t fmt, uint32_t x, uint32_t y) {
  const uint8_t* row = f->img + (intptr_t)y * f->img_stride;
  if (fmt == ORC_FMT_A8) return (uint32_t)row[x] * 0x01010101u;
  uint32_t v;
  memcpy(&v, row + (size_t)x * 4u, 4);
  if (fmt == ORC_FMT_XRGB32) v |= 0xFF000000u;
  return v;
}

/* Texel column of pattern position u: pad clamps; a repeat-or-reflect period
 * rx wraps (fetchgeneric_p.h:264-352: positions past the width index from the
 * far end, x ^ (x >> 31)). */
static uint32_t sp_index_x(const OrcFetch* f, int64_t u) {
  int64_t w = (int64_t)f->img_w;
  if (f->sp_rx == 0) return (uint32_t)(u < 0 ? 0 : u > w - 1 ? w - 1 : u);
  int64_t v = u % f->sp_rx;
  if (v < 0) v += f->sp_rx;
  if (v >= w) v -= f->sp_rx;
  return (uint32_t)(v ^ (v >> 63));
}

/* Texel row (fetchgeneric_p.h:80-165): pad clamps; repeat wraps by h;
 * reflect wraps by 2h and mirrors the second half. */
static uint32_t sp_index_y(const OrcFetch* f, int64_t v) {
  int64_t h = (int64_t)f->img_h;
  if (f->sp_ry == 0) return (uint32_t)(v < 0 ? 0 : v > h - 1 ? h - 1 : v);
  int64_t r = v % f->sp_ry;
  if (r < 0) r += f->sp_ry;
  if (r >= h) r = f->sp_ry - 1 - r;
  return (uint32_t)r;
}

/* FetchPatternAffineCtx::normalize_px_py (fetchgeneric_p.h:373-384) */
static inline uint64_t pa_norm(int32_t tw, int32_t rx, int32_t ox, uint64_t v) {
  uint32_t x = (uint32_t)((int32_t)(v >> 32) % tw);
  if ((int32_t)x < 0) x += (uint32_t)rx;
  if ((int32_t)x > ox) x -= (uint32_t)rx;
  return ((uint64_t)x << 32) | (v & 0xFFFFFFFFu);
}

static inline int32_t pa_index(int32_t v, int32_t mn, int32_t mx, int32_t cor) {
  if (v < mn) v = mn;
  if (v > mx) v = cor;
  return v ^ (v >> 31);
}

/* bayer_matrix_16x16 (tables/tables_p.h:448-469): the dither thresholds of
 * pixel (x & 15, y & 15). */
static const uint8_t kBayer16[16][16] = {
  {  0, 191,  48, 239,  12, 203,  60, 251,   3, 194,  51, 242,  15, 206,  63, 254},
  {127,  64, 175, 112, 139,  76, 187, 124, 130,  67, 178, 115, 142,  79, 190, 127},
  { 32, 223,  16, 207,  44, 235,  28, 219,  35, 226,  19, 210,  47, 238,  31, 222},
  {159,  96, 143,  80, 171, 108, 155,  92, 162,  99, 146,  83, 174, 111, 158,  95},
  {  8, 199,  56, 247,   4, 195,  52, 243,  11, 202,  59, 250,   7, 198,  55, 246},
  {135,  72, 183, 120, 131,  68, 179, 116, 138,  75, 186, 123, 134,  71, 182, 119},
  { 40, 231,  24, 215,  36, 227,  20, 211,  43, 234,  27, 218,  39, 230,  23, 214},
  {167, 104, 151,  88, 163, 100, 147,  84, 170, 107, 154,  91, 166, 103, 150,  87},
  {  2, 193,  50, 241,  14, 205,  62, 253,   1, 192,  49, 240,  13, 204,  61, 252},
  {129,  66, 177, 114, 141,  78, 189, 126, 128,  65, 176, 113, 140,  77, 188, 125},
  { 34, 225,  18, 209,  46, 237,  30, 221,  33, 224,  17, 208,  45, 236,  29, 220},
  {161,  98, 145,  82, 173, 110, 157,  94, 160,  97, 144,  81, 172, 109, 156,  93},
  { 10, 201,  58, 249,   6, 197,  54, 245,   9, 200,  57, 248,   5, 196,  53, 244},
  {137,  74, 185, 122, 133,  70, 181, 118, 136,  73, 184, 121, 132,  69, 180, 117},
  { 42, 233,  26, 217,  38, 229,  22, 213,  41, 232,  25, 216,  37, 228,  21, 212},
  {169, 106, 153,  90, 165, 102, 149,  86, 168, 105, 152,  89, 164, 101, 148,  85}};

/* Entry `idx` of a gradient table at pixel (x, y): the 32-bit entry, or for a
 * dithered gradient the 64-bit one with the pixel's Bayer threshold added to
 * each colour channel, cut to 8 bits and limited by alpha
 * (FetchGradientBase<DITHER>::fetch_pixel, fetchgeneric_p.h:899-934; pixel
 * origin 0, 0). */
static uint32_t lut_px(const OrcFetch* f, uint32_t idx, uint32_t x, uint32_t y) {
  if (!f->lut64) return f->lut[idx];
  uint64_t v = f->lut64[idx];
  uint32_t dd = kBayer16[y & 15u][x & 15u];
  uint32_t a = (uint32_t)(v >> 56);
  uint32_t r = (uint32_t)(((v >> 32) & 0xFFFFu) + dd) >> 8;
  uint32_t g = (uint32_t)(((v >> 16) & 0xFFFFu) + dd) >> 8;
  uint32_t b = (uint32_t)((v & 0xFFFFu) + dd) >> 8;
  r = r < a ? r : a; g = g < a ? g : a; b = b < a ? b : a;
  return (a << 24) | (r << 16) | (g << 8) | b;
}

static uint32_t fetch_px(const OrcPipe* p, uint32_t x, uint32_t y) {
  const OrcFetch* f = p->fetch;
  switch (f->type) {
    case ORC_FETCH_SOLID:
      return f->solid_prgb32;

    case ORC_FETCH_LINEAR: {
      uint64_t pt = f->lin_pt0 + (uint64_t)y * f->lin_dy + (uint64_t)x * f->lin_dt;
      uint32_t idx = (uint32_t)(pt >> 32);
      if (f->extend == 0) {
        int32_t i = (int32_t)idx;
        if (i < 0) i = 0;
        if (i > (int32_t)f->maxi) i = (int32_t)f->maxi;
        idx = (uint32_t)i;
      }
      else {
        uint32_t a = idx & f->maxi, b = a ^ f->rori;
        idx = b < a ? b : a;
      }
      return lut_px(f, idx, x, y);
    }

    case ORC_FETCH_RADIAL: {
      double yd = (double)(int32_t)y;
      double ptx = f->rad_yx * yd + f->rad_tx;
      double pty = f->rad_yy * yd + f->rad_ty;
      double b = yd * f->rad_by + f->rad_b0;
      double sq_dist = ptx * ptx + pty * pty;
      float fb = (float)(b * f->rad_inv2a);
      float fd = (float)((f->rad_amul4 * (sq_dist - f->rad_sq_fr) + b * b) * f->rad_sq_inv2a);
      float fdd = (float)((yd * f->rad_ddy + f->rad_dd0) * f->rad_sq_inv2a);
      float fx = (float)(int32_t)x;
      float sq_x = fx * fx;
      float a = sqrtf(fabsf(sq_x * f->rad_ddd + (fx * fdd + fd)));
      float v = (fx * f->rad_bd + fb) + a;
      uint32_t idx = (uint32_t)trunc_f32_to_i32(v);
      if (f->extend == 0) {
        int32_t i = (int32_t)idx;
        if (i < 0) i = 0;
        if (i > (int32_t)f->maxi) i = (int32_t)f->maxi;
        idx = (uint32_t)i;
      }
      else {
        uint32_t aa = idx & f->maxi, bb = aa ^ f->rori;
        idx = bb < aa ? bb : aa;
      }
      return lut_px(f, idx, x, y);
    }

    case ORC_FETCH_CONIC: {
      float tx = (float)p->con_tp_x;
      float ay = (float)p->con_tp_y;
      float by = signbit(ay) ? f->con_n1 : 0.0f;
      ay = fabsf(ay);
      float fx = (float)(int32_t)x;
      float xv = fx * f->con_xx + tx;
      float ax = fabsf(xv);
      float xy_min = ay < ax ? ay : ax;
      float xy_max = ax < ay ? ay : ax;
      float s = (ax == xy_min) ? f->con_n4 : 0.0f;
      float pp = xy_min / xy_max;
      float p_sq = pp * pp;
      float v = p_sq * f->con_q[3] + f->con_q[2];
      v = v * p_sq + f->con_q[1];
      v = v * p_sq + f->con_q[0];
      v = fabsf(v * pp - s);
      v = fabsf(v - (signbit(xv) ? f->con_n2 : 0.0f));
      v = fabsf(v - by) + f->con_offset;
      int32_t i = rint_f32_to_i32(v);
      if (i > f->con_maxi) i = f->con_maxi;
      return lut_px(f, (uint32_t)i & (uint32_t)f->con_rori, x, y);
    }

    case ORC_FETCH_PATTERN_AFFINE_NN:
    case ORC_FETCH_PATTERN_AFFINE_BI: {
      uint64_t tx = f->pa_tx + f->pa_yx * (uint64_t)y;
      uint64_t ty = f->pa_ty + f->pa_yy * (uint64_t)y;
      uint64_t px = pa_norm(f->pa_tw, f->pa_rx, f->pa_ox, tx + f->pa_xx * (uint64_t)x);
      uint64_t py = pa_norm(f->pa_th, f->pa_ry, f->pa_oy, ty + f->pa_xy * (uint64_t)x);
      int32_t ix = (int32_t)(px >> 32), iy = (int32_t)(py >> 32);
      uint32_t fmt = p->src_fmt;
      if (f->type == ORC_FETCH_PATTERN_AFFINE_NN) {
        int32_t xi = pa_index(ix, f->pa_minx, f->pa_maxx, f->pa_corx);
        int32_t yi = pa_index(iy, f->pa_miny, f->pa_maxy, f->pa_cory);
        return tex_fetch(f, fmt, (uint32_t)xi, (uint32_t)yi);
      }
      int32_t x0 = pa_index(ix, f->pa_minx, f->pa_maxx, f->pa_corx);
      int32_t y0 = pa_index(iy, f->pa_miny, f->pa_maxy, f->pa_cory);
      int32_t x1 = pa_index(ix + 1, f->pa_minx, f->pa_maxx, f->pa_corx);
      int32_t y1 = pa_index(iy + 1, f->pa_miny, f->pa_maxy, f->pa_cory);
      uint32_t wx = (uint32_t)(px & 0xFFFFFFFFu) >> 24;
      uint32_t wy = (uint32_t)(py & 0xFFFFFFFFu) >> 24;
      uint32_t ixw = 256 - wx, iyw = 256 - wy;
      uint32_t t00 = tex_fetch(f, fmt, (uint32_t)x0, (uint32_t)y0);
      uint32_t t01 = tex_fetch(f, fmt, (uint32_t)x0, (uint32_t)y1);
      uint32_t t10 = tex_fetch(f, fmt, (uint32_t)x1, (uint32_t)y0);
      uint32_t t11 = tex_fetch(f, fmt, (uint32_t)x1, (uint32_t)y1);
      uint32_t out = 0;
      for (int c = 0; c < 4; c++) {
        uint32_t p0 = ((ch(t00, c) * iyw + ch(t01, c) * wy) >> 8) & 0xFFu;
        uint32_t p1 = ((ch(t10, c) * iyw + ch(t11, c) * wy) >> 8) & 0xFFu;
        p0 *= ixw;
        p1 *= wx;
        out |= (((p0 + p1) >> 8) & 0xFFu) << (c * 8);
      }
      return out;
    }

    /* fetchgeneric_p.h:80-184 (vertical extend), :186-352 (horizontal
     * extend: pad / repeat / repeat-or-reflect), :470-540 (aligned),
     * :543-667 (FxFy); each context is stepped per pixel and row there, here
     * its position is computed directly from the absolute x and y. */
    case ORC_FETCH_PATTERN_ALIGNED:
      return tex_fetch(f, p->src_fmt, sp_index_x(f, (int64_t)(int32_t)x + f->sp_tx),
                       sp_index_y(f, (int64_t)(int32_t)y + f->sp_ty));

    case ORC_FETCH_PATTERN_FXFY: {
      uint32_t fmt = p->src_fmt;
      uint32_t x0 = sp_index_x(f, (int64_t)(int32_t)x + f->sp_tx), x1 = sp_index_x(f, (int64_t)(int32_t)x + 1 + f->sp_tx);
      uint32_t y0 = sp_index_y(f, (int64_t)(int32_t)y + f->sp_ty), y1 = sp_index_y(f, (int64_t)(int32_t)y + 1 + f->sp_ty);
      uint32_t p00 = tex_fetch(f, fmt, x0, y0), p01 = tex_fetch(f, fmt, x0, y1);
      uint32_t p10 = tex_fetch(f, fmt, x1, y0), p11 = tex_fetch(f, fmt, x1, y1);
      uint32_t out = 0;
      for (int c = 0; c < 4; c++) {
        /* 16-bit lanes, as the unpacked pixel arithmetic */
        uint32_t acc = ((ch(p00, c) * f->sp_wa) & 0xFFFFu) + ((ch(p01, c) * f->sp_wc) & 0xFFFFu);
        acc = (((ch(p10, c) * f->sp_wb) & 0xFFFFu) + ((ch(p11, c) * f->sp_wd) & 0xFFFFu) + (acc & 0xFFFFu)) & 0xFFFFu;
        out |= ((acc >> 8) & 0xFFu) << (c * 8);
      }
      return out;
    }

    default:
      return 0;
  }
}

void orc_pipe_begin(OrcPipe* p, uint32_t y0) {
  const OrcFetch* f = p->fetch;
  p->y_start = y0;
  p->con_row = (int32_t)y0;
  if (f->type == ORC_FETCH_CONIC) {
    p->con_tp_x = f->con_tx + f->con_yx * (double)(int32_t)y0;
    p->con_tp_y = f->con_ty + f->con_yy * (double)(int32_t)y0;
  }
}

void orc_pipe_row(OrcPipe* p, uint32_t y) {
  const OrcFetch* f = p->fetch;
  if (f->type == ORC_FETCH_CONIC) {
    while (p->con_row < (int32_t)y) {
      p->con_tp_x += f->con_yx;
      p->con_tp_y += f->con_yy;
      p->con_row++;
    }
  }
}

/* ---- composition (per channel on 8-bit values) ---- */

static inline uint32_t op_channel(uint32_t op, uint32_t d, uint32_t s, uint32_t da, uint32_t sa, uint32_t m) {
  (void)da; (void)sa;
  switch (op) {
    case ORC_OP_SRC_COPY:
      return orc_div255(d * (m ^ 0xFFu) + s * m);
    case ORC_OP_PLUS: {
      uint32_t v = d + orc_div255(s * m);
      return v > 255u ? 255u : v;
    }
    default:
      return d;
  }
}

void orc_pipe_composite(OrcPipe* p, OrcCanvas* c, uint32_t x, uint32_t y, uint32_t m) {
  uint8_t* row = c->pixels + (intptr_t)y * c->stride;
  uint32_t s = fetch_px(p, x, y);
  uint32_t op = p->op;

  if (c->format == ORC_FMT_A8) {
    uint32_t d = row[x];
    uint32_t sa = s >> 24;
    uint32_t r;
    switch (op) {
      case ORC_OP_SRC_COPY: r = orc_div255(d * (m ^ 0xFFu) + sa * m); break;
      case ORC_OP_SRC_OVER: {
        uint32_t s1 = orc_div255(sa * m);
        r = (s1 + orc_div255(d * (s1 ^ 0xFFu))) & 0xFFu;
        break;
      }
      case ORC_OP_PLUS: { uint32_t v = d + orc_div255(sa * m); r = v > 255u ? 255u : v; break; }
      default: r = d; break;
    }
    row[x] = (uint8_t)r;
    return;
  }

  uint32_t d;
  memcpy(&d, row + (size_t)x * 4u, 4);
  uint32_t r = 0;
  switch (op) {
    case ORC_OP_SRC_COPY:
    case ORC_OP_PLUS:
      for (int i = 0; i < 4; i++) r |= op_channel(op, ch(d, i), ch(s, i), 0, 0, m) << (i * 8);
      break;
    case ORC_OP_SRC_OVER: {
      uint32_t s1 = 0;
      for (int i = 0; i < 4; i++) s1 |= orc_div255(ch(s, i) * m) << (i * 8);
      uint32_t ia = (s1 >> 24) ^ 0xFFu, t = 0;
      for (int i = 0; i < 4; i++) t |= orc_div255(ch(d, i) * ia) << (i * 8);
      r = s1 + t;   /* packed 32-bit add, as P32_8888::operator+ */
      break;
    }
    case ORC_OP_MULTIPLY: {
      /* Dca' = Dca.(Sca.m + 1 - Sa.m) + Sca.m.(1 - Da), one div255 per channel */
      uint32_t s1[4];
      for (int i = 0; i < 4; i++) s1[i] = orc_div255(ch(s, i) * m);
      uint32_t da = ch(d, 3);
      for (int i = 0; i < 4; i++) {
        uint32_t v = ch(d, i) * ((s1[i] + (s1[3] ^ 0xFFu)) & 0xFFFFu) + s1[i] * (da ^ 0xFFu);
        r |= orc_div255(v & 0xFFFFu) << (i * 8);
      }
      break;
    }
    case ORC_OP_SCREEN: {
      /* Dca' = Sca.m + Dca.(1 - Sca.m) */
      for (int i = 0; i < 4; i++) {
        uint32_t s1 = orc_div255(ch(s, i) * m);
        r |= ((orc_div255(ch(d, i) * (s1 ^ 0xFFu)) + s1) & 0xFFu) << (i * 8);
      }
      break;
    }
    default:
      r = d;
      break;
  }
  memcpy(row + (size_t)x * 4u, &r, 4);
}
