// Gradient lookup tables, one entry at a time.
//
// The reference fills a gradient's table span by span on the host
// (pixelops/interpolation_avx2.cpp:18-176, selected by pixelops/funcs.cpp:
// 31-35): with f = (size - 1) << 8, stop k sits at u_k = round(offset_k * f)
// and each span runs from the previous position p = u >> 8 (0 for the first)
// to the next one q, inclusive, writing the previous stop's colour at p, and
// for q - p >= 2 stepping a 9.23 fixed-point value by trunc((b - a) * 2^23 /
// (q - p)) per entry; a later span overwrites the entry it shares with the
// one before it.  Entries past the last span take the last colour and entry 0
// the first stop's colour.  Channels are the high bytes of the 16-bit stop
// colours; every entry is premultiplied with div255.
//
// Here each entry is computed on its own -- the last span that starts at or
// before it decides its value -- so a table is filled by many threads at
// once on the device, with the same integer and double arithmetic.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define B2AMD_LUT_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define B2AMD_LUT_HD inline
#endif

namespace b2amd {
namespace lut {

// A gradient stop: offset in [0, 1] and straight-alpha colour with 16 bits
// per channel (BLRgba64 layout: b in bits 0-15, g, r, a in bits 48-63).
struct Stop {
  double offset;
  uint64_t rgba64;
};

// One table to fill: `size` entries at `lut` (in 32-bit words) from
// `n_stops` stops at `stop0`, sorted by offset; `wide` tables hold 64-bit
// entries (two words each, low word first).
struct Job {
  uint32_t lut;
  uint32_t size;
  uint32_t stop0;
  uint32_t n_stops;
  uint32_t wide;
};

B2AMD_LUT_HD uint32_t div255(uint32_t x) { return ((x + 128u) * 257u) >> 16; }

// support/math_p.h:337-343, 369-370: rint within the int32 range, then
// halves rounded up.
B2AMD_LUT_HD int32_t round_to_int(double x) {
  int32_t y = (x >= -2147483648.0 && x < 2147483648.0) ? (int32_t)__builtin_rint(x) : INT32_MIN;
  return y + ((double)y - x == -0.5);
}

// The high bytes of a 16-bit colour as a premultiplied 0xAARRGGBB pixel.
B2AMD_LUT_HD uint32_t premultiplied(uint64_t c) {
  const uint32_t b = (uint32_t)(c >> 8) & 0xFFu, g = (uint32_t)(c >> 24) & 0xFFu;
  const uint32_t r = (uint32_t)(c >> 40) & 0xFFu, a = (uint32_t)(c >> 56) & 0xFFu;
  return (div255(255u * a) << 24) | (div255(r * a) << 16) | (div255(g * a) << 8) | div255(b * a);
}

// Position (entry index) of stop k.
B2AMD_LUT_HD uint32_t stop_pos(const Stop* st, uint32_t k, double f_width) {
  return (uint32_t)round_to_int(st[k].offset * f_width) >> 8;
}

// Entry `e` (0 <= e < size) of the table of `n` >= 1 stops.
B2AMD_LUT_HD uint32_t entry(const Stop* st, uint32_t n, uint32_t size, uint32_t e) {
  if (e == 0) return premultiplied(st[0].rgba64);
  const double f_width = (double)((int32_t)(size - 1u) << 8);
  // spans t = 0 .. m - 1 end at stop s0 + t; span t starts where span t - 1
  // ended (span 0 at entry 0) and interpolates from the stop before
  const uint32_t s0 = (st[0].offset == 0.0 && n > 1) ? 1u : 0u;
  const uint32_t m = n - s0;
  auto start = [&](uint32_t t) -> uint32_t { return t == 0 ? 0u : stop_pos(st, s0 + t - 1u, f_width); };
  const uint32_t last_end = stop_pos(st, n - 1u, f_width);
  if (e > last_end) return premultiplied(st[n - 1u].rgba64);
  // the last span starting at or before e (starts never decrease)
  uint32_t lo = 0, hi = m - 1u;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (start(mid) <= e) lo = mid;
    else hi = mid - 1u;
  }
  const uint32_t t = lo;
  const uint32_t p = start(t), q = stop_pos(st, s0 + t, f_width);
  const uint64_t a = st[t == 0 ? 0u : s0 + t - 1u].rgba64;
  const uint64_t b = st[s0 + t].rgba64;
  const uint32_t i = q - p;
  if (i <= 1u) return premultiplied(e == p ? a : b);
  const double scale = (double)(1 << 23) / (double)(int32_t)i;
  const uint32_t j = e - p;
  uint64_t out = 0;
  for (int k = 0; k < 4; k++) {
    const int32_t ca = (int32_t)((a >> (16 * k + 8)) & 0xFFu);
    const int32_t cb = (int32_t)((b >> (16 * k + 8)) & 0xFFu);
    const int32_t dx = (int32_t)((double)(cb - ca) * scale);
    const uint32_t cx = (uint32_t)(ca << 23) + (1u << 22);
    uint32_t v = (cx + (uint32_t)dx * j) >> 23;
    v = v > 255u ? 255u : v;
    out |= (uint64_t)(v << 8) << (16 * k);
  }
  return premultiplied(out);
}

// ---- 64-bit tables (dithered gradients) ----------------------------------
//
// The reference fills a dithered gradient's table with the generic
// interpolate_prgb64 (pixelops/interpolation.cpp:161-275, the only one
// pixelops/funcs.cpp:23 installs): span t (t = 0 .. n - 1) runs from entry
// start(t) (0 for the first, else the end of span t - 1) to its end
// pos(t) = round(offset_t * f) >> 8, f = (size - 1) << 8, inclusive.  It
// writes the previous stop's colour (stop 0's for span 0; the span's own for
// an empty span) at start(t) and, when the two colours differ, steps each
// 16-bit channel in 16.15 fixed point by (c1 - c0) / (pos - start)
// (truncated) from c0 + 1/2; later spans overwrite the entry they share.
// Entries past the last span take the last colour, entry 0 the first stop's.
// Interpolated entries of two opaque colours keep their 16-bit channels;
// others are premultiplied channel by channel with udiv65535, as every
// stop colour is (cvt_prgb64_8888_from_argb64_8888, pixelops/scalar_p.h:
// 179-222).  Layout: a in bits 48-63, r, g, b below (BLRgba64).

B2AMD_LUT_HD uint32_t div65535(uint32_t x) { return ((x + 0x8000u) + ((x + 0x8000u) >> 16)) >> 16; }

B2AMD_LUT_HD uint64_t premultiplied64(uint64_t c) {
  const uint32_t a = (uint32_t)(c >> 48);
  const uint64_t r = div65535((uint32_t)((c >> 32) & 0xFFFFu) * a);
  const uint64_t g = div65535((uint32_t)((c >> 16) & 0xFFFFu) * a);
  const uint64_t b = div65535((uint32_t)(c & 0xFFFFu) * a);
  return ((uint64_t)a << 48) | (r << 32) | (g << 16) | b;
}

// Entry `e` (0 <= e < size) of the 64-bit table of `n` >= 1 stops.
B2AMD_LUT_HD uint64_t entry64(const Stop* st, uint32_t n, uint32_t size, uint32_t e) {
  if (e == 0 || n == 1u) return premultiplied64(st[0].rgba64);
  const double f_width = (double)((int32_t)(size - 1u) << 8);
  auto pos = [&](uint32_t k) -> uint32_t { return (uint32_t)round_to_int(st[k].offset * f_width) >> 8; };
  auto start = [&](uint32_t t) -> uint32_t { return t == 0 ? 0u : pos(t - 1u); };
  if (e > pos(n - 1u)) return premultiplied64(st[n - 1u].rgba64);
  // the last span starting at or before e (starts never decrease)
  uint32_t lo = 0, hi = n - 1u;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (start(mid) <= e) lo = mid;
    else hi = mid - 1u;
  }
  const uint32_t p = start(lo), q = pos(lo);
  const uint64_t c1 = st[lo].rgba64;
  const uint64_t c0 = (q == p) ? c1 : st[lo == 0 ? 0u : lo - 1u].rgba64;
  const uint32_t j = e - p;
  if (c0 == c1 || j == 0) return premultiplied64(c0);
  const int32_t i = (int32_t)(q - p);
  const uint32_t kMask = 0xFFFFu << 15;
  auto step = [&](uint32_t a0, uint32_t a1) -> uint32_t {
    const uint32_t inc = (uint32_t)((int32_t)(a1 - a0) / i);
    return a0 + (1u << 14) + inc * j;
  };
  const uint32_t rp = step((uint32_t)(c0 >> 17) & kMask, (uint32_t)(c1 >> 17) & kMask);
  const uint32_t gp = step((uint32_t)(c0 >> 1) & kMask, (uint32_t)(c1 >> 1) & kMask);
  const uint32_t bp = step((uint32_t)(c0 << 15) & kMask, (uint32_t)(c1 << 15) & kMask);
  if (((c0 & c1) >> 48) == 0xFFFFu)
    return ((uint64_t)(rp & kMask) << 17) | ((uint64_t)(gp & kMask) << 1) | ((uint64_t)(bp & kMask) >> 15) |
           0xFFFF000000000000ull;
  const uint32_t ap = step((uint32_t)(c0 >> 33) & kMask, (uint32_t)(c1 >> 33) & kMask);
  const uint32_t ca = ap >> 15;
  const uint64_t cr = div65535((rp >> 15) * ca), cg = div65535((gp >> 15) * ca), cb = div65535((bp >> 15) * ca);
  return ((uint64_t)ca << 48) | (cr << 32) | (cg << 16) | cb;
}

}  // namespace lut
}  // namespace b2amd
