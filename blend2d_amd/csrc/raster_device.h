// Coverage cells of one edge segment on one pixel row, in closed form.
//
// The reference rasterizes an edge by walking it scanline by scanline with
// two incremental divisions (raster/analyticrasterizer_p.h:289-1165): one
// steps x down the row boundaries, the other steps y across the column
// boundaries, each rounding its quotient to the nearest value.  A tile of
// the render kernel needs one row of one edge at a time, so this file
// evaluates those two sequences directly at any index instead of stepping
// to it:
//
//   x on the row boundary y = 256 k     X(k) = x0 +- floor(((256 k - y0) dx + ceil(dy / 2)) / dy)
//   y on the column boundary x = 256 m  Y(m) = y0 + floor((|256 m - x0| dy + ceil(dx / 2)) / dx)
//
// (both are the fixed points of the reference's lift / remainder / error
// accumulators, whose error term starts at floor(d / 2) - 1; the start and
// end rows use the edge's own end points).  Within a row the edge enters at
// X(r) and leaves at X(r + 1); every cell between gets the cover of the y
// interval between the column crossings around it and the area of its x
// extent.  How the crossing y is taken relative to the row differs with the
// slope, as in the reference: an edge at least as tall as wide (at most one
// crossing per row) keeps it modulo 256, a flatter one as is.
//
// Only prefix sums of a row's cells are ever used (the coverage of pixel x
// is the sum of cells 0..x), so a tile keeps the cells inside it and folds
// every cell left of it into one backdrop value, its cover << 9; the area a
// cell left of the tile carries into the first cell of the tile cancels out
// of every prefix sum inside the tile and is dropped on both sides.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define B2AMD_HD __host__ __device__ __forceinline__
#else
#define B2AMD_HD inline
#endif

namespace b2amd {
namespace ras {

// floor(n / d) for d > 0.  Below 2^53 the quotient comes from the correctly
// rounded double division: with Q = n / d < 2^53 / d the rounding error is
// at most 2^-53 * Q < 1 / d, while a non-integral Q lies at least 1 / d away
// from every integer, so truncating the rounded quotient gives floor(Q).
// The 24.8 coordinates the rasterizer sees keep every dividend far below
// 2^53; the integer division stays as the exact path for anything larger.
B2AMD_HD uint64_t udiv_floor(uint64_t n, uint64_t d) {
  if (n < (1ull << 53)) return (uint64_t)__builtin_trunc((double)n / (double)d);
  return n / d;
}

// Coordinates beyond this magnitude never come out of the edge builder for a
// canvas (it clips to the canvas in 24.8); an edge carrying one is ignored.
constexpr int32_t kCoordLimit = 1 << 30;

struct Edge {
  int32_t x0, y0, x1, y1;  // 24.8, y0 < y1
  int32_t dx, dy;          // |x1 - x0|, y1 - y0
  int32_t dir;             // +1: x grows downwards (or constant), -1: x shrinks
  int32_t row0, row1;      // first and last pixel row it touches
  int32_t top0;            // start y within row0 (0..255)
  int32_t bot1;            // end y within row1 (1..256)
  bool one_column;         // vertical, or within one cell: one cell column, constant area factor
  bool steep;              // dy >= dx: at most one column crossing per row
};

// Returns false for edges that add nothing (horizontal) or that no clipped
// geometry can produce.
B2AMD_HD bool edge_init(Edge& e, int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
  if (y0 >= y1) return false;
  if (x0 < -kCoordLimit || x0 > kCoordLimit || x1 < -kCoordLimit || x1 > kCoordLimit ||
      y0 < -kCoordLimit || y1 > kCoordLimit)
    return false;
  e.x0 = x0; e.y0 = y0; e.x1 = x1; e.y1 = y1;
  e.dir = x1 < x0 ? -1 : 1;
  e.dx = x1 < x0 ? x0 - x1 : x1 - x0;
  e.dy = y1 - y0;
  e.row0 = y0 >> 8;
  e.row1 = (y1 - 1) >> 8;
  e.top0 = y0 & 255;
  e.bot1 = ((y1 - 1) & 255) + 1;
  e.one_column = (x0 >> 8) == (x1 >> 8) && (e.row0 == e.row1 || e.dx == 0);
  e.steep = e.dy >= e.dx;
  return true;
}

// x where the edge crosses the row boundary y = 256 k, row0 < k <= row1.
B2AMD_HD int32_t x_at_row_boundary(const Edge& e, int32_t k) {
  const uint64_t t = (uint64_t)(int64_t)(256 * (int64_t)k - e.y0);
  const uint64_t q = udiv_floor(t * (uint64_t)e.dx + (uint64_t)(e.dy - (e.dy >> 1)), (uint64_t)e.dy);
  return e.x0 + e.dir * (int32_t)q;
}

// Crossings of consecutive column boundaries in the direction of travel:
// the boundary at distance `dist` from x0 is crossed at y0 + q, and each
// next boundary is 256 further.
struct ColumnWalk {
  uint64_t q, rem;          // quotient and remainder of dist * dy + ceil(dx / 2) by dx
  uint64_t lift, lift_rem;  // the same of 256 * dy
  uint64_t dx;
  B2AMD_HD void start(const Edge& e, uint64_t dist) {
    dx = (uint64_t)e.dx;
    const uint64_t num = dist * (uint64_t)e.dy + (dx - (dx >> 1));
    q = udiv_floor(num, dx);
    rem = num - q * dx;
    const uint64_t step = 256u * (uint64_t)e.dy;
    lift = udiv_floor(step, dx);
    lift_rem = step - lift * dx;
  }
  B2AMD_HD void next() {
    q += lift;
    rem += lift_rem;
    if (rem >= dx) { q++; rem -= dx; }
  }
};

// Row `r` (row0 <= r <= row1) of edge `e`, with cover sign `sgn` (+1 or -1),
// against a tile whose pixel columns are [x_left, x_left + width): calls
// sink(i, v) to add v to tile cell i (0 <= i < width) and returns the
// backdrop, the cover << 9 of the row's cells left of the tile.
template<typename Sink>
B2AMD_HD uint32_t emit_row(const Edge& e, int32_t r, int32_t x_left, int32_t width, int32_t sgn, Sink&& sink) {
  const int32_t ytop = r == e.row0 ? e.top0 : 0;
  const int32_t ybot = r == e.row1 ? e.bot1 : 256;
  const int32_t x_end = x_left + width;   // first column right of the tile

  // One cell: cover over the row, area factor fsum.
  auto single = [&](int32_t c, int32_t fsum) -> uint32_t {
    const int32_t cover = sgn * (ybot - ytop);
    if (c < x_left) return (uint32_t)cover << 9;
    if (c < x_end) {
      const int32_t area = cover * fsum;
      sink(c - x_left, ((uint32_t)cover << 9) - (uint32_t)area);
      if (c + 1 < x_end) sink(c + 1 - x_left, (uint32_t)area);
    }
    return 0u;
  };

  if (e.one_column) return single(e.x0 >> 8, (e.x0 & 255) + (e.x1 & 255));

  const int32_t xe = r == e.row0 ? e.x0 : x_at_row_boundary(e, r);
  const int32_t xx = r == e.row1 ? e.x1 : x_at_row_boundary(e, r + 1);
  const int32_t run = e.dir > 0 ? xx - xe : xe - xx;   // >= 0

  // Entry cell and entry offset: moving right a point on a boundary starts
  // the cell right of it, moving left the cell left of it (offset 256).
  int32_t ce, fe;
  if (e.dir > 0) { ce = xe >> 8; fe = xe & 255; }
  else { ce = (xe - 1) >> 8; fe = ((xe - 1) & 255) + 1; }
  if (e.dir > 0 ? fe + run <= 256 : fe - run >= 0) return single(ce, e.dir > 0 ? 2 * fe + run : 2 * fe - run);

  // Several cells, ce to cx in the direction of travel.
  const int32_t cx = e.dir > 0 ? (xx - 1) >> 8 : xx >> 8;
  const int32_t fx = e.dir > 0 ? ((xx - 1) & 255) + 1 : xx & 255;
  const int32_t lo = e.dir > 0 ? ce : cx, hi = e.dir > 0 ? cx : ce;
  if (hi < x_left) return (uint32_t)(sgn * (ybot - ytop)) << 9;
  if (lo >= x_end) return 0u;

  // In-tile cells in the direction of travel: first k0, last k1.
  const int32_t k0 = e.dir > 0 ? (ce > x_left ? ce : x_left) : (ce < x_end - 1 ? ce : x_end - 1);
  const int32_t k1 = e.dir > 0 ? (cx < x_end - 1 ? cx : x_end - 1) : (cx > x_left ? cx : x_left);
  // Crossing y of the boundary a cell is entered through, row-relative.
  const int32_t row_y = 256 * r;
  auto rel = [&](uint64_t q) -> int32_t {
    const int32_t y = e.y0 + (int32_t)q;
    return e.steep ? (y & 255) : y - row_y;
  };
  ColumnWalk walk;
  int32_t y_in;
  if (k0 == ce) {
    y_in = ytop;
    // the next boundary: right of ce / left of ce
    const int32_t m = e.dir > 0 ? ce + 1 : ce;
    walk.start(e, (uint64_t)(int64_t)(e.dir > 0 ? 256 * (int64_t)m - e.x0 : e.x0 - 256 * (int64_t)m));
  }
  else {
    // entered through the tile edge: the left edge x_left moving right, the
    // right edge x_end moving left
    const int32_t m = e.dir > 0 ? k0 : k0 + 1;
    walk.start(e, (uint64_t)(int64_t)(e.dir > 0 ? 256 * (int64_t)m - e.x0 : e.x0 - 256 * (int64_t)m));
    y_in = rel(walk.q);
    walk.next();
  }
  // Backdrop: cells left of the tile telescope to the y interval they span.
  uint32_t backdrop = 0;
  if (lo < x_left) {
    // moving right they are the first cells (ytop .. y_in of the first tile
    // cell), moving left the last ones (the crossing at x_left .. ybot)
    if (e.dir > 0) backdrop = (uint32_t)(sgn * (y_in - ytop)) << 9;
  }
  int32_t carry = 0;   // moving right: area the previous tile cell passes into this one
  for (int32_t k = k0;; k += e.dir) {
    const bool last = k == cx;
    const int32_t y_out = last ? ybot : rel(walk.q);
    const int32_t f_in = k == ce ? fe : (e.dir > 0 ? 0 : 256);
    const int32_t f_out = last ? fx : (e.dir > 0 ? 256 : 0);
    const int32_t cover = sgn * (y_out - y_in);
    const int32_t area = cover * (f_in + f_out);
    const uint32_t own = ((uint32_t)cover << 9) - (uint32_t)area;
    if (e.dir > 0) {
      sink(k - x_left, own + (uint32_t)carry);
      carry = area;
    }
    else {
      // moving left, the cell right of this one was visited just before
      sink(k - x_left, own);
      if (k + 1 < x_end) sink(k + 1 - x_left, (uint32_t)area);
    }
    if (k == k1) {
      if (e.dir > 0 && k + 1 < x_end) sink(k + 1 - x_left, (uint32_t)carry);
      // moving left past the tile: the rest of the row lies left of it
      if (e.dir < 0 && !last) backdrop = (uint32_t)(sgn * (ybot - y_out)) << 9;
      break;
    }
    y_in = y_out;
    walk.next();
  }
  return backdrop;
}

}  // namespace ras
}  // namespace b2amd
