// Checks the closed-form row emitter of the render kernel
// (blend2d_amd/csrc/raster_device.h, compiled here for the host) against the
// oracle's restatement of the reference rasterizer (oracle/orc_raster.c,
// itself pinned by the reference-rendered golden canvases): for random and
// boundary-hugging segments, every row and tiles at many offsets, the
// coverage prefix sum at every pixel of the tile (backdrop + tile cells)
// must equal the sum of the oracle's cells up to that pixel.
//
//   raster_rows_check <segments> <seed>     exit status 0 when all agree
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../blend2d_amd/csrc/raster_device.h"

extern "C" int orc_raster_segment(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign, uint32_t* cells,
                                  size_t stride, uint32_t rows);

namespace {

constexpr int kW = 300;       // canvas width in pixels
constexpr int kH = 40;        // canvas height in pixels
constexpr int kStride = kW + 8;

struct Gen {
  std::mt19937_64 rng;
  explicit Gen(uint64_t seed) : rng(seed) {}
  int uni(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); }
  // a coordinate, often on or next to a pixel boundary
  int coord(int max_fixed) {
    switch (uni(0, 5)) {
      case 0: return uni(0, max_fixed / 256) * 256;
      case 1: return std::min(max_fixed, uni(0, max_fixed / 256) * 256 + uni(-1, 1) + 1);
      default: return uni(0, max_fixed);
    }
  }
};

}  // namespace

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 20000;
  const uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  Gen g(seed);
  std::vector<uint32_t> cells((size_t)kStride * kH);
  long bad = 0, rows_checked = 0;
  for (long i = 0; i < n; i++) {
    int x0, y0, x1, y1;
    const int kind = g.uni(0, 7);
    y0 = g.coord(kH * 256);
    y1 = g.coord(kH * 256);
    x0 = g.coord(kW * 256);
    if (kind == 0) x1 = std::min(kW * 256, std::max(0, x0 + g.uni(-300, 300)));          // near vertical
    else if (kind == 1) { x1 = g.coord(kW * 256); y1 = std::min(kH * 256, y0 + g.uni(1, 300)); }  // flat
    else if (kind == 2) x1 = x0;                                                           // vertical
    else if (kind == 7) {                                                                  // 45 degrees (dx = dy or dy - 1)
      const int d = g.uni(2, 256 * 12);
      y1 = std::min(kH * 256, y0 + d);
      const int dxx = (y1 - y0) - g.uni(0, 1);
      x1 = g.uni(0, 1) ? x0 + dxx : x0 - dxx;
      if (x1 < 0 || x1 > kW * 256) x1 = x0;
    }
    else if (kind == 6) {                                                                  // very steep, tall
      y0 = g.uni(0, 256 * 4);
      y1 = std::min(kH * 256, y0 + g.uni(256 * 8, 256 * 30));
      x1 = std::min(kW * 256, std::max(0, x0 + g.uni(-12, 12)));
    }
    else x1 = g.coord(kW * 256);
    // some end points exactly on cell boundaries
    if (g.uni(0, 3) == 0) x0 &= ~255;
    if (g.uni(0, 3) == 0) x1 &= ~255;
    if (y0 == y1) continue;
    if (y0 > y1) { std::swap(y0, y1); std::swap(x0, x1); }
    const uint32_t sign = (uint32_t)g.uni(0, 1);
    std::fill(cells.begin(), cells.end(), 0u);
    if (!orc_raster_segment(x0, y0, x1, y1, sign, cells.data(), kStride, kH)) continue;
    b2amd::ras::Edge e;
    if (!b2amd::ras::edge_init(e, x0, y0, x1, y1)) { bad++; continue; }
    const int32_t sgn = sign ? -1 : 1;
    int lefts[8] = {0, 64, 128, 192, 256, g.uni(0, kW - 1), g.uni(0, kW - 1), std::max(0, (x0 >> 8) - g.uni(0, 64))};
    for (int li = 0; li < 8; li++) {
      const int xl = lefts[li];
      const int width = li < 6 ? 64 : g.uni(1, 64);
      // row by row, in chunks of 1..4 rows
      std::vector<uint32_t> tile((size_t)(e.row1 - e.row0 + 1) * 64, 0u), bdr(e.row1 - e.row0 + 1, 0u);
      for (int r = e.row0; r <= e.row1;) {
        const int last = std::min(e.row1, r + g.uni(0, 3));
        for (int row = r; row <= last; row++) {
          bdr[row - e.row0] += b2amd::ras::emit_row(e, row, xl, width, sgn, [&](int32_t k, uint32_t v) {
            if (k < 0 || k >= width) { bad++; return; }
            tile[(size_t)(row - e.row0) * 64 + k] += v;
          });
        }
        r = last + 1;
      }
      for (int r = e.row0; r <= e.row1; r++) {
        const uint32_t* row = &cells[(size_t)r * kStride];
        uint32_t want = 0;
        for (int x = 0; x < xl && x < kStride; x++) want += row[x];
        uint32_t got = bdr[r - e.row0];
        for (int k = 0; k < width && xl + k < kStride; k++) {
          want += row[xl + k];
          got += tile[(size_t)(r - e.row0) * 64 + k];
          if (want != got) {
            if (bad < 10)
              fprintf(stderr, "mismatch: seg (%d,%d)-(%d,%d) sign %u row %d tile %d+%d pixel %d: want %d got %d\n",
                      x0, y0, x1, y1, sign, r, xl, width, xl + k, (int)want, (int)got);
            bad++;
            break;
          }
        }
        rows_checked++;
      }
    }
  }
  printf("segments %ld rows*tiles %ld mismatches %ld\n", n, rows_checked, bad);
  return bad ? 1 : 0;
}
