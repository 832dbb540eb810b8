// Device analytic rasterizer: the reference's cover/area cell rasterizer
// (raster/analyticrasterizer_p.h:289-1165) with banding, where the band is the
// 16-row tile the calling workgroup owns.  A segment that starts above the
// tile is first advanced with advance_to_y() (analyticrasterizer_p.h:371-459);
// the reference's own test (raster/analyticrasterizer_test.cpp:34-158) pins
// that advancing equals iterated banded rasterization, which is what lets
// every tile start a segment at its own first row.  Cell writes go to a Sink
// (tile-local LDS cells plus a per-row backdrop for cells left of the tile).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b2amd {
namespace ras {

enum { A8_SHIFT = 8, A8_SCALE = 256, A8_MASK = 255 };
enum { F_INITIAL = 1, F_VERT_OR_SINGLE = 2, F_RTL = 4 };

struct RasState {
  int ex0, ey0, ex1, ey1;
  int fx0, fy0, fx1, fy1;
  int xErr, yErr, xDlt, yDlt, xRem, yRem, xLift, yLift;
  int dx, dy, savedFy1;
  uint32_t flags;
};

__device__ __forceinline__ uint32_t apply_sign(uint32_t x, uint32_t mask) { return (x ^ mask) - mask; }

// floor(n / d) for d > 0.  Below 2^53 the quotient comes from the correctly
// rounded double division: with Q = n / d < 2^53 / d the rounding error is
// at most 2^-53 * Q < 1 / d, while a non-integral Q lies at least 1 / d away
// from every integer, so truncating the rounded quotient gives floor(Q).
// The 24.8 coordinates the rasterizer sees keep every dividend far below
// 2^53; the integer division stays as the exact path for anything larger.
__device__ __forceinline__ uint64_t udiv_floor(uint64_t n, uint32_t d) {
  if (n < (1ull << 53)) return (uint64_t)__builtin_trunc((double)n / (double)d);
  return n / (uint64_t)d;
}

__device__ __forceinline__ void acc_err_step(int* acc, int* iter, int step, int correction) {
  *iter -= step;
  int mask = *iter >> 31;
  *acc -= mask;
  *iter += mask & correction;
}

__device__ __forceinline__ void acc_err_multi_step(int* acc, int* iter, int step, int correction, int count) {
  int64_t i = (int64_t)(uint32_t)(*iter);
  i -= (int64_t)((uint64_t)(uint32_t)step * (uint32_t)count);
  if (i < 0) {
    int n = (int)udiv_floor((uint64_t)(-i) + (uint32_t)correction - 1u, (uint32_t)correction);
    *acc += n;
    i += (int64_t)correction * n;
  }
  *iter = (int)i;
}

template<typename Sink>
__device__ __forceinline__ void merge(Sink& sink, int row, int x, uint32_t cover, uint32_t area) {
  sink.add(row, x, (cover << 9) - area);
  sink.add(row, x + 1, area);
}

struct Rasterizer {
  RasState st;
  uint32_t sign_mask;
  int band_end;   // last row this call may write

  /* analyticrasterizer_p.h:289-360 */
  __device__ __forceinline__ bool prepare(int p0x, int p0y, int p1x, int p1y) {
    RasState* s = &st;
    if (p0y == p1y) return 0;

    s->dx = p1x - p0x;
    s->dy = p1y - p0y;
    s->flags = F_INITIAL;
    if (s->dx < 0) { s->flags |= F_RTL; s->dx = -s->dx; }

    s->ex0 = p0x >> A8_SHIFT;
    s->ey0 = p0y >> A8_SHIFT;
    s->ex1 = p1x >> A8_SHIFT;
    s->ey1 = (p1y - 1) >> A8_SHIFT;

    s->fx0 = p0x & A8_MASK;
    s->fy0 = p0y & A8_MASK;
    s->fx1 = p1x & A8_MASK;
    s->fy1 = ((p1y - 1) & A8_MASK) + 1;

    s->savedFy1 = s->fy1;
    if (s->ey0 != s->ey1) s->fy1 = A8_SCALE;

    if (s->ex0 == s->ex1 && (s->ey0 == s->ey1 || s->dx == 0)) {
      s->flags |= F_VERT_OR_SINGLE;
      return 1;
    }

    uint64_t x_base = (uint64_t)(uint32_t)s->dx * A8_SCALE;
    uint64_t y_base = (uint64_t)(uint32_t)s->dy * A8_SCALE;

    const uint64_t xq = udiv_floor(x_base, (uint32_t)s->dy);
    const uint64_t yq = udiv_floor(y_base, (uint32_t)s->dx);
    s->xLift = (int)xq;
    s->xRem = (int)(x_base - xq * (uint32_t)s->dy);
    s->yLift = (int)yq;
    s->yRem = (int)(y_base - yq * (uint32_t)s->dx);

    s->xDlt = s->dx;
    s->yDlt = s->dy;
    s->xErr = (s->dy >> 1) - 1;
    s->yErr = (s->dx >> 1) - 1;

    if (s->ey0 != s->ey1) {
      uint64_t p = (uint64_t)(A8_SCALE - (uint32_t)s->fy0) * (uint32_t)s->dx;
      const uint64_t q = udiv_floor(p, (uint32_t)s->dy);
      s->xDlt = (int)q;
      s->xErr -= (int)(p - q * (uint32_t)s->dy);
      acc_err_step(&s->xDlt, &s->xErr, 0, s->dy);
    }

    if (s->ex0 != s->ex1) {
      uint64_t p = (uint64_t)((s->flags & F_RTL) ? (uint32_t)s->fx0 : A8_SCALE - (uint32_t)s->fx0) * (uint32_t)s->dy;
      const uint64_t q = udiv_floor(p, (uint32_t)s->dx);
      s->yDlt = (int)q;
      s->yErr -= (int)(p - q * (uint32_t)s->dx);
      acc_err_step(&s->yDlt, &s->yErr, 0, s->dx);
    }

    s->yDlt += s->fy0;
    return 1;
  }

  /* analyticrasterizer_p.h:371-459 */
  __device__ __forceinline__ void advance_to_y(int y_target) {
    RasState* s = &st;
    if (y_target <= s->ey0) return;

    if (!(s->flags & F_VERT_OR_SINGLE)) {
      int ny = y_target - s->ey0;
      s->xDlt += s->xLift * (ny - 1);
      acc_err_multi_step(&s->xDlt, &s->xErr, s->xRem, s->dy, ny - 1);

      if (s->flags & F_RTL) {
        s->fx0 -= s->xDlt;
        if (s->fx0 < 0) {
          int nx = -(s->fx0 >> A8_SHIFT);
          s->ex0 -= nx;
          s->fx0 &= A8_MASK;
          acc_err_multi_step(&s->yDlt, &s->yErr, s->yRem, s->dx, nx);
          s->yDlt += s->yLift * nx;
        }
        if (!(s->dy >= s->dx)) {
          if (!s->fx0) {
            s->fx0 = A8_SCALE;
            s->ex0--;
            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            s->yDlt += s->yLift;
          }
        }
        if (y_target == s->ey1 && s->dy >= s->dx) {
          s->fy1 = s->savedFy1;
          s->xDlt = ((s->ex0 - s->ex1) << A8_SHIFT) + s->fx0 - s->fx1;
        }
        else {
          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
        }
      }
      else {
        s->fx0 += s->xDlt;
        if (s->fx0 >= A8_SCALE) {
          int nx = s->fx0 >> A8_SHIFT;
          s->ex0 += nx;
          s->fx0 &= A8_MASK;
          acc_err_multi_step(&s->yDlt, &s->yErr, s->yRem, s->dx, nx);
          s->yDlt += s->yLift * nx;
        }
        if (y_target == s->ey1 && s->dy >= s->dx) {
          s->fy1 = s->savedFy1;
          s->xDlt = ((s->ex1 - s->ex0) << A8_SHIFT) + s->fx1 - s->fx0;
        }
        else {
          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
        }
      }

      if (s->dy >= s->dx) {
        s->yDlt &= A8_MASK;
      }
      else {
        int y = ny;
        if (s->flags & F_INITIAL) y--;
        s->yDlt -= y * A8_SCALE;
      }
    }
    else {
      if (y_target == s->ey1) s->fy1 = s->savedFy1;
    }

    s->fy0 = 0;
    s->ey0 = y_target;
    s->flags &= ~(uint32_t)F_INITIAL;
  }

  /* analyticrasterizer_p.h:466-1165, with kOptionBandOffset | kOptionRecordMinXMaxX
   * and kOptionBandingMode when `banding` is set.  Returns 1 when the edge is
   * fully rasterized. */
  template<typename Sink>
  __device__ __forceinline__ bool rasterize(Sink& sink) {
    const int banding = 1;
    RasState* s = &st;
    int ey1_end = s->ey1;
    if (banding && ey1_end > band_end) ey1_end = band_end;
    if (ey1_end < s->ey0) return true;

    size_t i = (unsigned)ey1_end - (unsigned)s->ey0;
    int cell = s->ey0;

    if (banding) s->ey0 += (int)i + 1;

    const uint32_t sm = sign_mask;
    const uint32_t full_cover = apply_sign(A8_SCALE, sm);

    if (s->flags & F_VERT_OR_SINGLE) {
      uint32_t cover;
      uint32_t area = (uint32_t)s->fx0 + (uint32_t)s->fx1;
        
      cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
      merge(sink, cell, s->ex0, cover, cover * area);

      if (banding) {
        if (!i) {
          if (s->ey0 > s->ey1) return 1;
          s->fy0 = 0;
          s->fy1 = s->ey0 == s->ey1 ? s->savedFy1 : A8_SCALE;
          return 0;
        }
      }
      else {
        if (!i) return 1;
      }

      cell++;
      cover = full_cover;
      while (--i) {
        merge(sink, cell, s->ex0, cover, cover * area);
        cell++;
      }

      if (banding) {
        if (s->ey0 <= s->ey1) {
          merge(sink, cell, s->ex0, cover, cover * area);
          s->fy0 = 0;
          s->fy1 = s->ey0 == s->ey1 ? s->savedFy1 : A8_SCALE;
          return 0;
        }
      }

      cover = apply_sign((uint32_t)s->savedFy1, sm);
      merge(sink, cell, s->ex0, cover, cover * area);
      return 1;
    }
    else if (s->dy >= s->dx) {
      if (banding) i += (s->ey0 <= s->ey1);

      if (s->flags & F_RTL) {
              for (;;) {
          uint32_t area = (uint32_t)s->fx0;
          uint32_t cov0, cov1;

          s->fx0 -= s->xDlt;
          if (s->fx0 < 0) {
            s->ex0--;
            s->fx0 += A8_SCALE;
            s->yDlt &= A8_MASK;

            if (!area) {
              area = A8_SCALE;
              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
              s->yDlt += s->yLift;
              goto VertRightToLeftSingleFirstOrLast;
            }

            cov0 = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
            area = cov0 * area;
            merge(sink, cell, s->ex0 + 1, cov0, area);

            cov0 = apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
            area = cov0 * ((uint32_t)s->fx0 + A8_SCALE);
            merge(sink, cell, s->ex0, cov0, area);

            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            s->yDlt += s->yLift;
          }
          else {
  VertRightToLeftSingleFirstOrLast:
            cov0 = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
            area = cov0 * (area + (uint32_t)s->fx0);
            merge(sink, cell, s->ex0, cov0, area);
          }

          s->fy0 = 0;
          cell++;

          if (!i) {
                      if (banding) {
              if (s->ey0 > s->ey1) return 1;
              s->xDlt = s->xLift;
              acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
              return 0;
            }
            return 1;
          }

          while (--i) {
            s->xDlt = s->xLift;
            acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);

            area = (uint32_t)s->fx0;
            s->fx0 -= s->xDlt;

            if (s->fx0 < 0) {
              s->ex0--;
              s->fx0 += A8_SCALE;
              s->yDlt &= A8_MASK;

              if (!area) {
                area = A8_SCALE;
                acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
                s->yDlt += s->yLift;
                goto VertRightToLeftSingleInLoop;
              }

              cov1 = apply_sign((uint32_t)s->yDlt, sm);
              area = cov1 * area;
              sink.add(cell, s->ex0 + 2, area);

              cov0 = full_cover - cov1;
              cov1 = (cov1 << 9) - area;
              area = cov0 * ((uint32_t)s->fx0 + A8_SCALE);

              cov0 = (cov0 << 9) - area;
              cov1 = cov1 + area;

              sink.add(cell, s->ex0 + 0, cov0);
              sink.add(cell, s->ex0 + 1, cov1);
              cell++;

              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
              s->yDlt += s->yLift;
            }
            else {
  VertRightToLeftSingleInLoop:
              area = full_cover * (area + (uint32_t)s->fx0);
              merge(sink, cell, s->ex0, full_cover, area);
              cell++;
            }
          }

          if (banding) {
            if (s->ey0 >= s->ey1) {
              s->fy1 = s->savedFy1;
              s->xDlt = ((s->ex0 - s->ex1) << A8_SHIFT) + s->fx0 - s->fx1;
              if (s->ey0 == s->ey1) {
                              return 0;
              }
            }
            else {
                          s->xDlt = s->xLift;
              acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
              return 0;
            }
          }
          else {
            s->fy1 = s->savedFy1;
            s->xDlt = ((s->ex0 - s->ex1) << A8_SHIFT) + s->fx0 - s->fx1;
          }
        }
      }
      else {
              for (;;) {
          uint32_t area = (uint32_t)s->fx0;
          uint32_t cov0, cov1;

          s->fx0 += s->xDlt;

          if (s->fx0 <= A8_SCALE) {
            cov0 = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
            area = cov0 * (area + (uint32_t)s->fx0);
            merge(sink, cell, s->ex0, cov0, area);

            if (s->fx0 == A8_SCALE) {
              s->ex0++;
              s->fx0 = 0;
              s->yDlt += s->yLift;
              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            }
          }
          else {
            s->ex0++;
            s->fx0 &= A8_MASK;
            s->yDlt &= A8_MASK;

            cov0 = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
            area = cov0 * (area + A8_SCALE);
            merge(sink, cell, s->ex0 - 1, cov0, area);

            cov0 = apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
            area = cov0 * (uint32_t)s->fx0;
            merge(sink, cell, s->ex0, cov0, area);

            s->yDlt += s->yLift;
            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
          }

          s->fy0 = 0;
          cell++;

          if (!i) {
                      if (banding) {
              if (s->ey0 > s->ey1) return 1;
              s->yDlt += s->yLift;
              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
              return 0;
            }
            return 1;
          }

          while (--i) {
            s->xDlt = s->xLift;
            acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);

            area = (uint32_t)s->fx0;
            s->fx0 += s->xDlt;

            if (s->fx0 <= A8_SCALE) {
              area = full_cover * (area + (uint32_t)s->fx0);
              merge(sink, cell, s->ex0, full_cover, area);
              cell++;

              if (s->fx0 < A8_SCALE) continue;

              s->ex0++;
              s->fx0 = 0;
            }
            else {
              s->fx0 &= A8_MASK;
              s->yDlt &= A8_MASK;

              cov0 = apply_sign((uint32_t)s->yDlt, sm);
              cov1 = cov0 * (area + A8_SCALE);

              cov0 = (cov0 << 9) - cov1;
              sink.add(cell, s->ex0 + 0, cov0);
              s->ex0++;

              cov0 = apply_sign(A8_SCALE - (uint32_t)s->yDlt, sm);
              area = cov0 * (uint32_t)s->fx0;

              cov0 = (cov0 << 9) - area + cov1;
              sink.add(cell, s->ex0 + 0, cov0);
              sink.add(cell, s->ex0 + 1, area);
              cell++;
            }

            s->yDlt += s->yLift;
            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
          }

          if (banding) {
            if (s->ey0 >= s->ey1) {
              s->fy1 = s->savedFy1;
              s->xDlt = ((s->ex1 - s->ex0) << A8_SHIFT) + s->fx1 - s->fx0;
              if (s->ey0 == s->ey1) {
                              return 0;
              }
            }
            else {
                          s->xDlt = s->xLift;
              acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
              return 0;
            }
          }
          else {
            s->fy1 = s->savedFy1;
            s->xDlt = ((s->ex1 - s->ex0) << A8_SHIFT) + s->fx1 - s->fx0;
          }
        }
      }
    }
    else {
      size_t j = 1;
      int x_local = (s->ex0 << A8_SHIFT) + s->fx0;
      uint32_t cover, area;

      if (s->flags & F_RTL) {
      
        if (s->flags & F_INITIAL) {
          s->flags &= ~(uint32_t)F_INITIAL;
          j = i;
          i = 1;

          cover = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
          if (s->fx0 - s->xDlt < 0) goto HorzRightToLeftInside;

          x_local -= s->xDlt;
          cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
          area = cover * (uint32_t)(s->fx0 * 2 - s->xDlt);
          merge(sink, cell, s->ex0, cover, area);

          if ((x_local & A8_MASK) == 0) {
            s->yDlt += s->yLift;
            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
          }

          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
          cell++;
          i--;
        }

        for (;;) {
          while (i) {
            s->ex0 = (x_local - 1) >> A8_SHIFT;
            s->fx0 = ((x_local - 1) & A8_MASK) + 1;

  HorzRightToLeftSkip:
            s->yDlt -= A8_SCALE;
            cover = apply_sign((uint32_t)s->yDlt, sm);

  HorzRightToLeftInside:
            x_local -= s->xDlt;
            {
              int ex_local = x_local >> A8_SHIFT;
              int fx_local = x_local & A8_MASK;

              area = cover * (uint32_t)s->fx0;
              for (int guard = 0; s->ex0 != ex_local && guard < 65536; guard++) {
                merge(sink, cell, s->ex0, cover, area);

                cover = (uint32_t)s->yLift;
                acc_err_step((int*)&cover, &s->yErr, s->yRem, s->dx);
                s->yDlt += (int32_t)cover;

                cover = apply_sign(cover, sm);
                area = cover * A8_SCALE;
                s->ex0--;
              }

              cover += apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
              area = cover * ((uint32_t)fx_local + A8_SCALE);
              merge(sink, cell, s->ex0, cover, area);

              if (fx_local == 0) {
                s->yDlt += s->yLift;
                acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
              }
            }

            s->xDlt = s->xLift;
            acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
            cell++;
            i--;
          }

          s->fy0 = 0;
          s->fy1 = A8_SCALE;

          if (!j) {
                      if (banding) {
              s->ex0 = (x_local - 1) >> A8_SHIFT;
              s->fx0 = ((x_local - 1) & A8_MASK) + 1;
              return s->ey0 > s->ey1;
            }
            return 1;
          }

          i = j - 1;
          j = 1;

          if (!i) {
            i = 1;
            j = 0;

            int is_last = banding ? s->ey0 > s->ey1 : 1;
            if (!is_last) continue;

            s->xDlt = x_local - ((s->ex1 << A8_SHIFT) + s->fx1);
            s->fy1 = s->savedFy1;

            s->ex0 = (x_local - 1) >> A8_SHIFT;
            s->fx0 = ((x_local - 1) & A8_MASK) + 1;

            if (s->fx0 - s->xDlt >= 0) {
              cover = apply_sign((uint32_t)s->fy1, sm);
              area = cover * (uint32_t)(s->fx0 * 2 - s->xDlt);
              merge(sink, cell, s->ex0, cover, area);
                          return 1;
            }
            goto HorzRightToLeftSkip;
          }
        }
      }
      else {
      
        if (s->flags & F_INITIAL) {
          s->flags &= ~(uint32_t)F_INITIAL;
          j = i;
          i = 1;

          cover = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
          if (s->fx0 + s->xDlt > A8_SCALE) goto HorzLeftToRightInside;

          x_local += s->xDlt;
          cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
          area = cover * ((uint32_t)s->fx0 * 2 + (uint32_t)s->xDlt);
          merge(sink, cell, s->ex0, cover, area);

          if (s->fx0 + s->xDlt == A8_SCALE) {
            s->yDlt += s->yLift;
            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
          }

          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
          cell++;
          i--;
        }

        for (;;) {
          while (i) {
            s->ex0 = x_local >> A8_SHIFT;
            s->fx0 = x_local & A8_MASK;

  HorzLeftToRightSkip:
            s->yDlt -= A8_SCALE;
            cover = apply_sign((uint32_t)s->yDlt, sm);

  HorzLeftToRightInside:
            x_local += s->xDlt;
            {
              int ex_local = (x_local - 1) >> A8_SHIFT;
              int fx_local = ((x_local - 1) & A8_MASK) + 1;

              area = cover * ((uint32_t)s->fx0 + A8_SCALE);
              for (int guard = 0; s->ex0 != ex_local && guard < 65536; guard++) {
                merge(sink, cell, s->ex0, cover, area);

                cover = (uint32_t)s->yLift;
                acc_err_step((int*)&cover, &s->yErr, s->yRem, s->dx);
                s->yDlt += (int32_t)cover;

                cover = apply_sign(cover, sm);
                area = cover * A8_SCALE;
                s->ex0++;
              }

              cover += apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
              area = cover * (uint32_t)fx_local;
              merge(sink, cell, s->ex0, cover, area);

              if (fx_local == A8_SCALE) {
                s->yDlt += s->yLift;
                acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
              }
            }

            s->xDlt = s->xLift;
            acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
            cell++;
            i--;
          }

          s->fy0 = 0;
          s->fy1 = A8_SCALE;

          if (!j) {
                      if (banding) {
              s->ex0 = x_local >> A8_SHIFT;
              s->fx0 = x_local & A8_MASK;
              return s->ey0 > s->ey1;
            }
            return 1;
          }

          i = j - 1;
          j = 1;

          if (!i) {
            i = 1;
            j = 0;

            int is_last = banding ? s->ey0 > s->ey1 : 1;
            if (!is_last) continue;

            s->xDlt = ((s->ex1 << A8_SHIFT) + s->fx1) - x_local;
            s->fy1 = s->savedFy1;

            s->ex0 = x_local >> A8_SHIFT;
            s->fx0 = x_local & A8_MASK;

            if (s->fx0 + s->xDlt <= A8_SCALE) {
              cover = apply_sign((uint32_t)s->fy1, sm);
              area = cover * ((uint32_t)s->fx0 * 2 + (uint32_t)s->xDlt);
              merge(sink, cell, s->ex0, cover, area);
                          return 1;
            }
            goto HorzLeftToRightSkip;
          }
        }
      }
    }
  }

};

}  // namespace ras
}  // namespace b2amd
