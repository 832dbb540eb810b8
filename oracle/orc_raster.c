/* CPU oracle -- analytic rasterizer, band driver and fill loops.
 *
 * TEST INFRASTRUCTURE (see orc_core.h).  Restates
 *   raster/analyticrasterizer_p.h   (prepare_ref, advanceToY, rasterize, cell_merge)
 *   raster/rendercommandprocsync_p.h:87-254 (band loop with active edges)
 *   pipeline/reference/fillgeneric_p.h  (FillAnalytic, FillBoxA, FillMask)
 *   pipeline/pipedefs_p.h:644-815   (unaligned box -> mask rows)
 * The shadow bit-vector of the reference only lets the compositor skip zero
 * cells; compositing with mask 0 is the identity for every comp op handled
 * here, so the fill loop below visits every pixel of [0, x_end) instead.
 */
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"

enum { A8_SHIFT = 8, A8_SCALE = 256, A8_MASK = 255 };
enum { F_INITIAL = 1, F_VERT_OR_SINGLE = 2, F_RTL = 4 };

typedef struct {
  int ex0, ey0, ex1, ey1;
  int fx0, fy0, fx1, fy1;
  int xErr, yErr, xDlt, yDlt, xRem, yRem, xLift, yLift;
  int dx, dy, savedFy1;
  uint32_t flags;
} RasState;

typedef struct {
  RasState s;
  uint32_t* cells;            /* band cell buffer, row 0 = band_offset */
  size_t cell_stride;         /* in cells */
  uint32_t sign_mask;
  uint32_t band_height, band_offset, band_end;
  uint32_t cell_min_x, cell_max_x;
} Ras;

/* analyticrasterizer_p.h:57-100 */
static inline uint32_t apply_sign(uint32_t x, uint32_t mask) { return (x ^ mask) - mask; }

static inline void acc_err_step(int* acc, int* iter, int step, int correction) {
  *iter -= step;
  int mask = *iter >> 31;
  *acc -= mask;
  *iter += mask & correction;
}

static inline void acc_err_multi_step(int* acc, int* iter, int step, int correction, int count) {
  int64_t i = (int64_t)(uint32_t)(*iter);
  i -= (int64_t)((uint64_t)(uint32_t)step * (uint32_t)count);
  if (i < 0) {
    int n = (int)(((uint64_t)(-i) + (uint32_t)correction - 1u) / (uint64_t)(uint32_t)correction);
    *acc += n;
    i += (int64_t)correction * n;
  }
  *iter = (int)i;
}

static uint32_t g_touch_min;   /* smallest cell index written in the current band */

static inline void cell_merge(Ras* r, uint32_t* row, int x, uint32_t cover, uint32_t area) {
  (void)r;
  if ((unsigned)x < g_touch_min) g_touch_min = (unsigned)x;
  row[(unsigned)x + 0] += (cover << 9) - area;
  row[(unsigned)x + 1] += area;
}
static inline void cell_add(uint32_t* row, int x, uint32_t v) {
  if ((unsigned)x < g_touch_min) g_touch_min = (unsigned)x;
  row[(unsigned)x] += v;
}

static inline void upd_min(Ras* r, int x) { if ((unsigned)x < r->cell_min_x) r->cell_min_x = (unsigned)x; }
static inline void upd_max(Ras* r, int x) { if ((unsigned)x > r->cell_max_x) r->cell_max_x = (unsigned)x; }

/* analyticrasterizer_p.h:289-360 */
static int ras_prepare(Ras* r, int p0x, int p0y, int p1x, int p1y) {
  RasState* s = &r->s;
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

  s->xLift = (int)(x_base / (unsigned)s->dy);
  s->xRem = (int)(x_base % (unsigned)s->dy);
  s->yLift = (int)(y_base / (unsigned)s->dx);
  s->yRem = (int)(y_base % (unsigned)s->dx);

  s->xDlt = s->dx;
  s->yDlt = s->dy;
  s->xErr = (s->dy >> 1) - 1;
  s->yErr = (s->dx >> 1) - 1;

  if (s->ey0 != s->ey1) {
    uint64_t p = (uint64_t)(A8_SCALE - (uint32_t)s->fy0) * (uint32_t)s->dx;
    s->xDlt = (int)(p / (unsigned)s->dy);
    s->xErr -= (int)(p % (unsigned)s->dy);
    acc_err_step(&s->xDlt, &s->xErr, 0, s->dy);
  }

  if (s->ex0 != s->ex1) {
    uint64_t p = (uint64_t)((s->flags & F_RTL) ? (uint32_t)s->fx0 : A8_SCALE - (uint32_t)s->fx0) * (uint32_t)s->dy;
    s->yDlt = (int)(p / (unsigned)s->dx);
    s->yErr -= (int)(p % (unsigned)s->dx);
    acc_err_step(&s->yDlt, &s->yErr, 0, s->dx);
  }

  s->yDlt += s->fy0;
  return 1;
}

/* analyticrasterizer_p.h:371-459 */
static void ras_advance_to_y(Ras* r, int y_target) {
  RasState* s = &r->s;
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
static int ras_rasterize(Ras* r, int banding) {
  RasState* s = &r->s;
  int ey1_end = s->ey1;
  if (banding && ey1_end > (int)r->band_end) ey1_end = (int)r->band_end;

  size_t i = (unsigned)ey1_end - (unsigned)s->ey0;
  uint32_t y_offset = (unsigned)s->ey0 - r->band_offset;
  uint32_t* cell = r->cells + (size_t)y_offset * r->cell_stride;
  const size_t cs = r->cell_stride;

  if (banding) s->ey0 += (int)i + 1;

  const uint32_t sm = r->sign_mask;
  const uint32_t full_cover = apply_sign(A8_SCALE, sm);

  if (s->flags & F_VERT_OR_SINGLE) {
    uint32_t cover;
    uint32_t area = (uint32_t)s->fx0 + (uint32_t)s->fx1;
    upd_min(r, s->ex0);
    upd_max(r, s->ex0);

    cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
    cell_merge(r, cell, s->ex0, cover, cover * area);

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

    cell += cs;
    cover = full_cover;
    while (--i) {
      cell_merge(r, cell, s->ex0, cover, cover * area);
      cell += cs;
    }

    if (banding) {
      if (s->ey0 <= s->ey1) {
        cell_merge(r, cell, s->ex0, cover, cover * area);
        s->fy0 = 0;
        s->fy1 = s->ey0 == s->ey1 ? s->savedFy1 : A8_SCALE;
        return 0;
      }
    }

    cover = apply_sign((uint32_t)s->savedFy1, sm);
    cell_merge(r, cell, s->ex0, cover, cover * area);
    return 1;
  }
  else if (s->dy >= s->dx) {
    if (banding) i += (s->ey0 <= s->ey1);

    if (s->flags & F_RTL) {
      upd_max(r, s->ex0);
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
          cell_merge(r, cell, s->ex0 + 1, cov0, area);

          cov0 = apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
          area = cov0 * ((uint32_t)s->fx0 + A8_SCALE);
          cell_merge(r, cell, s->ex0, cov0, area);

          acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
          s->yDlt += s->yLift;
        }
        else {
VertRightToLeftSingleFirstOrLast:
          cov0 = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
          area = cov0 * (area + (uint32_t)s->fx0);
          cell_merge(r, cell, s->ex0, cov0, area);
        }

        s->fy0 = 0;
        cell += cs;

        if (!i) {
          upd_min(r, s->ex0);
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
            cell_add(cell, s->ex0 + 2, area);

            cov0 = full_cover - cov1;
            cov1 = (cov1 << 9) - area;
            area = cov0 * ((uint32_t)s->fx0 + A8_SCALE);

            cov0 = (cov0 << 9) - area;
            cov1 = cov1 + area;

            cell_add(cell, s->ex0 + 0, cov0);
            cell_add(cell, s->ex0 + 1, cov1);
            cell += cs;

            acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            s->yDlt += s->yLift;
          }
          else {
VertRightToLeftSingleInLoop:
            area = full_cover * (area + (uint32_t)s->fx0);
            cell_merge(r, cell, s->ex0, full_cover, area);
            cell += cs;
          }
        }

        if (banding) {
          if (s->ey0 >= s->ey1) {
            s->fy1 = s->savedFy1;
            s->xDlt = ((s->ex0 - s->ex1) << A8_SHIFT) + s->fx0 - s->fx1;
            if (s->ey0 == s->ey1) {
              upd_min(r, s->ex0);
              return 0;
            }
          }
          else {
            upd_min(r, s->ex0);
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
      upd_min(r, s->ex0);
      for (;;) {
        uint32_t area = (uint32_t)s->fx0;
        uint32_t cov0, cov1;

        s->fx0 += s->xDlt;

        if (s->fx0 <= A8_SCALE) {
          cov0 = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
          area = cov0 * (area + (uint32_t)s->fx0);
          cell_merge(r, cell, s->ex0, cov0, area);

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
          cell_merge(r, cell, s->ex0 - 1, cov0, area);

          cov0 = apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
          area = cov0 * (uint32_t)s->fx0;
          cell_merge(r, cell, s->ex0, cov0, area);

          s->yDlt += s->yLift;
          acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
        }

        s->fy0 = 0;
        cell += cs;

        if (!i) {
          upd_max(r, s->ex0);
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
            cell_merge(r, cell, s->ex0, full_cover, area);
            cell += cs;

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
            cell_add(cell, s->ex0 + 0, cov0);
            s->ex0++;

            cov0 = apply_sign(A8_SCALE - (uint32_t)s->yDlt, sm);
            area = cov0 * (uint32_t)s->fx0;

            cov0 = (cov0 << 9) - area + cov1;
            cell_add(cell, s->ex0 + 0, cov0);
            cell_add(cell, s->ex0 + 1, area);
            cell += cs;
          }

          s->yDlt += s->yLift;
          acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
        }

        if (banding) {
          if (s->ey0 >= s->ey1) {
            s->fy1 = s->savedFy1;
            s->xDlt = ((s->ex1 - s->ex0) << A8_SHIFT) + s->fx1 - s->fx0;
            if (s->ey0 == s->ey1) {
              upd_max(r, s->ex0);
              return 0;
            }
          }
          else {
            upd_max(r, s->ex0);
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
      upd_max(r, s->ex0);

      if (s->flags & F_INITIAL) {
        s->flags &= ~(uint32_t)F_INITIAL;
        j = i;
        i = 1;

        cover = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
        if (s->fx0 - s->xDlt < 0) goto HorzRightToLeftInside;

        x_local -= s->xDlt;
        cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
        area = cover * (uint32_t)(s->fx0 * 2 - s->xDlt);
        cell_merge(r, cell, s->ex0, cover, area);

        if ((x_local & A8_MASK) == 0) {
          s->yDlt += s->yLift;
          acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
        }

        s->xDlt = s->xLift;
        acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
        cell += cs;
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
            while (s->ex0 != ex_local) {
              cell_merge(r, cell, s->ex0, cover, area);

              cover = (uint32_t)s->yLift;
              acc_err_step((int*)&cover, &s->yErr, s->yRem, s->dx);
              s->yDlt += (int32_t)cover;

              cover = apply_sign(cover, sm);
              area = cover * A8_SCALE;
              s->ex0--;
            }

            cover += apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
            area = cover * ((uint32_t)fx_local + A8_SCALE);
            cell_merge(r, cell, s->ex0, cover, area);

            if (fx_local == 0) {
              s->yDlt += s->yLift;
              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            }
          }

          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
          cell += cs;
          i--;
        }

        s->fy0 = 0;
        s->fy1 = A8_SCALE;

        if (!j) {
          upd_min(r, s->ex0);
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
            cell_merge(r, cell, s->ex0, cover, area);
            upd_min(r, s->ex0);
            return 1;
          }
          goto HorzRightToLeftSkip;
        }
      }
    }
    else {
      upd_min(r, s->ex0);

      if (s->flags & F_INITIAL) {
        s->flags &= ~(uint32_t)F_INITIAL;
        j = i;
        i = 1;

        cover = apply_sign((uint32_t)(s->yDlt - s->fy0), sm);
        if (s->fx0 + s->xDlt > A8_SCALE) goto HorzLeftToRightInside;

        x_local += s->xDlt;
        cover = apply_sign((uint32_t)(s->fy1 - s->fy0), sm);
        area = cover * ((uint32_t)s->fx0 * 2 + (uint32_t)s->xDlt);
        cell_merge(r, cell, s->ex0, cover, area);

        if (s->fx0 + s->xDlt == A8_SCALE) {
          s->yDlt += s->yLift;
          acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
        }

        s->xDlt = s->xLift;
        acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
        cell += cs;
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
            while (s->ex0 != ex_local) {
              cell_merge(r, cell, s->ex0, cover, area);

              cover = (uint32_t)s->yLift;
              acc_err_step((int*)&cover, &s->yErr, s->yRem, s->dx);
              s->yDlt += (int32_t)cover;

              cover = apply_sign(cover, sm);
              area = cover * A8_SCALE;
              s->ex0++;
            }

            cover += apply_sign((uint32_t)(s->fy1 - s->yDlt), sm);
            area = cover * (uint32_t)fx_local;
            cell_merge(r, cell, s->ex0, cover, area);

            if (fx_local == A8_SCALE) {
              s->yDlt += s->yLift;
              acc_err_step(&s->yDlt, &s->yErr, s->yRem, s->dx);
            }
          }

          s->xDlt = s->xLift;
          acc_err_step(&s->xDlt, &s->xErr, s->xRem, s->dy);
          cell += cs;
          i--;
        }

        s->fy0 = 0;
        s->fy1 = A8_SCALE;

        if (!j) {
          upd_max(r, s->ex0);
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
            cell_merge(r, cell, s->ex0, cover, area);
            upd_max(r, s->ex0);
            return 1;
          }
          goto HorzLeftToRightSkip;
        }
      }
    }
  }
}

/* Test hook: the invariant of raster/analyticrasterizer_test.cpp:34-158.
 * For each edge (x0, y0, x1, y1; 24.8, y0 < y1) one rasterizer is stepped
 * band by band with rasterize() while a second one jumps to the same row
 * with advance_to_y(); their states must agree after every band (yDlt is
 * compared modulo 256 when dy >= dx).  Returns the number of mismatches. */
static int ras_state_equal(const RasState* a, const RasState* b) {
  int m = (a->dy >= a->dx) ? 255 : -1;
  return a->dx == b->dx && a->dy == b->dy && a->ex0 == b->ex0 && a->ey0 == b->ey0 &&
         a->ex1 == b->ex1 && a->ey1 == b->ey1 && a->fx0 == b->fx0 && a->fy0 == b->fy0 &&
         a->fx1 == b->fx1 && a->fy1 == b->fy1 && a->xErr == b->xErr && a->yErr == b->yErr &&
         a->xDlt == b->xDlt && (a->yDlt & m) == (b->yDlt & m) && a->xRem == b->xRem &&
         a->yRem == b->yRem && a->xLift == b->xLift && a->yLift == b->yLift && a->savedFy1 == b->savedFy1;
}

size_t orc_raster_advance_check(const int32_t* edges, size_t n, uint32_t band_height, uint32_t w) {
  const size_t stride = (size_t)w + 2;
  uint32_t* cells = (uint32_t*)calloc(stride * band_height, sizeof(uint32_t));
  size_t bad = 0;
  if (!cells) return (size_t)-1;
  for (size_t i = 0; i < n; i++) {
    const int x0 = edges[4 * i + 0], y0 = edges[4 * i + 1], x1 = edges[4 * i + 2], y1 = edges[4 * i + 3];
    Ras a, b;
    memset(&a, 0, sizeof(a));
    memset(&b, 0, sizeof(b));
    a.cells = b.cells = cells;
    a.cell_stride = b.cell_stride = stride;
    a.band_height = b.band_height = band_height;
    uint32_t band_y0 = (uint32_t)(y0 >> A8_SHIFT), band_y1 = band_y0 + band_height;
    a.band_offset = b.band_offset = band_y0;
    a.band_end = b.band_end = band_y0 + band_height - 1;
    a.cell_min_x = b.cell_min_x = 0xFFFFFFFFu;
    int pa = ras_prepare(&a, x0, y0, x1, y1), pb = ras_prepare(&b, x0, y0, x1, y1);
    if (!pa || !pb || !ras_state_equal(&a.s, &b.s)) { bad++; continue; }
    for (int guard = 0; guard < 1 << 20; guard++) {
      if (ras_rasterize(&a, 1)) break;
      ras_advance_to_y(&b, (int)band_y1);
      if (!ras_state_equal(&a.s, &b.s)) { bad++; break; }
      band_y0 = band_y1;
      band_y1 += band_height;
      uint32_t last = (uint32_t)((y1 - 1) >> A8_SHIFT);
      a.band_offset = b.band_offset = band_y0;
      a.band_end = b.band_end = (band_y1 - 1 < last) ? band_y1 - 1 : last;
    }
  }
  free(cells);
  return bad;
}

/* Test hook: the cells one segment (24.8, y0 < y1, sign 1 when its source
 * went up) leaves in a cell buffer of `rows` rows of `stride` cells starting
 * at pixel row 0, rasterized in one pass.  Returns 0 when prepare() rejects
 * it (horizontal). */
int orc_raster_segment(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign, uint32_t* cells, size_t stride,
                       uint32_t rows) {
  Ras r;
  memset(&r, 0, sizeof(r));
  r.cells = cells;
  r.cell_stride = stride;
  r.sign_mask = (uint32_t)0 - sign;
  r.band_height = rows;
  r.band_offset = 0;
  r.band_end = rows - 1;
  r.cell_min_x = UINT32_MAX;
  g_touch_min = UINT32_MAX;
  if (!ras_prepare(&r, x0, y0, x1, y1)) return 0;
  ras_rasterize(&r, 0);
  return 1;
}

/* rastercontext.cpp:4154-4193 with thread_count == 0 */
uint32_t orc_band_height(uint32_t width) {
  uint32_t bh = 64;
  size_t limit = 1024 * 256;
  size_t pixel_count = (size_t)width * bh;
  do {
    if (pixel_count * 4 <= limit) break;
    bh >>= 1;
    pixel_count >>= 1;
  } while (bh > 8);
  return bh;
}

/* fillgeneric_p.h:381-387 */
static inline uint32_t calc_mask(uint32_t cov, uint32_t rule_mask, uint32_t alpha) {
  uint32_t m = ((uint32_t)((int32_t)cov >> 9) & rule_mask) - 256u;
  int32_t mi = (int32_t)m;
  uint32_t a = (uint32_t)(mi < 0 ? -mi : mi);
  if (a > 256u) a = 256u;
  return (a * alpha) >> 8;
}

typedef struct {
  RasState state;
  uint32_t sign;
  const OrcSeg* cur;
  const OrcSeg* end;
} Active;

/* Runs the fill function for one band (fillgeneric_p.h:174-379). */
static void fill_band(OrcCanvas* c, OrcPipe* p, Ras* r, uint32_t alpha, uint32_t rule_mask,
                      uint32_t y0, uint32_t y1) {
  uint32_t x_end = r->cell_max_x + 1;
  x_end = (x_end + 3u) & ~3u;
  if (x_end > c->w) x_end = c->w;

  orc_pipe_begin(p, y0);
  for (uint32_t y = y0; y < y1; y++) {
    uint32_t* row = r->cells + (size_t)(y - r->band_offset) * r->cell_stride;
    orc_pipe_row(p, y);
    uint32_t cov = 256u << 9;
    /* cells left of the first written cell are zero, so the prefix starts there */
    for (uint32_t x = g_touch_min; x < x_end; x++) {
      cov += row[x];
      uint32_t m = calc_mask(cov, rule_mask, alpha);
      if (m) orc_pipe_composite(p, c, x, y, m);
    }
    memset(row, 0, r->cell_stride * sizeof(uint32_t));
  }
  g_touch_min = UINT32_MAX;
}

/* rendercommandprocsync_p.h:87-254.  Segments are fed per source chain in
 * the reference; each segment is an independent edge here (a chain of one
 * point pair), which produces the same prepare/rasterize calls. */
void orc_fill_analytic(OrcCanvas* c, OrcEdges* e, OrcPipe* p, uint32_t alpha, uint32_t fill_rule) {
  if (e->n == 0 || e->bbox_y0 >= e->bbox_y1) return;

  const uint32_t bh = orc_band_height(c->w);
  uint32_t shift = 0;
  while ((1u << shift) < bh) shift++;
  const uint32_t band_count = (c->h + bh - 1) >> shift;

  const uint32_t y_start = (uint32_t)e->bbox_y0 >> 8;
  const uint32_t y_end = ((uint32_t)e->bbox_y1 + 255u) >> 8;
  const uint32_t rule_mask = fill_rule == 0 ? 0xFFFFFFFFu : 0x1FFu;

  size_t required_width = ((size_t)c->w + 1u + 4u + 3u) & ~(size_t)3u;
  Ras r;
  memset(&r, 0, sizeof(r));
  r.cell_stride = required_width;
  r.cells = (uint32_t*)calloc(required_width * bh, sizeof(uint32_t));
  r.band_height = bh;

  /* Sort segments by their top band, the order edges are linked in. */
  uint32_t band_id = (uint32_t)e->bbox_y0 >> (shift + 8);
  uint32_t band_end = ((uint32_t)e->bbox_y1 >> (shift + 8)) + 1;
  if (band_end > band_count) band_end = band_count;

  size_t nb = band_count + 1;
  size_t* start = (size_t*)calloc(nb + 1, sizeof(size_t));
  for (size_t k = 0; k < e->n; k++) {
    uint32_t b = (uint32_t)e->segs[k].y0 >> (shift + 8);
    if (b > band_count) b = band_count;
    start[b + 1]++;
  }
  for (size_t k = 0; k < nb; k++) start[k + 1] += start[k];
  const OrcSeg** by_band = (const OrcSeg**)malloc(e->n * sizeof(OrcSeg*));
  size_t* fillp = (size_t*)malloc(nb * sizeof(size_t));
  memcpy(fillp, start, nb * sizeof(size_t));
  for (size_t k = 0; k < e->n; k++) {
    uint32_t b = (uint32_t)e->segs[k].y0 >> (shift + 8);
    if (b > band_count) b = band_count;
    by_band[fillp[b]++] = &e->segs[k];
  }

  Active* active = (Active*)malloc(e->n * sizeof(Active));
  size_t n_active = 0;

  r.band_offset = y_start;
  g_touch_min = UINT32_MAX;
  do {
    r.cell_min_x = UINT32_MAX;
    r.cell_max_x = 0;
    uint32_t band_last = (band_id + 1) * bh;
    if (band_last > y_end) band_last = y_end;
    r.band_end = band_last - 1;

    /* continue edges from previous bands */
    size_t kept = 0;
    for (size_t k = 0; k < n_active; k++) {
      Active* a = &active[k];
      r.s = a->state;
      r.sign_mask = (uint32_t)0 - a->sign;
      if (ras_rasterize(&r, 1)) continue;
      a->state = r.s;
      active[kept++] = *a;
    }
    n_active = kept;

    /* new edges starting in this band */
    for (size_t k = start[band_id]; k < start[band_id + 1]; k++) {
      const OrcSeg* sg = by_band[k];
      r.sign_mask = (uint32_t)0 - sg->sign;
      if (!ras_prepare(&r, sg->x0, sg->y0, sg->x1, sg->y1)) continue;
      if ((uint32_t)r.s.ey1 <= r.band_end) {
        ras_rasterize(&r, 0);
      }
      else {
        if ((uint32_t)r.s.ey0 <= r.band_end) {
          if (ras_rasterize(&r, 1)) continue;
        }
        Active* a = &active[n_active++];
        a->state = r.s;
        a->sign = sg->sign;
      }
    }

    if (r.cell_min_x <= r.cell_max_x)
      fill_band(c, p, &r, alpha, rule_mask, r.band_offset, r.band_end + 1);

    r.band_offset = (r.band_offset + bh) & ~(bh - 1);
  } while (++band_id < band_end);

  free(active);
  free(fillp);
  free(by_band);
  free(start);
  free(r.cells);
}

/* FillBoxA (fillgeneric_p.h:22-65): constant mask `alpha`. */
void orc_fill_box_a(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1) {
  orc_pipe_begin(p, (uint32_t)y0);
  for (int y = y0; y < y1; y++) {
    orc_pipe_row(p, (uint32_t)y);
    for (int x = x0; x < x1; x++) orc_pipe_composite(p, c, (uint32_t)x, (uint32_t)y, alpha);
  }
}

/* FillMask_Base with one VMaskA8 command repeated over the rows
 * (rendercommandprocsync_p.h:65-84, fillgeneric_p.h:67-158,
 * compopgeneric_p.h:165-183): the mask of pixel x of row y is byte x - x0 of
 * mask row y - y0, multiplied by `alpha` (udiv255) below full alpha. */
void orc_fill_box_masked(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1,
                         const uint8_t* mask, intptr_t mask_stride) {
  orc_pipe_begin(p, (uint32_t)y0);
  for (int y = y0; y < y1; y++) {
    orc_pipe_row(p, (uint32_t)y);
    const uint8_t* row = mask + (intptr_t)(y - y0) * mask_stride;
    for (int x = x0; x < x1; x++) {
      uint32_t v = row[x - x0];
      orc_pipe_composite(p, c, (uint32_t)x, (uint32_t)y, alpha >= 255u ? v : orc_div255(v * alpha));
    }
  }
}

/* FillData::init_box_u_8bpc_24x8 (pipedefs_p.h:644-815) followed by
 * FillMask_Base: the per-pixel mask is (fx(x) * fy(y) * alpha) >> 16 with
 * border coverages fx/fy; rows whose interior mask is zero are skipped, which
 * moves the fetcher's start row. */
void orc_fill_box_u(OrcCanvas* c, OrcPipe* p, uint32_t alpha, int x0, int y0, int x1, int y1) {
  uint32_t ax0 = (uint32_t)x0 >> 8, ay0 = (uint32_t)y0 >> 8;
  uint32_t ax1 = (uint32_t)(x1 + 0xFF) >> 8, ay1 = (uint32_t)(y1 + 0xFF) >> 8;
  uint32_t fx0 = (uint32_t)x0 & 0xFFu, fy0 = (uint32_t)y0 & 0xFFu;
  uint32_t fx1 = ((uint32_t)(x1 - 1) & 0xFFu) + 1u, fy1 = ((uint32_t)(y1 - 1) & 0xFFu) + 1u;
  uint32_t w = ax1 - ax0, h = ay1 - ay0;

  fy0 = (h == 1 ? fy1 : 256u) - fy0;
  uint32_t fy0_a = fy0 * alpha, fy1_a = fy1 * alpha;

  /* row masks for the three row kinds: top, middle, bottom */
  uint32_t row_m[3][3]; /* [row kind][left, inner, right] */
  if (w == 1) {
    uint32_t fx = fx1 - fx0;
    row_m[0][0] = row_m[0][1] = row_m[0][2] = (fx * fy0_a) >> 16;
    row_m[1][0] = row_m[1][1] = row_m[1][2] = (fx * alpha) >> 8;
    row_m[2][0] = row_m[2][1] = row_m[2][2] = (fx * fy1_a) >> 16;
  }
  else {
    uint32_t lx = 256 - fx0;
    row_m[0][0] = (lx * fy0_a) >> 16; row_m[0][1] = fy0_a >> 8; row_m[0][2] = (fx1 * fy0_a) >> 16;
    row_m[1][0] = (lx * alpha) >> 8;  row_m[1][1] = alpha;      row_m[1][2] = (fx1 * alpha) >> 8;
    row_m[2][0] = (lx * fy1_a) >> 16; row_m[2][1] = fy1_a >> 8; row_m[2][2] = (fx1 * fy1_a) >> 16;
  }

  uint32_t by0 = ay0, by1 = ay0 + h;
  uint32_t top_inner = w == 1 ? row_m[0][0] : row_m[0][1];
  uint32_t mid_inner = w == 1 ? row_m[1][0] : row_m[1][1];
  uint32_t bot_inner = w == 1 ? row_m[2][0] : row_m[2][1];
  if (h == 1) {
    if (top_inner == 0) return;
  }
  else {
    by0 += (uint32_t)(top_inner == 0);
    by1 -= (uint32_t)(bot_inner == 0);
    if (!(by0 < by1)) return;
    if (w == 1 && mid_inner == 0) return;
  }

  orc_pipe_begin(p, by0);
  for (uint32_t y = by0; y < by1; y++) {
    int kind = (y == ay0) ? 0 : (y == ay0 + h - 1) ? 2 : 1;
    orc_pipe_row(p, y);
    for (uint32_t x = ax0; x < ax1; x++) {
      int col = (w == 1) ? 0 : (x == ax0) ? 0 : (x == ax1 - 1) ? 2 : 1;
      uint32_t m = row_m[kind][col];
      if (m) orc_pipe_composite(p, c, x, y, m);
    }
  }
}
