// Data layouts shared by the host frontend (context.cpp) and the HIP kernels
// (kernels.hip).  Everything here is plain-old-data copied to HBM once per
// flush; the kernels read it with scalar (uniform) loads.
#pragma once

#include <stdint.h>

namespace b2amd {

// Tile geometry of the render kernel: 64 x 16 pixels, one 256-thread
// workgroup per tile, 4 pixels per thread.  A tile row is 16 lanes, which is
// one DPP row of a wave64, so the per-scanline coverage scan stays inside DPP.
constexpr int kTileW = 64;
constexpr int kTileH = 16;
constexpr int kTileThreads = 256;
// Bands with box fills only are composited by tile groups: up to 4
// consecutive tiles of one band (256 x 16 pixels) per workgroup.
constexpr int kGroupTiles = 4;

// Executed comp ops (after the frontend's simplification).
enum ExecOp : uint32_t {
  kOpSrcCopy = 0,
  kOpSrcOver = 1,
  kOpPlus = 2,
  kOpMultiply = 3,
  kOpScreen = 4,
};

// Pixel formats (BLFormat values).
enum Fmt : uint32_t {
  kFmtPRGB32 = 1,
  kFmtXRGB32 = 2,
  kFmtA8 = 3,
};

// Fetch kinds.
enum FetchKind : uint32_t {
  kFetchSolid = 0,
  kFetchLinear = 1,
  kFetchRadial = 2,
  kFetchConic = 3,
  kFetchPatternNN = 4,
  kFetchPatternBI = 5,
  kFetchPatternAligned = 6,   // pure translation, integer texel offset
  kFetchPatternFxFy = 7,      // pure translation, fractional offset, bilinear
};

// Command kinds.
enum CmdKind : uint32_t {
  kCmdAnalytic = 0,   // edges -> cells -> coverage
  kCmdBoxA = 1,       // pixel-aligned box, constant mask
  kCmdBoxU = 2,       // 24.8 box, border-coverage masks
  kCmdBoxM = 3,       // pixel-aligned box, per-pixel masks read from a mask
                      // image (box_mask[0]: its image index, the view's row 0
                      // and byte 0 at the box's top-left pixel)
};

// Why a fill command failed on the device (bits, per command).
enum CmdFail : uint32_t {
  // The line clipper met a clip-flag combination that no case handles (NaN
  // coordinates), or a curve has a non-finite control point: the reference
  // fails the fill with BL_ERROR_INVALID_GEOMETRY and reverts its edges
  // (raster/edgebuilder_p.h:1494-1496, 1584-1586,
  // raster/rastercontextops.cpp:38-45).
  kFailInvalidGeometry = 1,
  // One edge work item flattened to more than kMaxItemSegments segments
  // (control points astronomically far apart): the fill is dropped and
  // reported as BL_ERROR_OUT_OF_MEMORY instead of being flattened on to the
  // subdivision limit.
  kFailTooManySegments = 2,
};
constexpr uint32_t kMaxItemSegments = 1u << 22;

// Source geometry unit processed by one edge-builder thread.
enum SrcKind : uint32_t {
  kSrcLine = 0,
  kSrcQuad = 1,
  kSrcCubic = 2,
};

struct SrcSeg {
  uint32_t cmd;       // command index (global within the batch)
  uint32_t kind;      // SrcKind
  uint32_t i0;        // index of the start vertex
  uint32_t i1;        // index of the next vertex (control or end point)
};

// Per-command transform to 24.8 fixed point (user transform * 256).
struct CmdXform {
  double m00, m01, m10, m11, m20, m21;
  uint32_t scale_only;   // transform type <= SCALE: x*m00+m20, y*m11+m21
  uint32_t pad;
};

struct DevCmd {
  uint32_t kind;         // CmdKind
  uint32_t op;           // ExecOp
  uint32_t alpha;        // 0..255
  uint32_t rule_mask;    // 0xFFFFFFFF non-zero, 0x1FF even-odd
  uint32_t style;        // index into the style table
  uint32_t canvas;       // canvas index
  uint32_t src_fmt;      // source pixel format after simplification (pattern fetch)
  uint32_t fetch_row0;   // box fills: row the fill call starts at
  int32_t box[4];        // box fills: pixel box [x0, y0, x1, y1) that is composited
  int32_t box_row_first; // BoxU: first and last row of the unclipped mask box
  int32_t box_row_last;
  uint32_t box_mask[3];  // BoxU: per row kind (top, middle, bottom) the masks of
                         //       the left, inner and right columns, one byte each
  uint32_t pad;
};

// Per-command bounds written by the edge builder (atomics), used for tile
// culling and for the first band row of the fill call.
struct CmdBounds {
  int32_t min_x, max_x;  // cell range touched (fixed >> 8), inclusive
  int32_t min_y, max_y;  // fixed-point y range of non-horizontal segments
};

struct DevStyle {
  uint32_t kind;         // FetchKind
  uint32_t solid;        // PRGB32 (solid)
  uint32_t lut;          // LUT offset in the LUT pool (entries)
  uint32_t extend;       // 0 pad, 1 repeat, 2 reflect
  uint32_t maxi, rori;
  uint32_t img;          // image index
  uint32_t dither;       // gradients: 1 samples a 64-bit table (two words an
                         // entry at `lut`) with the pixel's Bayer threshold
  // Fetcher parameters; only those of `kind` are meaningful.
  union {
    struct {             // linear
      uint64_t lin_pt0, lin_dt, lin_dy;
    };
    struct {             // radial
      double rad_tx, rad_ty, rad_yx, rad_yy, rad_b0, rad_by, rad_dd0, rad_ddy;
      double rad_inv2a, rad_amul4, rad_sq_inv2a, rad_sq_fr;
      float rad_bd, rad_ddd;
    };
    struct {             // conic
      double con_tx, con_ty, con_yx, con_yy;
      float con_q[4];
      float con_n1, con_n2, con_n4, con_offset, con_xx;
      int32_t con_maxi, con_rori;
    };
    struct {             // affine pattern
      uint64_t pa_xx, pa_xy, pa_yx, pa_yy, pa_tx, pa_ty;
      int32_t pa_ox, pa_oy, pa_rx, pa_ry, pa_minx, pa_miny, pa_maxx, pa_maxy;
      int32_t pa_corx, pa_cory, pa_tw, pa_th;
    };
    struct {             // translated pattern (aligned / FxFy)
      int32_t sp_tx, sp_ty;    // texel offset added to the pixel position
      int32_t sp_rx, sp_ry;    // repeat (w) or reflect (2w) period, 0 = pad
      int32_t sp_w, sp_h;      // image size
      uint32_t sp_wa, sp_wb, sp_wc, sp_wd;   // FxFy weights, sum <= 256
    };
  };
};
static_assert(sizeof(DevStyle) == 136, "DevStyle layout");

struct DevImage {
  uint64_t ptr;          // device address of pixel row 0
  int64_t stride;        // bytes
  uint32_t w, h, format, pad;
};

// A render target within a batch.
struct DevCanvas {
  uint64_t ptr;          // device address of pixel row 0
  int64_t stride;        // bytes
  uint32_t w, h, format;
  uint32_t tiles_x, bands;        // tile grid
  uint32_t ref_band_h;            // band height the reference would use (conic row origin)
  uint32_t cmd_base, n_cmds;      // commands of this canvas: [cmd_base, cmd_base + n_cmds)
  uint32_t band_base;             // first global band index
  uint32_t tile_base;             // first global tile index
  uint64_t mat_base;              // (band, cmd) matrix offset
  uint32_t group_base;            // first global tile group (4 tiles of a band)
  uint32_t groups_x;              // tile groups per band
};

// One emitted edge segment, 24.8 fixed point, y0 < y1.
struct DevSeg {
  int32_t x0, y0, x1, y1;
};

}  // namespace b2amd
