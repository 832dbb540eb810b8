// Launch interface of the fill-path kernels (kernels.hip), used by the host
// frontend (context.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "device_types.h"
#include "lut_device.h"

namespace b2amd {

// Device views of one flush ("frame") of recorded commands.
struct FrameDev {
  const DevCanvas* canvases; uint32_t n_canvases;
  const DevCmd* cmds; uint32_t n_cmds;
  const CmdXform* xforms;
  const SrcSeg* src; uint32_t n_src;
  const double* vtx;
  const DevStyle* styles;
  const uint32_t* luts;
  uint32_t dither;         // some style samples a dithered gradient (64-bit table)
  const DevImage* images;
  uint32_t n_tiles;
  uint32_t n_groups;       // tile groups: up to 4 consecutive tiles of one band
  uint32_t n_bands;        // bands over all canvases
  uint64_t n_mat;          // (band, cmd) matrix entries over all canvases
  double tol_sq;           // squared flattening tolerance in 24.8 units
};

// Scratch buffers the pipeline needs; sized by the caller.
// One edge work item (see k_items_count): up to 4 points (x, y pairs).
struct EdgeItem {
  double p[8];
  uint32_t cmd, kind, any, pad;
};

struct FrameScratch {
  uint32_t* item_cnt;      // n_src + 1: edge work items per source segment
  uint32_t* item_off;      // n_src + 1: their exclusive scan
  EdgeItem* items;         // <= item cap: the work items
  uint32_t* src_count;     // item cap + 1: edge segments per work item (pool mode: per source)
  uint32_t* src_off;       // item cap + 1: their exclusive scan
  CmdBounds* bounds;       // n_cmds
  DevSeg* segs;            // n_segs
  uint32_t* seg_cmd;       // n_segs
  uint32_t* mat_cnt;       // n_mat + 1
  uint32_t* mat_off;       // n_mat + 1
  uint32_t* mat_pres;      // n_mat + 1
  uint32_t* list_off;      // n_mat + 1
  uint4* list_ent;         // n_list: per listed (band, command) cell {command, tile-column range, binned offset, binned count}
  uint32_t* list_style;    // n_list: style index of the entry's command
  DevSeg* binned;          // n_binned
  uint32_t* scan_tmp;      // scan_tmp_words(max n)
  unsigned long long* pixels;  // composited pixels (mask > 0), summed after render
  uint32_t* cmd_tiles;     // n_cmds: tile-column range per command
  // Tiles of bands with analytic fills, tile groups of bands with solid box
  // fills only and tile groups of the other box-only bands, each compacted in
  // band order: per band the tile (group) counts and their exclusive scans
  // (totals at [n_bands]).
  uint32_t* band_nonsolid; // n_bands + 1: a box fill with a non-solid style touches the band
  uint32_t* band_slow;     // n_bands + 1: a box fill other than a solid SRC_OVER / SRC_COPY aligned box on a
                           // 32-bit canvas touches the band
  uint32_t* band_tiles;    // n_bands + 1
  uint32_t* band_sgroups;  // n_bands + 1
  uint32_t* band_ggroups;  // n_bands + 1
  uint32_t* tile_off;      // n_bands + 1
  uint32_t* sgroup_off;    // n_bands + 1
  uint32_t* ggroup_off;    // n_bands + 1
  uint32_t* band_fgroups;  // n_bands + 1: tile groups of bands with fast box fills only
  uint32_t* fgroup_off;    // n_bands + 1
  uint4* tile_hdr;         // <= n_tiles: {canvas, band list begin, band list end, tile column | band << 16}
  uint4* group_hdr;        // <= 2 n_groups: two header words per box tile group (see k_group_setup),
                           // the solid bands' groups first
  unsigned long long* tile_px; // n_tiles: composited pixels of each tile
  // Failed fills: per command the kFailInvalidGeometry / kFailTooManySegments
  // bits of its work items (the command's segments are not binned), and one
  // word set when any command failed.  Both start at zero.
  uint32_t* cmd_err;       // n_cmds
  uint32_t* err_any;       // 1
  // Pool mode (sizes predicted ahead of time; null / 0 otherwise):
  uint32_t* seg_cursor;    // segments taken from the pool
  uint32_t* flags;         // bit 0: segment pool overflow, bit 1: bin pool overflow
  uint32_t binned_cap;     // capacity of `binned`
  // Work queue of the analytic tiles (k_render_queue; null otherwise):
  uint32_t* queue;         // <= n_tiles: tile slots in queue order (null: band order)
  uint32_t* queue_head;    // 1: next queue position
  uint32_t* queue_hist;    // kQueueBuckets: bucket sizes, then bucket cursors
  uint32_t* tile_key;      // <= n_tiles: bucket of each slot
  float* tile_feat;        // <= n_tiles x 16: tile cost-model features
  uint32_t* tile_cycles;   // <= n_tiles: logged wall-clock ticks per slot (null: not logged)
};

size_t scan_tmp_words(size_t n);
void scan_exclusive(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp, hipStream_t s);

// Stage 1: edge building.  Work items (lines, monotonic curve pieces) per
// source segment, scanned and written; then count == true writes per-item
// segment counts and command bounds, count == false writes the segments at
// the scanned offsets.  `cap` bounds the number of items.
void launch_edge_items(const FrameDev& f, const FrameScratch& s, hipStream_t st);
void launch_edges(const FrameDev& f, const FrameScratch& s, bool count, uint32_t cap, hipStream_t st);
// Single pass into a pool of `cap` segments (pool mode).
void launch_edges_pool(const FrameDev& f, const FrameScratch& s, uint32_t cap, uint32_t item_cap, hipStream_t st);
// Stage 2: binning of segments into (band, command) cells.
void launch_bin_count(const FrameDev& f, const FrameScratch& s, uint32_t n_segs, hipStream_t st);
void launch_presence(const FrameDev& f, const FrameScratch& s, hipStream_t st);
void launch_bin_scatter(const FrameDev& f, const FrameScratch& s, uint32_t n_segs, uint32_t* cursor, hipStream_t st);
void launch_list_build(const FrameDev& f, const FrameScratch& s, hipStream_t st);
void launch_cmd_tiles(const FrameDev& f, const FrameScratch& s, hipStream_t st);
// Stage 3: band classes (after the binned scan), per-tile and per-group
// headers (setup, after the list build), tile rasterization + fetch +
// composite (render: n_a analytic tiles, n_s solid-box and n_g other box
// tile groups, or upper bounds of them), composited pixels summed into *s.pixels, which must
// start at zero (stats).
void launch_band_class(const FrameDev& f, const FrameScratch& s, hipStream_t st);
void launch_render_setup(const FrameDev& f, const FrameScratch& s, hipStream_t st);
void launch_render(const FrameDev& f, const FrameScratch& s, uint32_t n_a, uint32_t n_f, uint32_t n_s, uint32_t n_g,
                   hipStream_t st);
void launch_render_stats(const FrameDev& f, const FrameScratch& s, hipStream_t st);
// Work-queue form of the analytic-tile render: features of the (at most n_a)
// analytic tiles, the queue sorted by predicted cost (`pred`: the tile cost
// model's outputs, n_a x 2, or null for the segment-count heuristic), and
// the persistent render over the queue with `grid` workgroups, `wgs_per_cu`
// of them per CU (s.queue null: band order).  Box tile groups still go through launch_render.
void launch_tile_features(const FrameDev& f, const FrameScratch& s, uint32_t n_a, hipStream_t st);
void launch_tile_queue(const FrameDev& f, const FrameScratch& s, uint32_t n_a, const float* pred, hipStream_t st);
void launch_render_queue(const FrameDev& f, const FrameScratch& s, uint32_t grid, uint32_t wgs_per_cu, hipStream_t st);

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
// Device-to-device copy of n words by a kernel (a graph node like any other).
void launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t st);
// Fills the gradient tables of `n_jobs` jobs (one workgroup per table).
void launch_gradient_luts(const lut::Job* jobs, const lut::Stop* stops, uint32_t n_jobs, uint32_t* luts, hipStream_t st);

// The device scratch region of bl_amd_context_replay_noalloc on device
// `dev`: its address (looked up once per device) and size in bytes.
hipError_t noalloc_scratch(int dev, void** ptr, size_t* bytes);
void launch_init_bounds(CmdBounds* b, uint32_t n, hipStream_t st);

}  // namespace b2amd
