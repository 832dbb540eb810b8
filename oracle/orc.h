/* CPU oracle for the Blend2D fill path -- public entry points (ctypes).
 *
 * TEST INFRASTRUCTURE: a plain-C restatement of the reference's non-JIT fill
 * pipeline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker; the MI355X library never links it.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replays a scene (blend2d_amd/csrc/scene_format.h) and writes the canvas rows, packed,
 * to `out`.  Returns 0 or the first BLResult error a call produced. */
uint32_t orc_render(const uint8_t* scene, size_t size, uint8_t* out, size_t out_size);

/* Edge segments (x0, y0, x1, y1, sign) of a filled path, 24.8 fixed point,
 * for a user transform `m6` on a `width` x `height` canvas.  Returns the count. */
size_t orc_path_segments(const uint8_t* cmds, const double* vtx, size_t n, const double* m6,
                         uint32_t width, uint32_t height, int32_t* out, size_t cap);

/* Gradient LUT for stops given as straight-alpha rgba32 (add_stop order). */
void orc_gradient_lut(const double* offsets, const uint32_t* rgba32, size_t n, uint32_t* out, uint32_t* lut_size_out);

/* One pixel of an executed comp op (after simplification). */
uint32_t orc_composite_pixel(uint32_t op, uint32_t dst_format, uint32_t d, uint32_t s, uint32_t m);

/* Comp-op simplification: op | dst_fmt << 8 | src_fmt << 16 | solid_id << 24. */
uint32_t orc_simplify_op(uint32_t op, uint32_t dst_fmt, uint32_t src_fmt);

uint32_t orc_div255(uint32_t x);
void orc_segments_mask(const int32_t* segs, size_t n, uint32_t w, uint32_t h, uint32_t fill_rule, uint8_t* out);
uint32_t orc_premultiply(uint32_t argb32);

/* Rasterizer invariant of raster/analyticrasterizer_test.cpp:34-158: band-wise
 * rasterize() and advance_to_y() reach the same state.  Returns mismatches. */
size_t orc_raster_advance_check(const int32_t* edges, size_t n, uint32_t band_height, uint32_t w);

#ifdef __cplusplus
}
#endif

#endif /* ORC_H */
