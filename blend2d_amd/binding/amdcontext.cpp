// Reference-side binding of the MI355X fill backend: a BLContext
// implementation (BLContextImpl with its BLContextVirt table,
// core/context.h:454-626) whose fill calls forward to the C ABI of
// libblend2d_amd.so (include/blend2d_amd.h).
//
// A maintainer adds this file to the reference tree (e.g. as
// blend2d/raster/amdcontext.cpp, built when BL_BUILD_AMD is defined and
// linked against libblend2d_amd.so) and makes bl_context_init_as()
// (core/context.cpp) call bl_amd_context_init_impl() when the create-info
// asks for the GPU -- the same way it calls bl_raster_context_init_impl()
// for the CPU raster context.  It uses the reference's public API plus the
// object allocator of core/object_p.h; tests/test_binding.py compiles it
// against the reference headers.
//
// Supported here: every fill entry (fill_all / rect / path / geometry with
// the context style, an explicit rgba32 / rgba64 colour or an explicit style
// object), clear_*, the fill style slot (colours, BLGradient, BLPattern),
// fill alpha, global alpha, comp op, fill rule, transforms, the pattern and
// gradient quality hints, image blits (blit_image / blit_scaled_image) and mask
// fills (fill_mask in all its forms).  Every other entry (stroking, text,
// clipping, save / restore) reports BL_ERROR_NOT_IMPLEMENTED.

#include <blend2d/blend2d.h>
#include <blend2d/core/object_p.h>

#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "blend2d_amd.h"

namespace bl_amd_binding {

// A pattern image uploaded to the device, kept while the context lives.
struct DeviceImage {
  BLImage host;            // keeps the host image (and its impl address) alive
  BLAmdImage* device = nullptr;
};

struct AmdContextImpl : public BLContextImpl {
  BLAmdContext* gpu = nullptr;
  BLAmdImage* target = nullptr;
  BLImage host_target;     // keeps the target alive while the context lives
  BLImageData dst{};       // its pixels, made mutable at attach; written back at flush(SYNC) and at the end
  BLContextState state_data{};
  std::unordered_map<const void*, DeviceImage> images;   // by image impl
  uint32_t error_flags = 0;   // BLContextErrorFlags accumulated by flushes
};

AmdContextImpl* self_of(BLContextImpl* impl) { return static_cast<AmdContextImpl*>(impl); }

BLResult result_of(BLAmdResult r) { return r ? BLResult(r) : BL_SUCCESS; }

// Writes the device pixels back into the caller's image.  The pixel pointer
// was taken when the image was made mutable at attach, as the raster context
// does (rastercontext.cpp:4306-4342): the context's own reference must not
// make the image shared-and-copied at write-back time.
BLResult download(AmdContextImpl* self) {
  return result_of(bl_amd_image_download(self->target, self->dst.pixel_data, self->dst.stride));
}

BLAmdImage* device_image(AmdContextImpl* self, const BLImage& img) {
  const void* key = img._d.impl;
  auto it = self->images.find(key);
  if (it != self->images.end()) return it->second.device;
  BLImageData d;
  if (img.get_data(&d) != BL_SUCCESS) return nullptr;
  BLAmdImage* dev = nullptr;
  if (bl_amd_image_create(&dev, d.size.w, d.size.h, d.format) != BL_AMD_SUCCESS) return nullptr;
  if (bl_amd_image_upload(dev, d.pixel_data, d.stride) != BL_AMD_SUCCESS) {
    bl_amd_image_destroy(dev);
    return nullptr;
  }
  self->images.emplace(key, DeviceImage{img, dev});
  return dev;
}

// A style object as the C ABI's BLAmdStyle; `offs`, `cols` and `m` hold the
// arrays it points to.
BLResult style_of(AmdContextImpl* self, const BLObjectCore* style, BLAmdStyle& out, std::vector<double>& offs,
                  std::vector<uint32_t>& cols, BLMatrix2D& m) {
  out = BLAmdStyle{};
  // Dispatch on the object type: the is_rgba32() / is_gradient() family
  // tests type bits (a gradient, type 5, also passes is_rgba32(), type 1).
  switch (style->_d.get_type()) {
    case BL_OBJECT_TYPE_RGBA32:
      out.kind = BL_AMD_STYLE_RGBA32;
      out.rgba32 = style->_d.rgba32.value;
      return BL_SUCCESS;
    case BL_OBJECT_TYPE_RGBA64:
      out.kind = BL_AMD_STYLE_RGBA64;
      out.rgba64 = style->_d.rgba64.value;
      return BL_SUCCESS;
    case BL_OBJECT_TYPE_RGBA:
      // float colours become 8-bit ones, as the raster context resolves them
      // (rastercontext.cpp:1277-1310)
      out.kind = BL_AMD_STYLE_RGBA32;
      out.rgba32 = style->_d.rgba.to_rgba32().value;
      return BL_SUCCESS;
    case BL_OBJECT_TYPE_GRADIENT: {
      const BLGradient& g = *static_cast<const BLGradient*>(static_cast<const BLGradientCore*>(style));
      offs.resize(g.size());
      cols.resize(g.size());
      for (size_t i = 0; i < g.size(); i++) {
        offs[i] = g.stops()[i].offset;
        cols[i] = BLRgba32(g.stops()[i].rgba).value;   // the tables use the high byte of each channel
      }
      out.kind = BL_AMD_STYLE_GRADIENT;
      out.gradient_type = uint32_t(g.type());
      out.extend = uint32_t(g.extend_mode());
      out.values = &g.linear().x0;   // the value array of the union (x0 y0 x1 y1 r0 r1 / angle repeat)
      out.stop_offsets = offs.data();
      out.stop_rgba32 = cols.data();
      out.n_stops = g.size();
      if (g.has_transform()) {
        m = g.transform();
        out.transform = &m.m00;
      }
      return BL_SUCCESS;
    }
    case BL_OBJECT_TYPE_PATTERN: {
      const BLPattern& p = *static_cast<const BLPattern*>(static_cast<const BLPatternCore*>(style));
      BLAmdImage* img = device_image(self, p.get_image());
      if (!img) return BL_ERROR_INVALID_VALUE;
      out.kind = BL_AMD_STYLE_PATTERN;
      out.extend = uint32_t(p.extend_mode());
      out.image = img;
      if (p.has_transform()) {
        m = p.transform();
        out.transform = &m.m00;
      }
      return BL_SUCCESS;
    }
    default:
      break;
  }
  return BL_ERROR_INVALID_VALUE;
}

// Calls `fill` with the explicit style converted; rgba64 colours take the
// _rgba64 entry, everything else the _ext one.
template<typename Ext, typename Rgba64>
BLResult with_style(BLContextImpl* impl, const BLObjectCore* style, Ext&& ext, Rgba64&& rgba64) {
  AmdContextImpl* self = self_of(impl);
  BLAmdStyle st;
  std::vector<double> offs;
  std::vector<uint32_t> cols;
  BLMatrix2D m;
  BL_PROPAGATE(style_of(self, style, st, offs, cols, m));
  if (st.kind == BL_AMD_STYLE_RGBA64) return result_of(rgba64(self->gpu, st.rgba64));
  return result_of(ext(self->gpu, &st));
}

// ---- BLContextVirt entries ------------------------------------------------

BLResult BL_CDECL destroy_impl(BLObjectImpl* obj) noexcept {
  AmdContextImpl* self = static_cast<AmdContextImpl*>(static_cast<BLContextImpl*>(obj));
  if (self->gpu) {
    bl_amd_context_end(self->gpu);
    download(self);
    bl_amd_context_destroy(self->gpu);
  }
  if (self->target) bl_amd_image_destroy(self->target);
  for (auto& kv : self->images) bl_amd_image_destroy(kv.second.device);
  bl_stroke_options_destroy(&self->state_data.stroke_options);
  self->~AmdContextImpl();
  return bl::ObjectInternal::free_impl(obj);
}

// Fills whose geometry turns out invalid on the device (NaN reaching the line
// clipper) are reported by the flush that renders them through the
// accumulated error flags, as the raster context's asynchronous mode reports
// errors of deferred work (context.h:156-189); the fill call itself has
// already returned BL_SUCCESS.
BLResult BL_CDECL flush_impl(BLContextImpl* impl, BLContextFlushFlags flags) noexcept {
  AmdContextImpl* self = self_of(impl);
  BL_PROPAGATE(result_of(bl_amd_context_flush(self->gpu, uint32_t(flags))));
  uint32_t ef = 0;
  if (bl_amd_context_get_error_flags(self->gpu, &ef) == BL_AMD_SUCCESS) self->error_flags |= ef;
  return (flags & BL_CONTEXT_FLUSH_SYNC) ? download(self) : BL_SUCCESS;
}

bool property_is(const char* name, size_t size, const char* key) {
  const size_t n = strlen(key);
  return size == n && memcmp(name, key, n) == 0;
}

// The two properties the raster context answers (rastercontext.cpp:1497-1511).
BLResult BL_CDECL get_property_impl(const BLObjectImpl* impl, const char* name, size_t name_size, BLVarCore* value_out) noexcept {
  const AmdContextImpl* self = static_cast<const AmdContextImpl*>(static_cast<const BLContextImpl*>(impl));
  if (property_is(name, name_size, "thread_count")) return bl_var_assign_uint64(value_out, 0u);
  if (property_is(name, name_size, "accumulated_error_flags")) return bl_var_assign_uint64(value_out, self->error_flags);
  return BL_ERROR_INVALID_KEY;
}

BLResult BL_CDECL set_property_impl(BLObjectImpl*, const char*, size_t, const BLVarCore*) noexcept {
  return BL_ERROR_INVALID_KEY;
}

BLResult BL_CDECL apply_transform_op_impl(BLContextImpl* impl, BLTransformOp op, const void* data) noexcept {
  AmdContextImpl* self = self_of(impl);
  BLMatrix2D m = self->state_data.user_transform;
  BL_PROPAGATE(bl_matrix2d_apply_op(&m, op, data));
  BL_PROPAGATE(result_of(bl_amd_context_set_transform(self->gpu, &m.m00)));
  self->state_data.user_transform = m;
  self->state_data.final_transform = m;   // the meta transform stays identity
  return BL_SUCCESS;
}

BLResult BL_CDECL set_hint_impl(BLContextImpl* impl, BLContextHint hint, uint32_t value) noexcept {
  AmdContextImpl* self = self_of(impl);
  if (hint == BL_CONTEXT_HINT_PATTERN_QUALITY) BL_PROPAGATE(result_of(bl_amd_context_set_pattern_quality(self->gpu, value)));
  else if (hint == BL_CONTEXT_HINT_GRADIENT_QUALITY) BL_PROPAGATE(result_of(bl_amd_context_set_gradient_quality(self->gpu, value)));
  else if (uint32_t(hint) > BL_CONTEXT_HINT_MAX_VALUE) return BL_ERROR_INVALID_VALUE;
  self->state_data.hints.hints[hint] = uint8_t(value);
  return BL_SUCCESS;
}

BLResult BL_CDECL set_style_impl(BLContextImpl* impl, BLContextStyleSlot slot, const BLObjectCore* style,
                                 BLContextStyleTransformMode mode) noexcept {
  AmdContextImpl* self = self_of(impl);
  if (slot != BL_CONTEXT_STYLE_SLOT_FILL || mode != BL_CONTEXT_STYLE_TRANSFORM_MODE_USER) return BL_ERROR_NOT_IMPLEMENTED;
  BLAmdStyle st;
  std::vector<double> offs;
  std::vector<uint32_t> cols;
  BLMatrix2D m;
  BL_PROPAGATE(style_of(self, style, st, offs, cols, m));
  BLAmdResult r = BL_AMD_SUCCESS;
  switch (st.kind) {
    case BL_AMD_STYLE_RGBA32: r = bl_amd_context_set_fill_style_rgba32(self->gpu, st.rgba32); break;
    case BL_AMD_STYLE_RGBA64: r = bl_amd_context_set_fill_style_rgba64(self->gpu, st.rgba64); break;
    case BL_AMD_STYLE_GRADIENT:
      r = bl_amd_context_set_fill_style_gradient(self->gpu, st.gradient_type, st.values, st.extend, st.stop_offsets,
                                                 st.stop_rgba32, st.n_stops, st.transform);
      break;
    default: r = bl_amd_context_set_fill_style_pattern(self->gpu, st.image, st.extend, st.transform); break;
  }
  return result_of(r);
}

BLResult BL_CDECL set_style_rgba_impl(BLContextImpl* impl, BLContextStyleSlot slot, const BLRgba* rgba) noexcept {
  if (slot != BL_CONTEXT_STYLE_SLOT_FILL) return BL_ERROR_NOT_IMPLEMENTED;
  const float v[4] = {rgba->r, rgba->g, rgba->b, rgba->a};
  return result_of(bl_amd_context_set_fill_style_rgba(self_of(impl)->gpu, v));
}

BLResult BL_CDECL set_style_rgba32_impl(BLContextImpl* impl, BLContextStyleSlot slot, uint32_t rgba32) noexcept {
  if (slot != BL_CONTEXT_STYLE_SLOT_FILL) return BL_ERROR_NOT_IMPLEMENTED;
  return result_of(bl_amd_context_set_fill_style_rgba32(self_of(impl)->gpu, rgba32));
}

BLResult BL_CDECL set_style_rgba64_impl(BLContextImpl* impl, BLContextStyleSlot slot, uint64_t rgba64) noexcept {
  if (slot != BL_CONTEXT_STYLE_SLOT_FILL) return BL_ERROR_NOT_IMPLEMENTED;
  return result_of(bl_amd_context_set_fill_style_rgba64(self_of(impl)->gpu, rgba64));
}

BLResult BL_CDECL set_style_alpha_impl(BLContextImpl* impl, BLContextStyleSlot slot, double alpha) noexcept {
  if (slot != BL_CONTEXT_STYLE_SLOT_FILL) return BL_ERROR_NOT_IMPLEMENTED;
  BL_PROPAGATE(result_of(bl_amd_context_set_fill_alpha(self_of(impl)->gpu, alpha)));
  self_of(impl)->state_data.style_alpha[BL_CONTEXT_STYLE_SLOT_FILL] = alpha;
  return BL_SUCCESS;
}

BLResult BL_CDECL set_global_alpha_impl(BLContextImpl* impl, double alpha) noexcept {
  BL_PROPAGATE(result_of(bl_amd_context_set_global_alpha(self_of(impl)->gpu, alpha)));
  self_of(impl)->state_data.global_alpha = alpha;
  return BL_SUCCESS;
}

BLResult BL_CDECL set_comp_op_impl(BLContextImpl* impl, BLCompOp op) noexcept {
  BL_PROPAGATE(result_of(bl_amd_context_set_comp_op(self_of(impl)->gpu, uint32_t(op))));
  self_of(impl)->state_data.comp_op = uint8_t(op);
  return BL_SUCCESS;
}

BLResult BL_CDECL set_fill_rule_impl(BLContextImpl* impl, BLFillRule rule) noexcept {
  BL_PROPAGATE(result_of(bl_amd_context_set_fill_rule(self_of(impl)->gpu, uint32_t(rule))));
  self_of(impl)->state_data.fill_rule = uint8_t(rule);
  return BL_SUCCESS;
}

// Blits: the image (uploaded once per context) and its area as x, y, w, h.
template<typename Fn>
BLResult with_image(BLContextImpl* impl, const BLImageCore* img, const BLRectI* area, Fn&& fn) {
  AmdContextImpl* self = self_of(impl);
  BLAmdImage* dev = device_image(self, *static_cast<const BLImage*>(img));
  if (!dev) return BL_ERROR_INVALID_VALUE;
  int32_t a[4];
  if (area) { a[0] = area->x; a[1] = area->y; a[2] = area->w; a[3] = area->h; }
  return result_of(fn(self->gpu, dev, area ? a : nullptr));
}

BLResult BL_CDECL blit_image_i_impl(BLContextImpl* impl, const BLPointI* o, const BLImageCore* img, const BLRectI* area) noexcept {
  return with_image(impl, img, area, [o](BLAmdContext* g, BLAmdImage* d, const int32_t* a) {
    return bl_amd_context_blit_image_i(g, o->x, o->y, d, a);
  });
}
BLResult BL_CDECL blit_image_d_impl(BLContextImpl* impl, const BLPoint* o, const BLImageCore* img, const BLRectI* area) noexcept {
  return with_image(impl, img, area, [o](BLAmdContext* g, BLAmdImage* d, const int32_t* a) {
    return bl_amd_context_blit_image_d(g, o->x, o->y, d, a);
  });
}
BLResult BL_CDECL blit_scaled_image_i_impl(BLContextImpl* impl, const BLRectI* r, const BLImageCore* img, const BLRectI* area) noexcept {
  return with_image(impl, img, area, [r](BLAmdContext* g, BLAmdImage* d, const int32_t* a) {
    return bl_amd_context_blit_scaled_image_i(g, r->x, r->y, r->w, r->h, d, a);
  });
}
BLResult BL_CDECL blit_scaled_image_d_impl(BLContextImpl* impl, const BLRect* r, const BLImageCore* img, const BLRectI* area) noexcept {
  return with_image(impl, img, area, [r](BLAmdContext* g, BLAmdImage* d, const int32_t* a) {
    return bl_amd_context_blit_scaled_image_d(g, r->x, r->y, r->w, r->h, d, a);
  });
}

// Mask fills: the mask image (uploaded once per context) and its area; an
// explicit style is converted as for the other fills.

BLResult BL_CDECL fill_mask_i_impl(BLContextImpl* impl, const BLPointI* o, const BLImageCore* mask, const BLRectI* area) noexcept {
  return with_image(impl, mask, area, [o](BLAmdContext* g, BLAmdImage* m, const int32_t* a) {
    return bl_amd_context_fill_mask_i(g, o->x, o->y, m, a);
  });
}
BLResult BL_CDECL fill_mask_i_rgba32_impl(BLContextImpl* impl, const BLPointI* o, const BLImageCore* mask, const BLRectI* area,
                                          uint32_t c) noexcept {
  return with_image(impl, mask, area, [o, c](BLAmdContext* g, BLAmdImage* m, const int32_t* a) {
    return bl_amd_context_fill_mask_i_rgba32(g, o->x, o->y, m, a, c);
  });
}
BLResult BL_CDECL fill_mask_i_ext_impl(BLContextImpl* impl, const BLPointI* o, const BLImageCore* mask, const BLRectI* area,
                                       const BLObjectCore* style) noexcept {
  return with_image(impl, mask, area, [impl, o, style](BLAmdContext*, BLAmdImage* m, const int32_t* a) {
    return with_style(impl, style,
                      [o, m, a](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_mask_i_ext(g, o->x, o->y, m, a, s); },
                      [o, m, a](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_mask_i_rgba64(g, o->x, o->y, m, a, c); });
  });
}
BLResult BL_CDECL fill_mask_d_impl(BLContextImpl* impl, const BLPoint* o, const BLImageCore* mask, const BLRectI* area) noexcept {
  return with_image(impl, mask, area, [o](BLAmdContext* g, BLAmdImage* m, const int32_t* a) {
    return bl_amd_context_fill_mask_d(g, o->x, o->y, m, a);
  });
}
BLResult BL_CDECL fill_mask_d_rgba32_impl(BLContextImpl* impl, const BLPoint* o, const BLImageCore* mask, const BLRectI* area,
                                          uint32_t c) noexcept {
  return with_image(impl, mask, area, [o, c](BLAmdContext* g, BLAmdImage* m, const int32_t* a) {
    return bl_amd_context_fill_mask_d_rgba32(g, o->x, o->y, m, a, c);
  });
}
BLResult BL_CDECL fill_mask_d_ext_impl(BLContextImpl* impl, const BLPoint* o, const BLImageCore* mask, const BLRectI* area,
                                       const BLObjectCore* style) noexcept {
  return with_image(impl, mask, area, [impl, o, style](BLAmdContext*, BLAmdImage* m, const int32_t* a) {
    return with_style(impl, style,
                      [o, m, a](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_mask_d_ext(g, o->x, o->y, m, a, s); },
                      [o, m, a](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_mask_d_rgba64(g, o->x, o->y, m, a, c); });
  });
}

BLResult BL_CDECL clear_all_impl(BLContextImpl* impl) noexcept { return result_of(bl_amd_context_clear_all(self_of(impl)->gpu)); }
BLResult BL_CDECL clear_recti_impl(BLContextImpl* impl, const BLRectI* r) noexcept {
  return result_of(bl_amd_context_clear_rect_i(self_of(impl)->gpu, r->x, r->y, r->w, r->h));
}
BLResult BL_CDECL clear_rectd_impl(BLContextImpl* impl, const BLRect* r) noexcept {
  return result_of(bl_amd_context_clear_rect_d(self_of(impl)->gpu, r->x, r->y, r->w, r->h));
}

BLResult BL_CDECL fill_all_impl(BLContextImpl* impl) noexcept { return result_of(bl_amd_context_fill_all(self_of(impl)->gpu)); }
BLResult BL_CDECL fill_all_rgba32_impl(BLContextImpl* impl, uint32_t c) noexcept {
  return result_of(bl_amd_context_fill_all_rgba32(self_of(impl)->gpu, c));
}
BLResult BL_CDECL fill_all_ext_impl(BLContextImpl* impl, const BLObjectCore* style) noexcept {
  return with_style(impl, style, [](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_all_ext(g, s); },
                    [](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_all_rgba64(g, c); });
}

BLResult BL_CDECL fill_rect_i_impl(BLContextImpl* impl, const BLRectI* r) noexcept {
  return result_of(bl_amd_context_fill_rect_i(self_of(impl)->gpu, r->x, r->y, r->w, r->h));
}
BLResult BL_CDECL fill_rect_i_rgba32_impl(BLContextImpl* impl, const BLRectI* r, uint32_t c) noexcept {
  return result_of(bl_amd_context_fill_rect_i_rgba32(self_of(impl)->gpu, r->x, r->y, r->w, r->h, c));
}
BLResult BL_CDECL fill_rect_i_ext_impl(BLContextImpl* impl, const BLRectI* r, const BLObjectCore* style) noexcept {
  return with_style(impl, style,
                    [r](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_rect_i_ext(g, r->x, r->y, r->w, r->h, s); },
                    [r](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_rect_i_rgba64(g, r->x, r->y, r->w, r->h, c); });
}

BLResult BL_CDECL fill_rect_d_impl(BLContextImpl* impl, const BLRect* r) noexcept {
  return result_of(bl_amd_context_fill_rect_d(self_of(impl)->gpu, r->x, r->y, r->w, r->h));
}
BLResult BL_CDECL fill_rect_d_rgba32_impl(BLContextImpl* impl, const BLRect* r, uint32_t c) noexcept {
  return result_of(bl_amd_context_fill_rect_d_rgba32(self_of(impl)->gpu, r->x, r->y, r->w, r->h, c));
}
BLResult BL_CDECL fill_rect_d_ext_impl(BLContextImpl* impl, const BLRect* r, const BLObjectCore* style) noexcept {
  return with_style(impl, style,
                    [r](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_rect_d_ext(g, r->x, r->y, r->w, r->h, s); },
                    [r](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_rect_d_rgba64(g, r->x, r->y, r->w, r->h, c); });
}

// BLPath data: command bytes and vertices of the public BLPath view.
struct PathView {
  const uint8_t* cmds;
  const double* vtx;
  size_t n;
  explicit PathView(const BLPathCore* core) {
    const BLPath& p = *static_cast<const BLPath*>(core);
    cmds = p.command_data();
    vtx = &p.vertex_data()->x;
    n = p.size();
  }
};

BLResult BL_CDECL fill_path_d_impl(BLContextImpl* impl, const BLPoint* origin, const BLPathCore* path) noexcept {
  const PathView v(path);
  return result_of(bl_amd_context_fill_path_d(self_of(impl)->gpu, origin ? &origin->x : nullptr, v.cmds, v.vtx, v.n));
}
BLResult BL_CDECL fill_path_d_rgba32_impl(BLContextImpl* impl, const BLPoint* origin, const BLPathCore* path, uint32_t c) noexcept {
  const PathView v(path);
  return result_of(bl_amd_context_fill_path_d_rgba32(self_of(impl)->gpu, origin ? &origin->x : nullptr, v.cmds, v.vtx, v.n, c));
}
BLResult BL_CDECL fill_path_d_ext_impl(BLContextImpl* impl, const BLPoint* origin, const BLPathCore* path,
                                       const BLObjectCore* style) noexcept {
  const PathView v(path);
  const double* o = origin ? &origin->x : nullptr;
  return with_style(impl, style,
                    [&](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_path_d_ext(g, o, v.cmds, v.vtx, v.n, s); },
                    [&](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_path_d_rgba64(g, o, v.cmds, v.vtx, v.n, c); });
}

// Geometry data has the reference's own layout, which the C ABI takes as is;
// paths go through the path entry.
BLResult BL_CDECL fill_geometry_impl(BLContextImpl* impl, BLGeometryType type, const void* data) noexcept {
  if (type == BL_GEOMETRY_TYPE_PATH) return fill_path_d_impl(impl, nullptr, static_cast<const BLPathCore*>(data));
  return result_of(bl_amd_context_fill_geometry(self_of(impl)->gpu, uint32_t(type), data));
}
BLResult BL_CDECL fill_geometry_rgba32_impl(BLContextImpl* impl, BLGeometryType type, const void* data, uint32_t c) noexcept {
  if (type == BL_GEOMETRY_TYPE_PATH) return fill_path_d_rgba32_impl(impl, nullptr, static_cast<const BLPathCore*>(data), c);
  return result_of(bl_amd_context_fill_geometry_rgba32(self_of(impl)->gpu, uint32_t(type), data, c));
}
BLResult BL_CDECL fill_geometry_ext_impl(BLContextImpl* impl, BLGeometryType type, const void* data,
                                         const BLObjectCore* style) noexcept {
  if (type == BL_GEOMETRY_TYPE_PATH) return fill_path_d_ext_impl(impl, nullptr, static_cast<const BLPathCore*>(data), style);
  return with_style(impl, style,
                    [&](BLAmdContext* g, const BLAmdStyle* s) { return bl_amd_context_fill_geometry_ext(g, uint32_t(type), data, s); },
                    [&](BLAmdContext* g, uint64_t c) { return bl_amd_context_fill_geometry_rgba64(g, uint32_t(type), data, c); });
}

// Every other entry: not implemented by the GPU backend.
template<typename Fn> struct NotImplemented;
template<typename... Args> struct NotImplemented<BLResult (*)(Args...) noexcept> {
  static BLResult BL_CDECL fn(Args...) noexcept { return bl_make_error(BL_ERROR_NOT_IMPLEMENTED); }
};
template<typename Fn> void not_implemented(Fn& slot) { slot = &NotImplemented<Fn>::fn; }

BLContextVirt make_virt() {
  BLContextVirt v{};
  v.base.destroy = destroy_impl;
  v.base.get_property = get_property_impl;
  v.base.set_property = set_property_impl;

  v.apply_transform_op = apply_transform_op_impl;
  v.fill_rect_i = fill_rect_i_impl;
  v.fill_rect_i_rgba32 = fill_rect_i_rgba32_impl;
  v.fill_rect_i_ext = fill_rect_i_ext_impl;
  v.fill_rect_d = fill_rect_d_impl;
  v.fill_rect_d_rgba32 = fill_rect_d_rgba32_impl;
  v.fill_rect_d_ext = fill_rect_d_ext_impl;
  v.fill_path_d = fill_path_d_impl;
  v.fill_path_d_rgba32 = fill_path_d_rgba32_impl;
  v.fill_path_d_ext = fill_path_d_ext_impl;
  v.blit_image_i = blit_image_i_impl;
  v.blit_scaled_image_i = blit_scaled_image_i_impl;

  v.flush = flush_impl;
  not_implemented(v.save);
  not_implemented(v.restore);
  not_implemented(v.user_to_meta);
  v.set_hint = set_hint_impl;
  not_implemented(v.set_hints);
  not_implemented(v.set_flatten_mode);
  not_implemented(v.set_flatten_tolerance);
  not_implemented(v.set_approximation_options);
  not_implemented(v.get_style);
  v.set_style = set_style_impl;
  v.set_style_rgba = set_style_rgba_impl;
  v.set_style_rgba32 = set_style_rgba32_impl;
  v.set_style_rgba64 = set_style_rgba64_impl;
  not_implemented(v.disable_style);
  v.set_style_alpha = set_style_alpha_impl;
  not_implemented(v.swap_styles);
  v.set_global_alpha = set_global_alpha_impl;
  v.set_comp_op = set_comp_op_impl;
  v.set_fill_rule = set_fill_rule_impl;
  not_implemented(v.set_stroke_width);
  not_implemented(v.set_stroke_miter_limit);
  not_implemented(v.set_stroke_cap);
  not_implemented(v.set_stroke_caps);
  not_implemented(v.set_stroke_join);
  not_implemented(v.set_stroke_dash_offset);
  not_implemented(v.set_stroke_dash_array);
  not_implemented(v.set_stroke_transform_order);
  not_implemented(v.set_stroke_options);
  not_implemented(v.clip_to_rect_i);
  not_implemented(v.clip_to_rect_d);
  not_implemented(v.restore_clipping);
  v.clear_all = clear_all_impl;
  v.clear_recti = clear_recti_impl;
  v.clear_rectd = clear_rectd_impl;
  v.fill_all = fill_all_impl;
  v.fill_all_rgba32 = fill_all_rgba32_impl;
  v.fill_all_ext = fill_all_ext_impl;
  v.fill_geometry = fill_geometry_impl;
  v.fill_geometry_rgba32 = fill_geometry_rgba32_impl;
  v.fill_geometry_ext = fill_geometry_ext_impl;
  not_implemented(v.fill_text_op_i);
  not_implemented(v.fill_text_op_i_rgba32);
  not_implemented(v.fill_text_op_i_ext);
  not_implemented(v.fill_text_op_d);
  not_implemented(v.fill_text_op_d_rgba32);
  not_implemented(v.fill_text_op_d_ext);
  v.fill_mask_i = fill_mask_i_impl;
  v.fill_mask_i_rgba32 = fill_mask_i_rgba32_impl;
  v.fill_mask_i_ext = fill_mask_i_ext_impl;
  v.fill_mask_d = fill_mask_d_impl;
  v.fill_mask_d_Rgba32 = fill_mask_d_rgba32_impl;
  v.fill_mask_d_ext = fill_mask_d_ext_impl;
  not_implemented(v.stroke_path_d);
  not_implemented(v.stroke_path_d_rgba32);
  not_implemented(v.stroke_path_d_ext);
  not_implemented(v.stroke_geometry);
  not_implemented(v.stroke_geometry_rgba32);
  not_implemented(v.stroke_geometry_ext);
  not_implemented(v.stroke_text_op_i);
  not_implemented(v.stroke_text_op_i_rgba32);
  not_implemented(v.stroke_text_op_i_ext);
  not_implemented(v.stroke_text_op_d);
  not_implemented(v.stroke_text_op_d_rgba32);
  not_implemented(v.stroke_text_op_d_ext);
  v.blit_image_d = blit_image_d_impl;
  v.blit_scaled_image_d = blit_scaled_image_d_impl;
  return v;
}

const BLContextVirt amd_virt = make_virt();

}  // namespace bl_amd_binding

// Makes `self` a GPU context rendering into `image` (the counterpart of
// bl_raster_context_init_impl, rastercontext.cpp:4619-4638).  The image's
// pixels are uploaded; they are written back at flush(SYNC) and at the end.
BLResult bl_amd_context_init_impl(BLContextCore* self, BLImageCore* image, const BLContextCreateInfo* cci) noexcept {
  using namespace bl_amd_binding;
  (void)cci;
  BLImageData d;
  BL_PROPAGATE(bl_image_make_mutable(image, &d));
  if (d.format != BL_FORMAT_PRGB32 && d.format != BL_FORMAT_XRGB32 && d.format != BL_FORMAT_A8)
    return bl_make_error(BL_ERROR_NOT_IMPLEMENTED);

  BLObjectInfo info = BLObjectInfo::from_type_with_marker(BL_OBJECT_TYPE_CONTEXT);
  BL_PROPAGATE(bl::ObjectInternal::alloc_impl_t<AmdContextImpl>(self, info));
  AmdContextImpl* impl = new (static_cast<void*>(self->_d.impl)) AmdContextImpl();
  impl->virt = &amd_virt;
  impl->state = &impl->state_data;
  impl->context_type = BL_CONTEXT_TYPE_RASTER;   // renders into an image; the reference has no GPU type

  BLContextState& st = impl->state_data;
  impl->host_target = *static_cast<BLImage*>(image);
  impl->dst = d;
  st.target_image = &impl->host_target;
  st.target_size.reset(double(d.size.w), double(d.size.h));
  st.comp_op = uint8_t(BL_COMP_OP_SRC_OVER);
  st.fill_rule = uint8_t(BL_FILL_RULE_NON_ZERO);
  st.global_alpha = 1.0;
  st.style_alpha[0] = st.style_alpha[1] = 1.0;
  bl_stroke_options_init(&st.stroke_options);
  st.meta_transform.reset();
  st.user_transform.reset();
  st.final_transform.reset();

  BLResult r = result_of(bl_amd_context_create(&impl->gpu));
  if (r == BL_SUCCESS) r = result_of(bl_amd_image_create(&impl->target, d.size.w, d.size.h, d.format));
  if (r == BL_SUCCESS) r = result_of(bl_amd_image_upload(impl->target, d.pixel_data, d.stride));
  if (r == BL_SUCCESS) r = result_of(bl_amd_context_begin(impl->gpu, impl->target));
  if (r != BL_SUCCESS) destroy_impl(impl);
  return r;
}
