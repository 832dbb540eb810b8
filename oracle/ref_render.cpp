// Scene driver for the reference Blend2D library built by oracle/ref.mk.
//
// Test infrastructure only: it replays a scene (blend2d_amd/csrc/scene_format.h) through
// the reference's public C++ API (BLContext, BLPath, BLGradient, BLPattern)
// so the tests can compare the CPU oracle and the MI355X backend against the
// real reference, and bench.py can time the reference CPU pipeline.  It is
// never linked into the product library.

#include <blend2d/blend2d.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "scene_format.h"

namespace {

struct Reader {
  const uint8_t* p;
  const uint8_t* end;
  template<typename T> bool get(T& out) {
    if (size_t(end - p) < sizeof(T)) return false;
    std::memcpy(&out, p, sizeof(T));
    p += sizeof(T);
    return true;
  }
};

static inline size_t pad8(size_t n) { return (n + 7u) & ~size_t(7); }

static BLMatrix2D matrix_from(const double m[6]) {
  return BLMatrix2D(m[0], m[1], m[2], m[3], m[4], m[5]);
}

// A decoded call; objects are built before the timed region starts.
struct Call {
  uint32_t op = 0;
  uint32_t u0 = 0;
  double d[6] = {};
  BLPath path;
  std::vector<BLPoint> poly;
  BLGradient gradient;
  BLPattern pattern;
  uint64_t u64 = 0;
  float f[4] = {};
  int32_t i[4] = {};
  uint32_t rgba32 = 0;
  uint32_t flags = 0;
  uint32_t gtype = 0;
  std::vector<uint8_t> blob;   // geometry data
  // FILL_EXPLICIT
  SceneExplicit ex{};
  BLRectI rect_i;
  BLRect rect_d;
  // BLIT_IMAGE
  SceneBlit blit{};
  // FILL_MASK, FILL_EXPLICIT target 5
  SceneMask mask{};
};

// Builds the style object of a FILL_EXPLICIT payload (kinds 2, 3) into `call`;
// returns the bytes it used or 0 when malformed.
size_t decode_style(Call& call, const uint8_t* p, const std::vector<BLImage>& images) {
  if (call.ex.style_kind == 2) {
    SceneGradient g;
    std::memcpy(&g, p, sizeof(g));
    call.gradient = (
      g.type == 0 ? BLGradient(BLLinearGradientValues(g.values[0], g.values[1], g.values[2], g.values[3])) :
      g.type == 1 ? BLGradient(BLRadialGradientValues(g.values[0], g.values[1], g.values[2], g.values[3], g.values[4], g.values[5])) :
                    BLGradient(BLConicGradientValues(g.values[0], g.values[1], g.values[2], g.values[3])));
    call.gradient.set_extend_mode(BLExtendMode(g.extend));
    for (uint32_t k = 0; k < g.n_stops; k++) {
      SceneStop st;
      std::memcpy(&st, p + sizeof(g) + k * sizeof(SceneStop), sizeof(st));
      call.gradient.add_stop(st.offset, BLRgba32(st.rgba32));
    }
    if (g.has_transform) call.gradient.set_transform(matrix_from(g.transform));
    return sizeof(g) + g.n_stops * sizeof(SceneStop);
  }
  if (call.ex.style_kind == 3) {
    ScenePattern pt;
    std::memcpy(&pt, p, sizeof(pt));
    if (pt.image_index >= images.size()) return 0;
    call.pattern = BLPattern(images[pt.image_index], BLExtendMode(pt.extend));
    if (pt.has_transform) call.pattern.set_transform(matrix_from(pt.transform));
    return sizeof(pt);
  }
  return 0;
}

}  // namespace

// With REF_RENDER_AMD the same driver renders through the MI355X backend's
// BLContext implementation (blend2d_amd/binding/amdcontext.cpp): the calls
// go through the reference's own C++ API and its BLContextVirt table to the
// GPU, as they would for a user of the reference with the backend bound in.
#if defined(REF_RENDER_AMD)
BLResult bl_amd_context_init_impl(BLContextCore* self, BLImageCore* image, const BLContextCreateInfo* cci) noexcept;
#define REF_RENDER_ENTRY ref_render_amd
#else
#define REF_RENDER_ENTRY ref_render
#endif

extern "C" {

// Renders `scene` and writes the canvas rows (width * bytes-per-pixel each,
// no padding) to `out`.  Returns the first non-zero BLResult reported by any
// call (rendering continues after an error, as a caller ignoring errors
// would), or 0x7FFFFFFF when the scene cannot be parsed.  `elapsed_ms`
// receives the wall time from BLContext::begin() to BLContext::end().
uint32_t REF_RENDER_ENTRY(const uint8_t* scene, size_t size, uint8_t* out, size_t out_size,
                    uint32_t thread_count, double* elapsed_ms) {
  Reader r{scene, scene + size};
  SceneHeader hdr;
  if (!r.get(hdr) || hdr.magic != SCENE_MAGIC || hdr.version != SCENE_VERSION)
    return 0x7FFFFFFFu;

  const uint32_t bpp = hdr.format == SCENE_FMT_A8 ? 1u : 4u;
  if (out_size < size_t(hdr.width) * hdr.height * bpp)
    return 0x7FFFFFFFu;

  std::vector<BLImage> images;
  for (uint32_t i = 0; i < hdr.n_images; i++) {
    SceneImage si;
    if (!r.get(si)) return 0x7FFFFFFFu;
    BLImage img(int(si.width), int(si.height), BLFormat(si.format));
    BLImageData d;
    img.make_mutable(&d);
    uint32_t ibpp = si.format == SCENE_FMT_A8 ? 1u : 4u;
    for (uint32_t y = 0; y < si.height; y++)
      std::memcpy(static_cast<uint8_t*>(d.pixel_data) + intptr_t(y) * d.stride, r.p + size_t(y) * si.stride, size_t(si.width) * ibpp);
    r.p += pad8(size_t(si.height) * si.stride);
    images.push_back(img);
  }

  std::vector<Call> calls(hdr.n_cmds);
  for (uint32_t i = 0; i < hdr.n_cmds; i++) {
    SceneCmd c;
    if (!r.get(c) || size_t(r.end - r.p) < c.size) return 0x7FFFFFFFu;
    const uint8_t* payload = r.p;
    r.p += c.size;
    Call& call = calls[i];
    call.op = c.op;

    switch (c.op) {
      case SCENE_OP_SET_COMP_OP:
      case SCENE_OP_SET_FILL_RULE:
      case SCENE_OP_SET_STYLE_SOLID:
      case SCENE_OP_SET_PATTERN_QUALITY:
      case SCENE_OP_SET_GRADIENT_QUALITY:
        std::memcpy(&call.u0, payload, 4);
        break;
      case SCENE_OP_SET_TRANSFORM:
        std::memcpy(call.d, payload, 48);
        break;
      case SCENE_OP_SET_GLOBAL_ALPHA:
      case SCENE_OP_SET_FILL_ALPHA:
        std::memcpy(call.d, payload, 8);
        break;
      case SCENE_OP_SET_STYLE_RGBA64:
        std::memcpy(&call.u64, payload, 8);
        break;
      case SCENE_OP_SET_STYLE_RGBA:
        std::memcpy(call.f, payload, 16);
        break;
      case SCENE_OP_CLEAR_RECT_I:
        std::memcpy(call.i, payload, 16);
        break;
      case SCENE_OP_CLEAR_RECT_D:
        std::memcpy(call.d, payload, 32);
        break;
      case SCENE_OP_FILL_ALL_RGBA32:
        std::memcpy(&call.rgba32, payload, 4);
        break;
      case SCENE_OP_FILL_RECT_I_RGBA32:
        std::memcpy(call.i, payload, 16);
        std::memcpy(&call.rgba32, payload + 16, 4);
        break;
      case SCENE_OP_FILL_RECT_D_RGBA32:
        std::memcpy(call.d, payload, 32);
        std::memcpy(&call.rgba32, payload + 32, 4);
        break;
      case SCENE_OP_FILL_GEOMETRY: {
        SceneGeometry g;
        std::memcpy(&g, payload, sizeof(g));
        call.gtype = g.type;
        call.flags = g.has_rgba32;
        call.rgba32 = g.rgba32;
        call.u0 = g.count;
        call.blob.assign(payload + sizeof(g), payload + c.size);
        break;
      }
      case SCENE_OP_FILL_PATH_EX: {
        ScenePathEx h;
        std::memcpy(&h, payload, sizeof(h));
        call.flags = h.flags;
        call.rgba32 = h.rgba32;
        call.d[0] = h.origin[0];
        call.d[1] = h.origin[1];
        const uint8_t* cmds = payload + sizeof(h);
        const uint8_t* vtx = cmds + pad8(h.n);
        uint8_t* cmd_out;
        BLPoint* vtx_out;
        call.path.modify_op(BL_MODIFY_OP_ASSIGN_FIT, h.n, &cmd_out, &vtx_out);
        std::memcpy(cmd_out, cmds, h.n);
        std::memcpy(vtx_out, vtx, size_t(h.n) * sizeof(BLPoint));
        break;
      }
      case SCENE_OP_FILL_EXPLICIT: {
        std::memcpy(&call.ex, payload, sizeof(call.ex));
        const uint8_t* p = payload + sizeof(call.ex);
        if (call.ex.style_kind >= 2) {
          const size_t used = decode_style(call, p, images);
          if (!used) return 0x7FFFFFFFu;
          p += used;
        }
        switch (call.ex.target) {
          case 1: { int32_t v[4]; std::memcpy(v, p, 16); call.rect_i = BLRectI(v[0], v[1], v[2], v[3]); break; }
          case 2: { double v[4]; std::memcpy(v, p, 32); call.rect_d = BLRect(v[0], v[1], v[2], v[3]); break; }
          case 3: {
            const uint32_t n = call.ex.count;
            uint8_t* cmd_out;
            BLPoint* vtx_out;
            call.path.modify_op(BL_MODIFY_OP_ASSIGN_FIT, n, &cmd_out, &vtx_out);
            std::memcpy(cmd_out, p, n);
            std::memcpy(vtx_out, p + pad8(n), size_t(n) * sizeof(BLPoint));
            break;
          }
          case 4: call.blob.assign(p, payload + c.size); break;
          case 5:
            std::memcpy(&call.mask, p, sizeof(SceneMask));
            if (call.mask.image_index >= images.size()) return 0x7FFFFFFFu;
            break;
          default: break;
        }
        break;
      }
      case SCENE_OP_FILL_MASK:
        std::memcpy(&call.mask, payload, sizeof(SceneMask));
        if (call.mask.image_index >= images.size()) return 0x7FFFFFFFu;
        break;
      case SCENE_OP_FILL_RECT_D:
        std::memcpy(call.d, payload, 32);
        break;
      case SCENE_OP_FILL_RECT_I: {
        int32_t v[4];
        std::memcpy(v, payload, 16);
        for (int k = 0; k < 4; k++) call.d[k] = double(v[k]);
        break;
      }
      case SCENE_OP_FILL_PATH: {
        uint32_t n;
        std::memcpy(&n, payload, 4);
        const uint8_t* cmds = payload + 8;
        const uint8_t* vtx = cmds + pad8(n);
        uint8_t* cmd_out;
        BLPoint* vtx_out;
        call.path.modify_op(BL_MODIFY_OP_ASSIGN_FIT, n, &cmd_out, &vtx_out);
        std::memcpy(cmd_out, cmds, n);
        std::memcpy(vtx_out, vtx, size_t(n) * sizeof(BLPoint));
        break;
      }
      case SCENE_OP_FILL_POLYGON: {
        uint32_t n;
        std::memcpy(&n, payload, 4);
        call.poly.resize(n);
        std::memcpy(call.poly.data(), payload + 8, size_t(n) * sizeof(BLPoint));
        break;
      }
      case SCENE_OP_SET_STYLE_GRADIENT: {
        SceneGradient g;
        std::memcpy(&g, payload, sizeof(g));
        call.gradient = (
          g.type == 0 ? BLGradient(BLLinearGradientValues(g.values[0], g.values[1], g.values[2], g.values[3])) :
          g.type == 1 ? BLGradient(BLRadialGradientValues(g.values[0], g.values[1], g.values[2], g.values[3], g.values[4], g.values[5])) :
                        BLGradient(BLConicGradientValues(g.values[0], g.values[1], g.values[2], g.values[3])));
        call.gradient.set_extend_mode(BLExtendMode(g.extend));
        const uint8_t* sp = payload + sizeof(g);
        for (uint32_t k = 0; k < g.n_stops; k++) {
          SceneStop s;
          std::memcpy(&s, sp + k * sizeof(SceneStop), sizeof(s));
          call.gradient.add_stop(s.offset, BLRgba32(s.rgba32));
        }
        if (g.has_transform)
          call.gradient.set_transform(matrix_from(g.transform));
        break;
      }
      case SCENE_OP_BLIT_IMAGE:
        std::memcpy(&call.blit, payload, sizeof(SceneBlit));
        if (call.blit.image_index >= images.size()) return 0x7FFFFFFFu;
        break;
      case SCENE_OP_SET_STYLE_PATTERN: {
        ScenePattern p;
        std::memcpy(&p, payload, sizeof(p));
        if (p.image_index >= images.size()) return 0x7FFFFFFFu;
        call.pattern = BLPattern(images[p.image_index], BLExtendMode(p.extend));
        if (p.has_transform)
          call.pattern.set_transform(matrix_from(p.transform));
        break;
      }
      default:
        break;
    }
  }

  BLImage canvas(int(hdr.width), int(hdr.height), BLFormat(hdr.format));
  {
    BLImageData d;
    canvas.make_mutable(&d);
    for (uint32_t y = 0; y < hdr.height; y++) {
      uint8_t* row = static_cast<uint8_t*>(d.pixel_data) + intptr_t(y) * d.stride;
      if (bpp == 1) {
        std::memset(row, int(hdr.clear_value & 0xFFu), hdr.width);
      }
      else {
        uint32_t* row32 = reinterpret_cast<uint32_t*>(row);
        for (uint32_t x = 0; x < hdr.width; x++) row32[x] = hdr.clear_value;
      }
    }
  }

  BLContextCreateInfo ci {};
  ci.thread_count = thread_count;
  ci.flags = BL_CONTEXT_CREATE_FLAG_DISABLE_JIT;

  uint32_t first_error = 0;
  auto note = [&](BLResult res) { if (res != BL_SUCCESS && first_error == 0) first_error = res; };

  auto t0 = std::chrono::steady_clock::now();
  BLContext ctx;
#if defined(REF_RENDER_AMD)
  note(bl_amd_context_init_impl(&ctx, &canvas, &ci));
#else
  note(ctx.begin(canvas, ci));
#endif

  for (Call& call : calls) {
    switch (call.op) {
      case SCENE_OP_SET_COMP_OP: note(ctx.set_comp_op(BLCompOp(call.u0))); break;
      case SCENE_OP_SET_FILL_RULE: note(ctx.set_fill_rule(BLFillRule(call.u0))); break;
      case SCENE_OP_SET_TRANSFORM: note(ctx.set_transform(matrix_from(call.d))); break;
      case SCENE_OP_SET_GLOBAL_ALPHA: note(ctx.set_global_alpha(call.d[0])); break;
      case SCENE_OP_SET_STYLE_SOLID: note(ctx.set_fill_style(BLRgba32(call.u0))); break;
      case SCENE_OP_SET_STYLE_GRADIENT: note(ctx.set_fill_style(call.gradient)); break;
      case SCENE_OP_SET_STYLE_PATTERN: note(ctx.set_fill_style(call.pattern)); break;
      case SCENE_OP_SET_PATTERN_QUALITY: note(ctx.set_hint(BL_CONTEXT_HINT_PATTERN_QUALITY, call.u0)); break;
      case SCENE_OP_SET_GRADIENT_QUALITY: note(ctx.set_hint(BL_CONTEXT_HINT_GRADIENT_QUALITY, call.u0)); break;
      case SCENE_OP_FILL_PATH: note(ctx.fill_path(call.path)); break;
      case SCENE_OP_FILL_POLYGON: note(ctx.fill_polygon(call.poly.data(), call.poly.size())); break;
      case SCENE_OP_FILL_RECT_D: note(ctx.fill_rect(BLRect(call.d[0], call.d[1], call.d[2], call.d[3]))); break;
      case SCENE_OP_FILL_RECT_I: note(ctx.fill_rect(BLRectI(int(call.d[0]), int(call.d[1]), int(call.d[2]), int(call.d[3])))); break;
      case SCENE_OP_FILL_ALL: note(ctx.fill_all()); break;
      case SCENE_OP_BLIT_IMAGE: {
        const SceneBlit& b = call.blit;
        const BLImage& img = images[b.image_index];
        const BLRectI area(b.area[0], b.area[1], b.area[2], b.area[3]);
        switch (b.kind) {
          case 0: {
            const BLPointI o(int(b.dst[0]), int(b.dst[1]));
            note(b.has_area ? ctx.blit_image(o, img, area) : ctx.blit_image(o, img));
            break;
          }
          case 1: {
            const BLPoint o(b.dst[0], b.dst[1]);
            note(b.has_area ? ctx.blit_image(o, img, area) : ctx.blit_image(o, img));
            break;
          }
          case 2: {
            const BLRectI r(int(b.dst[0]), int(b.dst[1]), int(b.dst[2]), int(b.dst[3]));
            note(b.has_area ? ctx.blit_image(r, img, area) : ctx.blit_image(r, img));
            break;
          }
          default: {
            const BLRect r(b.dst[0], b.dst[1], b.dst[2], b.dst[3]);
            note(b.has_area ? ctx.blit_image(r, img, area) : ctx.blit_image(r, img));
            break;
          }
        }
        break;
      }
      case SCENE_OP_FILL_MASK: {
        const SceneMask& m = call.mask;
        const BLImage& img = images[m.image_index];
        const BLRectI area(m.area[0], m.area[1], m.area[2], m.area[3]);
        if (m.kind == 0) {
          const BLPointI o(int(m.dst[0]), int(m.dst[1]));
          note(m.has_area ? ctx.fill_mask(o, img, area) : ctx.fill_mask(o, img));
        }
        else {
          const BLPoint o(m.dst[0], m.dst[1]);
          note(m.has_area ? ctx.fill_mask(o, img, area) : ctx.fill_mask(o, img));
        }
        break;
      }
      case SCENE_OP_SET_FILL_ALPHA: note(ctx.set_fill_alpha(call.d[0])); break;
      case SCENE_OP_SET_STYLE_RGBA64: note(ctx.set_fill_style(BLRgba64(call.u64))); break;
      case SCENE_OP_SET_STYLE_RGBA: note(ctx.set_fill_style(BLRgba(call.f[0], call.f[1], call.f[2], call.f[3]))); break;
      case SCENE_OP_CLEAR_ALL: note(ctx.clear_all()); break;
      case SCENE_OP_CLEAR_RECT_I: note(ctx.clear_rect(BLRectI(call.i[0], call.i[1], call.i[2], call.i[3]))); break;
      case SCENE_OP_CLEAR_RECT_D: note(ctx.clear_rect(BLRect(call.d[0], call.d[1], call.d[2], call.d[3]))); break;
      case SCENE_OP_FILL_ALL_RGBA32: note(ctx.fill_all(BLRgba32(call.rgba32))); break;
      case SCENE_OP_FILL_RECT_I_RGBA32: note(ctx.fill_rect(BLRectI(call.i[0], call.i[1], call.i[2], call.i[3]), BLRgba32(call.rgba32))); break;
      case SCENE_OP_FILL_RECT_D_RGBA32: note(ctx.fill_rect(BLRect(call.d[0], call.d[1], call.d[2], call.d[3]), BLRgba32(call.rgba32))); break;
      case SCENE_OP_FILL_PATH_EX: {
        BLPoint origin(call.d[0], call.d[1]);
        if (!(call.flags & 1u)) origin.reset(0.0, 0.0);
        if (call.flags & 2u) note(ctx.fill_path(origin, call.path, BLRgba32(call.rgba32)));
        else note(ctx.fill_path(origin, call.path));
        break;
      }
      case SCENE_OP_FILL_EXPLICIT: {
        // the C++ API's style overloads route BLRgba32 / BLRgba64 to the
        // _rgba32 / _rgba64 entry points and BLGradient / BLPattern to _ext
        BLVar style;
        if (call.ex.style_kind == 0) style = BLRgba32(uint32_t(call.ex.rgba));
        else if (call.ex.style_kind == 1) style = BLRgba64(call.ex.rgba);
        else if (call.ex.style_kind == 2) style = call.gradient;
        else style = call.pattern;
        BLContextCore* core = &ctx;
        const BLUnknown* st = &style;
        switch (call.ex.target) {
          case 0:
            if (call.ex.style_kind == 1) note(bl_context_fill_all_rgba64(core, call.ex.rgba));
            else note(bl_context_fill_all_ext(core, st));
            break;
          case 1:
            if (call.ex.style_kind == 1) note(bl_context_fill_rect_i_rgba64(core, &call.rect_i, call.ex.rgba));
            else note(bl_context_fill_rect_i_ext(core, &call.rect_i, st));
            break;
          case 2:
            if (call.ex.style_kind == 1) note(bl_context_fill_rect_d_rgba64(core, &call.rect_d, call.ex.rgba));
            else note(bl_context_fill_rect_d_ext(core, &call.rect_d, st));
            break;
          case 3: {
            BLPoint origin(0.0, 0.0);
            if (call.ex.style_kind == 1) note(bl_context_fill_path_d_rgba64(core, &origin, &call.path, call.ex.rgba));
            else note(bl_context_fill_path_d_ext(core, &origin, &call.path, st));
            break;
          }
          case 4: {
            const void* data = call.blob.data();
            BLArrayView<uint8_t> view;
            if (call.ex.geom_type >= BL_GEOMETRY_TYPE_POLYLINEI && call.ex.geom_type <= BL_GEOMETRY_TYPE_ARRAY_VIEW_RECTD) {
              view.data = call.blob.data();
              view.size = call.ex.count;
              data = &view;
            }
            if (call.ex.style_kind == 1) note(bl_context_fill_geometry_rgba64(core, BLGeometryType(call.ex.geom_type), data, call.ex.rgba));
            else note(bl_context_fill_geometry_ext(core, BLGeometryType(call.ex.geom_type), data, st));
            break;
          }
          case 5: {
            const SceneMask& m = call.mask;
            const BLImageCore* img = &images[m.image_index];
            const BLRectI area(m.area[0], m.area[1], m.area[2], m.area[3]);
            const BLRectI* a = m.has_area ? &area : nullptr;
            if (m.kind == 0) {
              const BLPointI o(int(m.dst[0]), int(m.dst[1]));
              if (call.ex.style_kind == 1) note(bl_context_fill_mask_i_rgba64(core, &o, img, a, call.ex.rgba));
              else note(bl_context_fill_mask_i_ext(core, &o, img, a, st));
            }
            else {
              const BLPoint o(m.dst[0], m.dst[1]);
              if (call.ex.style_kind == 1) note(bl_context_fill_mask_d_rgba64(core, &o, img, a, call.ex.rgba));
              else note(bl_context_fill_mask_d_ext(core, &o, img, a, st));
            }
            break;
          }
          default: break;
        }
        break;
      }
      case SCENE_OP_FILL_GEOMETRY: {
        const void* data = call.blob.data();
        BLArrayView<uint8_t> view;
        bool is_array = (call.gtype >= BL_GEOMETRY_TYPE_POLYLINEI && call.gtype <= BL_GEOMETRY_TYPE_ARRAY_VIEW_RECTD);
        if (is_array) {
          view.data = call.blob.data();
          view.size = call.u0;
          data = &view;
        }
        BLContextCore* core = &ctx;
        if (call.flags)
          note(bl_context_fill_geometry_rgba32(core, BLGeometryType(call.gtype), data, call.rgba32));
        else
          note(bl_context_fill_geometry(core, BLGeometryType(call.gtype), data));
        break;
      }
      default: break;
    }
  }
#if defined(REF_RENDER_AMD)
  // The bound backend checks geometry while the device renders: a fill that
  // fails there is reported by the flush through the accumulated error
  // flags (the reference's own asynchronous mode does the same for deferred
  // work), so the driver flushes and turns the flags into the result the
  // synchronous mode returns from the failing call.
  note(ctx.flush(BL_CONTEXT_FLUSH_SYNC));
  const BLContextErrorFlags ef = ctx.accumulated_error_flags();
  if (ef & BL_CONTEXT_ERROR_FLAG_INVALID_GEOMETRY) note(BL_ERROR_INVALID_GEOMETRY);
  if (ef & BL_CONTEXT_ERROR_FLAG_OUT_OF_MEMORY) note(BL_ERROR_OUT_OF_MEMORY);
#endif
  note(ctx.end());
  auto t1 = std::chrono::steady_clock::now();
  if (elapsed_ms)
    *elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();

  BLImageData d;
  canvas.get_data(&d);
  for (uint32_t y = 0; y < hdr.height; y++)
    std::memcpy(out + size_t(y) * hdr.width * bpp, static_cast<const uint8_t*>(d.pixel_data) + intptr_t(y) * d.stride, size_t(hdr.width) * bpp);
  return first_error;
}

}  // extern "C"
