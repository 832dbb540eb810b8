// Scene player: issues recorded scenes (scene_format.h) through the C ABI of
// libblend2d_amd.so from native code, and times them the way the reference's
// own benchmark times a backend (blend2d-testing/bench/bl_bench_backend.cpp:
// 50-89): wall clock from the first begin() to the return of a synchronous
// flush / end(), with every call object built beforehand.  The same scenes
// go through the reference library by oracle/ref_render.cpp, so the two
// timings cover the same calls.
//
// It is a client of the rendering library (linked against it, built into
// libb2amd_player.so), not part of it: the bench and the tests use it to
// drive thousands of canvases without a Python call per fill.

#include <chrono>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <new>
#include <thread>
#include <vector>

#include "../../include/blend2d_amd.h"
#include "scene_format.h"

namespace b2p {

constexpr uint32_t kBadScene = 0x7FFFFFFFu;

inline size_t pad8(size_t n) { return (n + 7u) & ~size_t(7); }

// BLArrayView (core/api.h): data pointer and element count.
using ArrayView = BLAmdArrayView;

// A decoded call.  Pointers refer into the scene's own bytes, which the
// frame keeps; gradient stops are split into the arrays the C ABI takes.
struct Call {
  uint32_t op = 0;
  uint32_t u0 = 0;
  uint64_t u64 = 0;
  double d[6] = {};
  int32_t i[4] = {};
  float f[4] = {};
  const uint8_t* cmds = nullptr;
  const double* vtx = nullptr;
  size_t n = 0;
  const void* geom = nullptr;     // geometry struct, or &view for array types
  ArrayView view{};
  bool has_origin = false;
  bool has_rgba32 = false;
  uint32_t target = 0;            // FILL_EXPLICIT target
  uint32_t geom_type = 0;         // FILL_GEOMETRY / FILL_EXPLICIT: BLGeometryType
  // gradient / pattern (SET_STYLE_* and FILL_EXPLICIT)
  uint32_t gtype = 0, extend = 0;
  double values[6] = {};
  double transform[6] = {};
  bool has_transform = false;
  std::vector<double> stop_offsets;
  std::vector<uint32_t> stop_rgba;
  size_t image = 0;
  BLAmdStyle style{};
  bool has_area = false;          // BLIT_IMAGE / FILL_MASK: area in i[], kind in u0, destination in d[]
  size_t mask = 0;                // FILL_MASK / FILL_EXPLICIT target 5: the mask image (`image` may be
                                  // the explicit style's pattern)
};

struct Scene {
  std::vector<uint64_t> bytes;    // 8-aligned copy of the scene
  SceneHeader hdr{};
  std::vector<BLAmdImage*> images;
  BLAmdImage* target = nullptr;
  std::vector<Call> calls;
};

// Geometry types whose data is a BLArrayView (core/geometry.h): polylines,
// polygons and the box / rect array views.
bool is_array_geometry(uint32_t type) {
  return type >= BL_AMD_GEOMETRY_POLYLINEI && type <= BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD;
}

bool is_render_op(uint32_t op) {
  switch (op) {
    case SCENE_OP_FILL_PATH: case SCENE_OP_FILL_POLYGON: case SCENE_OP_FILL_RECT_D: case SCENE_OP_FILL_RECT_I:
    case SCENE_OP_FILL_ALL: case SCENE_OP_CLEAR_ALL: case SCENE_OP_CLEAR_RECT_I: case SCENE_OP_CLEAR_RECT_D:
    case SCENE_OP_FILL_GEOMETRY: case SCENE_OP_FILL_ALL_RGBA32: case SCENE_OP_FILL_RECT_I_RGBA32:
    case SCENE_OP_FILL_RECT_D_RGBA32: case SCENE_OP_FILL_PATH_EX: case SCENE_OP_FILL_EXPLICIT:
    case SCENE_OP_BLIT_IMAGE: case SCENE_OP_FILL_MASK:
      return true;
    default:
      return false;
  }
}

// SceneMask at p into `c` (kind in u0, area in i[], x, y in d[0..1]).
bool decode_mask(Call& c, const uint8_t* p, const uint8_t* end, size_t n_images) {
  SceneMask m;
  if (size_t(end - p) < sizeof(m)) return false;
  std::memcpy(&m, p, sizeof(m));
  if (m.image_index >= n_images) return false;
  c.u0 = m.kind;
  c.mask = m.image_index;
  c.has_area = m.has_area != 0;
  std::memcpy(c.i, m.area, 16);
  std::memcpy(c.d, m.dst, 16);
  return true;
}

// SceneGradient + stops at p into `c`; returns the bytes used (0: malformed).
size_t decode_gradient(Call& c, const uint8_t* p, const uint8_t* end) {
  SceneGradient g;
  if (size_t(end - p) < sizeof(g)) return 0;
  std::memcpy(&g, p, sizeof(g));
  const size_t used = sizeof(g) + size_t(g.n_stops) * sizeof(SceneStop);
  if (size_t(end - p) < used) return 0;
  c.gtype = g.type;
  c.extend = g.extend;
  std::memcpy(c.values, g.values, sizeof(c.values));
  c.has_transform = g.has_transform != 0;
  std::memcpy(c.transform, g.transform, sizeof(c.transform));
  c.stop_offsets.resize(g.n_stops);
  c.stop_rgba.resize(g.n_stops);
  for (uint32_t k = 0; k < g.n_stops; k++) {
    SceneStop s;
    std::memcpy(&s, p + sizeof(g) + k * sizeof(SceneStop), sizeof(s));
    c.stop_offsets[k] = s.offset;
    c.stop_rgba[k] = s.rgba32;
  }
  return used;
}

size_t decode_pattern(Call& c, const uint8_t* p, const uint8_t* end, size_t n_images) {
  ScenePattern pt;
  if (size_t(end - p) < sizeof(pt)) return 0;
  std::memcpy(&pt, p, sizeof(pt));
  if (pt.image_index >= n_images) return 0;
  c.image = pt.image_index;
  c.extend = pt.extend;
  c.has_transform = pt.has_transform != 0;
  std::memcpy(c.transform, pt.transform, sizeof(c.transform));
  return sizeof(pt);
}

uint32_t decode(Scene& sc) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(sc.bytes.data());
  const uint8_t* end = p + sc.bytes.size() * 8;
  auto take = [&](void* out, size_t n) {
    if (size_t(end - p) < n) return false;
    std::memcpy(out, p, n);
    p += n;
    return true;
  };
  if (!take(&sc.hdr, sizeof(sc.hdr)) || sc.hdr.magic != SCENE_MAGIC || sc.hdr.version != SCENE_VERSION)
    return kBadScene;
  for (uint32_t k = 0; k < sc.hdr.n_images; k++) {
    SceneImage si;
    if (!take(&si, sizeof(si))) return kBadScene;
    const size_t bytes = size_t(si.height) * si.stride;
    if (size_t(end - p) < pad8(bytes)) return kBadScene;
    BLAmdImage* img = nullptr;
    uint32_t r = bl_amd_image_create(&img, int(si.width), int(si.height), si.format);
    if (r) return r;
    sc.images.push_back(img);
    if ((r = bl_amd_image_upload(img, p, intptr_t(si.stride)))) return r;
    p += pad8(bytes);
  }
  sc.calls.resize(sc.hdr.n_cmds);
  for (uint32_t k = 0; k < sc.hdr.n_cmds; k++) {
    SceneCmd cmd;
    if (!take(&cmd, sizeof(cmd)) || size_t(end - p) < cmd.size) return kBadScene;
    const uint8_t* pl = p;
    const uint8_t* pe = p + cmd.size;
    p = pe;
    Call& c = sc.calls[k];
    c.op = cmd.op;
    switch (cmd.op) {
      case SCENE_OP_SET_COMP_OP: case SCENE_OP_SET_FILL_RULE: case SCENE_OP_SET_STYLE_SOLID:
      case SCENE_OP_SET_PATTERN_QUALITY: case SCENE_OP_SET_GRADIENT_QUALITY: case SCENE_OP_FILL_ALL_RGBA32:
        std::memcpy(&c.u0, pl, 4);
        break;
      case SCENE_OP_SET_TRANSFORM:
        std::memcpy(c.d, pl, 48);
        break;
      case SCENE_OP_SET_GLOBAL_ALPHA: case SCENE_OP_SET_FILL_ALPHA:
        std::memcpy(c.d, pl, 8);
        break;
      case SCENE_OP_SET_STYLE_RGBA64:
        std::memcpy(&c.u64, pl, 8);
        break;
      case SCENE_OP_SET_STYLE_RGBA:
        std::memcpy(c.f, pl, 16);
        break;
      case SCENE_OP_FILL_RECT_I: case SCENE_OP_CLEAR_RECT_I:
        std::memcpy(c.i, pl, 16);
        break;
      case SCENE_OP_FILL_RECT_D: case SCENE_OP_CLEAR_RECT_D:
        std::memcpy(c.d, pl, 32);
        break;
      case SCENE_OP_FILL_RECT_I_RGBA32:
        std::memcpy(c.i, pl, 16);
        std::memcpy(&c.u0, pl + 16, 4);
        break;
      case SCENE_OP_FILL_RECT_D_RGBA32:
        std::memcpy(c.d, pl, 32);
        std::memcpy(&c.u0, pl + 32, 4);
        break;
      case SCENE_OP_FILL_PATH: {
        uint32_t n;
        std::memcpy(&n, pl, 4);
        if (size_t(pe - pl) < 8 + pad8(n) + size_t(n) * 16) return kBadScene;
        c.n = n;
        c.cmds = pl + 8;
        c.vtx = reinterpret_cast<const double*>(pl + 8 + pad8(n));
        break;
      }
      case SCENE_OP_FILL_POLYGON: {
        uint32_t n;
        std::memcpy(&n, pl, 4);
        if (size_t(pe - pl) < 8 + size_t(n) * 16) return kBadScene;
        c.n = n;
        c.vtx = reinterpret_cast<const double*>(pl + 8);
        break;
      }
      case SCENE_OP_FILL_PATH_EX: {
        ScenePathEx h;
        std::memcpy(&h, pl, sizeof(h));
        if (size_t(pe - pl) < sizeof(h) + pad8(h.n) + size_t(h.n) * 16) return kBadScene;
        c.n = h.n;
        c.has_origin = (h.flags & 1u) != 0;
        c.has_rgba32 = (h.flags & 2u) != 0;
        c.u0 = h.rgba32;
        c.d[0] = h.origin[0];
        c.d[1] = h.origin[1];
        c.cmds = pl + sizeof(h);
        c.vtx = reinterpret_cast<const double*>(pl + sizeof(h) + pad8(h.n));
        break;
      }
      case SCENE_OP_FILL_GEOMETRY: {
        SceneGeometry g;
        std::memcpy(&g, pl, sizeof(g));
        c.geom_type = g.type;
        c.has_rgba32 = g.has_rgba32 != 0;
        c.u0 = g.rgba32;
        const uint8_t* data = pl + sizeof(g);
        if (is_array_geometry(g.type)) {
          c.view = ArrayView{data, g.count};
          c.geom = &c.view;
        }
        else {
          c.geom = data < pe ? data : nullptr;
        }
        break;
      }
      case SCENE_OP_SET_STYLE_GRADIENT:
        if (!decode_gradient(c, pl, pe)) return kBadScene;
        break;
      case SCENE_OP_SET_STYLE_PATTERN:
        if (!decode_pattern(c, pl, pe, sc.images.size())) return kBadScene;
        break;
      case SCENE_OP_BLIT_IMAGE: {
        SceneBlit b;
        if (size_t(pe - pl) < sizeof(b)) return kBadScene;
        std::memcpy(&b, pl, sizeof(b));
        if (b.image_index >= sc.images.size()) return kBadScene;
        c.u0 = b.kind;
        c.image = b.image_index;
        c.has_area = b.has_area != 0;
        std::memcpy(c.i, b.area, 16);
        std::memcpy(c.d, b.dst, 32);
        break;
      }
      case SCENE_OP_FILL_MASK:
        if (!decode_mask(c, pl, pe, sc.images.size())) return kBadScene;
        break;
      case SCENE_OP_FILL_EXPLICIT: {
        SceneExplicit ex;
        std::memcpy(&ex, pl, sizeof(ex));
        const uint8_t* q = pl + sizeof(ex);
        c.target = ex.target;
        c.u64 = ex.rgba;
        c.geom_type = ex.geom_type;
        BLAmdStyle& st = c.style;
        st.kind = ex.style_kind;
        st.rgba32 = uint32_t(ex.rgba);
        st.rgba64 = ex.rgba;
        if (ex.style_kind == BL_AMD_STYLE_GRADIENT) {
          const size_t used = decode_gradient(c, q, pe);
          if (!used) return kBadScene;
          q += used;
        }
        else if (ex.style_kind == BL_AMD_STYLE_PATTERN) {
          const size_t used = decode_pattern(c, q, pe, sc.images.size());
          if (!used) return kBadScene;
          q += used;
        }
        switch (ex.target) {
          case 1: std::memcpy(c.i, q, 16); break;
          case 2: std::memcpy(c.d, q, 32); break;
          case 3:
            c.n = ex.count;
            c.cmds = q;
            c.vtx = reinterpret_cast<const double*>(q + pad8(ex.count));
            break;
          case 4:
            if (is_array_geometry(ex.geom_type)) {
              c.view = ArrayView{q, ex.count};
              c.geom = &c.view;
            }
            else {
              c.geom = q < pe ? q : nullptr;
            }
            break;
          case 5:
            if (!decode_mask(c, q, pe, sc.images.size())) return kBadScene;
            break;
          default: break;
        }
        break;
      }
      default:
        break;
    }
  }
  // explicit styles point into their own call (stable once the vector is sized)
  for (Call& c : sc.calls) {
    if (c.op != SCENE_OP_FILL_EXPLICIT) continue;
    BLAmdStyle& st = c.style;
    st.extend = c.extend;
    if (st.kind == BL_AMD_STYLE_GRADIENT) {
      st.gradient_type = c.gtype;
      st.values = c.values;
      st.stop_offsets = c.stop_offsets.data();
      st.stop_rgba32 = c.stop_rgba.data();
      st.n_stops = c.stop_offsets.size();
    }
    if (st.kind == BL_AMD_STYLE_PATTERN) st.image = sc.images[c.image];
    st.transform = c.has_transform ? c.transform : nullptr;
  }
  return 0;
}

// First error a scene's calls returned: the rendering-call count when it
// came (the failing call included) and its code.
struct Issued {
  uint64_t first_index = ~0ull;   // rendering-call count when the first error came
  uint32_t first_code = 0;
};

void note(Issued& out, uint64_t n_render, uint32_t r) {
  if (r && out.first_index == ~0ull) {
    out.first_index = n_render;
    out.first_code = r;
  }
}

void issue(BLAmdContext* ctx, const Scene& sc, Issued& out) {
  uint64_t n_render = 0;
  for (const Call& c : sc.calls) {
    if (is_render_op(c.op)) n_render++;
    uint32_t r = 0;
    switch (c.op) {
      case SCENE_OP_SET_COMP_OP: r = bl_amd_context_set_comp_op(ctx, c.u0); break;
      case SCENE_OP_SET_FILL_RULE: r = bl_amd_context_set_fill_rule(ctx, c.u0); break;
      case SCENE_OP_SET_TRANSFORM: r = bl_amd_context_set_transform(ctx, c.d); break;
      case SCENE_OP_SET_GLOBAL_ALPHA: r = bl_amd_context_set_global_alpha(ctx, c.d[0]); break;
      case SCENE_OP_SET_PATTERN_QUALITY: r = bl_amd_context_set_pattern_quality(ctx, c.u0); break;
      case SCENE_OP_SET_GRADIENT_QUALITY: r = bl_amd_context_set_gradient_quality(ctx, c.u0); break;
      case SCENE_OP_SET_STYLE_SOLID: r = bl_amd_context_set_fill_style_rgba32(ctx, c.u0); break;
      case SCENE_OP_SET_STYLE_RGBA64: r = bl_amd_context_set_fill_style_rgba64(ctx, c.u64); break;
      case SCENE_OP_SET_STYLE_RGBA: r = bl_amd_context_set_fill_style_rgba(ctx, c.f); break;
      case SCENE_OP_SET_FILL_ALPHA: r = bl_amd_context_set_fill_alpha(ctx, c.d[0]); break;
      case SCENE_OP_SET_STYLE_GRADIENT:
        r = bl_amd_context_set_fill_style_gradient(ctx, c.gtype, c.values, c.extend, c.stop_offsets.data(),
                                                   c.stop_rgba.data(), c.stop_offsets.size(),
                                                   c.has_transform ? c.transform : nullptr);
        break;
      case SCENE_OP_SET_STYLE_PATTERN:
        r = bl_amd_context_set_fill_style_pattern(ctx, sc.images[c.image], c.extend,
                                                  c.has_transform ? c.transform : nullptr);
        break;
      case SCENE_OP_FILL_PATH: r = bl_amd_context_fill_path_d(ctx, nullptr, c.cmds, c.vtx, c.n); break;
      case SCENE_OP_FILL_POLYGON: r = bl_amd_context_fill_polygon_d(ctx, c.vtx, c.n); break;
      case SCENE_OP_FILL_RECT_D: r = bl_amd_context_fill_rect_d(ctx, c.d[0], c.d[1], c.d[2], c.d[3]); break;
      case SCENE_OP_FILL_RECT_I: r = bl_amd_context_fill_rect_i(ctx, c.i[0], c.i[1], c.i[2], c.i[3]); break;
      case SCENE_OP_FILL_ALL: r = bl_amd_context_fill_all(ctx); break;
      case SCENE_OP_CLEAR_ALL: r = bl_amd_context_clear_all(ctx); break;
      case SCENE_OP_CLEAR_RECT_I: r = bl_amd_context_clear_rect_i(ctx, c.i[0], c.i[1], c.i[2], c.i[3]); break;
      case SCENE_OP_CLEAR_RECT_D: r = bl_amd_context_clear_rect_d(ctx, c.d[0], c.d[1], c.d[2], c.d[3]); break;
      case SCENE_OP_FILL_ALL_RGBA32: r = bl_amd_context_fill_all_rgba32(ctx, c.u0); break;
      case SCENE_OP_BLIT_IMAGE: {
        const int32_t* area = c.has_area ? c.i : nullptr;
        BLAmdImage* img = sc.images[c.image];
        switch (c.u0) {
          case 0: r = bl_amd_context_blit_image_i(ctx, int(c.d[0]), int(c.d[1]), img, area); break;
          case 1: r = bl_amd_context_blit_image_d(ctx, c.d[0], c.d[1], img, area); break;
          case 2:
            r = bl_amd_context_blit_scaled_image_i(ctx, int(c.d[0]), int(c.d[1]), int(c.d[2]), int(c.d[3]), img, area);
            break;
          default: r = bl_amd_context_blit_scaled_image_d(ctx, c.d[0], c.d[1], c.d[2], c.d[3], img, area); break;
        }
        break;
      }
      case SCENE_OP_FILL_MASK: {
        const int32_t* area = c.has_area ? c.i : nullptr;
        BLAmdImage* mask = sc.images[c.mask];
        r = c.u0 == 0 ? bl_amd_context_fill_mask_i(ctx, int(c.d[0]), int(c.d[1]), mask, area)
                      : bl_amd_context_fill_mask_d(ctx, c.d[0], c.d[1], mask, area);
        break;
      }
      case SCENE_OP_FILL_RECT_I_RGBA32:
        r = bl_amd_context_fill_rect_i_rgba32(ctx, c.i[0], c.i[1], c.i[2], c.i[3], c.u0);
        break;
      case SCENE_OP_FILL_RECT_D_RGBA32:
        r = bl_amd_context_fill_rect_d_rgba32(ctx, c.d[0], c.d[1], c.d[2], c.d[3], c.u0);
        break;
      case SCENE_OP_FILL_PATH_EX: {
        const double* origin = c.has_origin ? c.d : nullptr;
        r = c.has_rgba32 ? bl_amd_context_fill_path_d_rgba32(ctx, origin, c.cmds, c.vtx, c.n, c.u0)
                         : bl_amd_context_fill_path_d(ctx, origin, c.cmds, c.vtx, c.n);
        break;
      }
      case SCENE_OP_FILL_GEOMETRY:
        r = c.has_rgba32 ? bl_amd_context_fill_geometry_rgba32(ctx, c.geom_type, c.geom, c.u0)
                         : bl_amd_context_fill_geometry(ctx, c.geom_type, c.geom);
        break;
      case SCENE_OP_FILL_EXPLICIT: {
        const bool r64 = c.style.kind == BL_AMD_STYLE_RGBA64;
        const BLAmdStyle* st = &c.style;
        switch (c.target) {
          case 0: r = r64 ? bl_amd_context_fill_all_rgba64(ctx, c.u64) : bl_amd_context_fill_all_ext(ctx, st); break;
          case 1:
            r = r64 ? bl_amd_context_fill_rect_i_rgba64(ctx, c.i[0], c.i[1], c.i[2], c.i[3], c.u64)
                    : bl_amd_context_fill_rect_i_ext(ctx, c.i[0], c.i[1], c.i[2], c.i[3], st);
            break;
          case 2:
            r = r64 ? bl_amd_context_fill_rect_d_rgba64(ctx, c.d[0], c.d[1], c.d[2], c.d[3], c.u64)
                    : bl_amd_context_fill_rect_d_ext(ctx, c.d[0], c.d[1], c.d[2], c.d[3], st);
            break;
          case 3:
            r = r64 ? bl_amd_context_fill_path_d_rgba64(ctx, nullptr, c.cmds, c.vtx, c.n, c.u64)
                    : bl_amd_context_fill_path_d_ext(ctx, nullptr, c.cmds, c.vtx, c.n, st);
            break;
          case 4:
            r = r64 ? bl_amd_context_fill_geometry_rgba64(ctx, c.geom_type, c.geom, c.u64)
                    : bl_amd_context_fill_geometry_ext(ctx, c.geom_type, c.geom, st);
            break;
          case 5: {
            const int32_t* area = c.has_area ? c.i : nullptr;
            BLAmdImage* mask = sc.images[c.mask];
            if (c.u0 == 0) {
              r = r64 ? bl_amd_context_fill_mask_i_rgba64(ctx, int(c.d[0]), int(c.d[1]), mask, area, c.u64)
                      : bl_amd_context_fill_mask_i_ext(ctx, int(c.d[0]), int(c.d[1]), mask, area, st);
            }
            else {
              r = r64 ? bl_amd_context_fill_mask_d_rgba64(ctx, c.d[0], c.d[1], mask, area, c.u64)
                      : bl_amd_context_fill_mask_d_ext(ctx, c.d[0], c.d[1], mask, area, st);
            }
            break;
          }
          default: break;
        }
        break;
      }
      default:
        break;
    }
    note(out, n_render, r);
  }
}

}  // namespace b2p

using namespace b2p;

// A set of scenes with their target canvases and source images on the
// device, and the contexts that render them, kept from frame to frame as a
// renderer serving many frames keeps them (their device buffers are reused).
struct B2PFrame {
  std::vector<Scene> scenes;
  BLAmdContext* plain = nullptr;          // PLAIN: one context, canvas after canvas
  std::vector<BLAmdContext*> contexts;    // BATCHED: one context per canvas
  BLAmdBatch* batch = nullptr;
};

namespace b2p {

void accumulate(BLAmdStats& sum, const BLAmdStats& s) {
  sum.commands += s.commands;
  sum.sources += s.sources;
  sum.segments += s.segments;
  sum.binned += s.binned;
  sum.tiles += s.tiles;
  sum.pixels += s.pixels;
  sum.frame_bytes += s.frame_bytes;
  sum.lut_bytes += s.lut_bytes;
  sum.image_bytes += s.image_bytes;
  sum.pred_segs += s.pred_segs;
  sum.pred_bins += s.pred_bins;
  sum.sizing = s.sizing;
  sum.fallback += s.fallback;
  sum.syncs += s.syncs;
  sum.render_kernels |= s.render_kernels;
  sum.render_ms += s.render_ms;
  sum.total_ms += s.total_ms;
  sum.merge_ms += s.merge_ms;
  sum.upload_ms += s.upload_ms;
  sum.pack_ms += s.pack_ms;
}

// First error of a canvas in call order: the calls' own returns merged with
// the error the device found for a call during the flush.
uint32_t first_error(const Issued& is, BLAmdContext* ctx) {
  uint64_t index = is.first_index;
  uint32_t code = is.first_code;
  uint64_t call = 0;
  uint32_t dcode = 0;
  if (bl_amd_context_get_deferred_error(ctx, &call, &dcode) == BL_AMD_SUCCESS && dcode) {
    // the deferred index is 0-based; a call's own return counts the call
    if (index == ~0ull || call + 1 < index || (call + 1 == index && dcode < code)) {
      index = call + 1;
      code = dcode;
    }
  }
  return code;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace b2p

extern "C" {

enum { B2P_PLAIN = 0, B2P_BATCHED = 1 };

void b2p_frame_destroy(B2PFrame* f) {
  if (!f) return;
  if (f->batch) bl_amd_batch_destroy(f->batch);
  for (BLAmdContext* c : f->contexts)
    if (c) bl_amd_context_destroy(c);
  if (f->plain) bl_amd_context_destroy(f->plain);
  for (Scene& sc : f->scenes) {
    if (sc.target) bl_amd_image_destroy(sc.target);
    for (BLAmdImage* img : sc.images) bl_amd_image_destroy(img);
  }
  delete f;
}

// Decodes `n` scenes, uploads their source images and creates one target
// canvas per scene plus the contexts and the batch.  Returns 0, a
// BLAmdResult, or 0x7FFFFFFF for a malformed scene.
uint32_t b2p_frame_create(B2PFrame** out, const uint8_t* const* scenes, const size_t* sizes, size_t n) {
  if (!out) return BL_AMD_ERROR_INVALID_VALUE;
  *out = nullptr;
  B2PFrame* f = new (std::nothrow) B2PFrame();
  if (!f) return BL_AMD_ERROR_OUT_OF_MEMORY;
  uint32_t r = 0;
  f->scenes.resize(n);
  f->contexts.assign(n, nullptr);
  for (size_t k = 0; k < n && !r; k++) {
    Scene& sc = f->scenes[k];
    sc.bytes.assign((sizes[k] + 7) / 8, 0);
    std::memcpy(sc.bytes.data(), scenes[k], sizes[k]);
    r = decode(sc);
    if (!r) r = bl_amd_image_create(&sc.target, int(sc.hdr.width), int(sc.hdr.height), sc.hdr.format);
    if (!r) r = bl_amd_context_create(&f->contexts[k]);
  }
  if (!r) r = bl_amd_context_create(&f->plain);
  if (!r) r = bl_amd_batch_create(&f->batch);
  if (r) {
    b2p_frame_destroy(f);
    return r;
  }
  *out = f;
  return 0;
}

// Resets every canvas to its scene's clear value (done when this returns).
uint32_t b2p_frame_clear(B2PFrame* f) {
  if (!f) return BL_AMD_ERROR_INVALID_VALUE;
  for (Scene& sc : f->scenes)
    if (uint32_t r = bl_amd_image_fill_raw(sc.target, sc.hdr.clear_value)) return r;
  return 0;
}

// Renders every scene onto its canvas.  PLAIN: one context, canvas after
// canvas (begin, calls, synchronous flush, end).  BATCHED: every canvas's
// context begins and takes its calls -- on `threads` client threads, each
// taking every threads-th canvas (contexts are independent) -- then one
// synchronous batch flush renders them all and every context ends.
// `elapsed_ms` receives the wall time from the first begin() to the last
// end(), `issue_ms` (may be NULL) the part of it until every call was issued
// (PLAIN: summed over the canvases), `errors` (n entries, may be NULL) each
// canvas's first error in call order, `stats` (may be NULL) the counters of
// the flush (PLAIN: summed over the canvases).  Returns the first error of a
// begin / flush / end / batch call.
uint32_t b2p_frame_render(B2PFrame* f, uint32_t mode, uint32_t threads, double* elapsed_ms, double* issue_ms,
                          uint32_t* errors, BLAmdStats* stats) {
  if (!f || mode > B2P_BATCHED) return BL_AMD_ERROR_INVALID_VALUE;
  const size_t n = f->scenes.size();
  std::vector<Issued> issued(n);
  BLAmdStats sum{};
  uint32_t result = 0;
  auto keep = [&](uint32_t r) { if (r && !result) result = r; };

  const auto t0 = std::chrono::steady_clock::now();
  double elapsed = 0.0, issuing = 0.0;
  if (mode == B2P_PLAIN) {
    BLAmdContext* ctx = f->plain;
    for (size_t k = 0; k < n; k++) {
      const auto ti = std::chrono::steady_clock::now();
      keep(bl_amd_context_begin(ctx, f->scenes[k].target));
      issue(ctx, f->scenes[k], issued[k]);
      issuing += ms_since(ti);
      keep(bl_amd_context_flush(ctx, 0x80000000u));   // BL_CONTEXT_FLUSH_SYNC
      if (errors) errors[k] = first_error(issued[k], ctx);
      BLAmdStats s{};
      if (bl_amd_context_get_stats(ctx, &s) == BL_AMD_SUCCESS) accumulate(sum, s);
      keep(bl_amd_context_end(ctx));
    }
    elapsed = ms_since(t0);
  }
  else {
    keep(bl_amd_batch_clear(f->batch));
    const size_t nt = std::max<size_t>(1, std::min<size_t>(threads, n));
    std::vector<uint32_t> begun(n, 0);
    auto record = [&](size_t t) {
      for (size_t k = t; k < n; k += nt) {
        begun[k] = bl_amd_context_begin(f->contexts[k], f->scenes[k].target);
        issue(f->contexts[k], f->scenes[k], issued[k]);
      }
    };
    if (nt == 1) {
      record(0);
    }
    else {
      std::vector<std::thread> pool;
      for (size_t t = 1; t < nt; t++) pool.emplace_back(record, t);
      record(0);
      for (std::thread& th : pool) th.join();
    }
    for (size_t k = 0; k < n; k++) {
      keep(begun[k]);
      keep(bl_amd_batch_add(f->batch, f->contexts[k]));
    }
    issuing = ms_since(t0);
    keep(bl_amd_batch_flush(f->batch));
    for (size_t k = 0; k < n; k++)
      if (errors) errors[k] = first_error(issued[k], f->contexts[k]);
    for (size_t k = 0; k < n; k++) keep(bl_amd_context_end(f->contexts[k]));
    elapsed = ms_since(t0);
    bl_amd_batch_get_stats(f->batch, &sum);
  }
  if (elapsed_ms) *elapsed_ms = elapsed;
  if (issue_ms) *issue_ms = issuing;
  if (stats) *stats = sum;
  return result;
}

// Renders the batch's last frame again from its device-resident copy
// (bl_amd_batch_replay): the device stages only.  `elapsed_ms` receives the
// wall time of the replay, `stats` its counters.
uint32_t b2p_frame_replay(B2PFrame* f, double* elapsed_ms, BLAmdStats* stats) {
  if (!f) return BL_AMD_ERROR_INVALID_VALUE;
  const auto t0 = std::chrono::steady_clock::now();
  uint32_t r = bl_amd_batch_replay(f->batch);
  if (elapsed_ms) *elapsed_ms = ms_since(t0);
  if (stats) bl_amd_batch_get_stats(f->batch, stats);
  return r;
}

// Runs the autotuner on the batched frame resident in the device
// (bl_amd_batch_autotune): its replays render onto the canvases again.
uint32_t b2p_frame_autotune(B2PFrame* f, uint32_t reps, BLAmdTune* out) {
  if (!f) return BL_AMD_ERROR_INVALID_VALUE;
  return bl_amd_batch_autotune(f->batch, reps, out);
}

size_t b2p_frame_count(const B2PFrame* f) { return f ? f->scenes.size() : 0; }

// Copies canvas `k` to `out` (rows of width * bytes-per-pixel, no padding).
uint32_t b2p_frame_download(const B2PFrame* f, size_t k, void* out, size_t size) {
  if (!f || k >= f->scenes.size()) return BL_AMD_ERROR_INVALID_VALUE;
  const Scene& sc = f->scenes[k];
  const size_t bpp = sc.hdr.format == SCENE_FMT_A8 ? 1 : 4;
  if (size < size_t(sc.hdr.width) * sc.hdr.height * bpp) return BL_AMD_ERROR_INVALID_VALUE;
  return bl_amd_image_download(sc.target, out, intptr_t(sc.hdr.width * bpp));
}

}  // extern "C"
