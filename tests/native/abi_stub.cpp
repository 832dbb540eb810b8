// A recording stand-in for the C ABI of include/blend2d_amd.h, for CPU tests
// of the callers of that ABI: every entry point appends one line naming the
// call and the arguments it received (pointed-to data hashed with FNV-1a) to
// a log and succeeds without rendering.  The scene player
// (blend2d_amd/csrc/scene_player.cpp) is compiled into the same library, so
// the Python mirror and the native player can drive the same scene through
// it and their call logs compared line by line.
//
//   g++ -std=c++17 -O1 -fPIC -shared tests/native/abi_stub.cpp -o <out>.so

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/blend2d_amd.h"

namespace {

std::string g_log;
uint32_t g_serial = 0;

struct StubImage { uint32_t serial; int w, h; uint32_t format; };
struct StubContext { uint32_t serial; };

void logf(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_log += buf;
  g_log += '\n';
}

uint64_t fnv(const void* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  const uint8_t* b = static_cast<const uint8_t*>(p);
  for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
  return h;
}

std::string dbl(const double* v, size_t n) {
  std::string s;
  char buf[40];
  for (size_t i = 0; i < n; i++) {
    snprintf(buf, sizeof(buf), "%s%.17g", i ? "," : "", v[i]);
    s += buf;
  }
  return s;
}

std::string transform(const double* m) { return m ? dbl(m, 6) : std::string("none"); }

uint32_t serial_of(const BLAmdImage* img) {
  return img ? reinterpret_cast<const StubImage*>(img)->serial : 0u;
}

std::string gradient(uint32_t type, const double* values, uint32_t extend, const double* offs,
                     const uint32_t* cols, size_t n, const double* tr) {
  std::string s = "type=" + std::to_string(type) + " extend=" + std::to_string(extend) +
                  " values=" + dbl(values, 6) + " stops=";
  char buf[64];
  for (size_t i = 0; i < n; i++) {
    snprintf(buf, sizeof(buf), "%s%.17g:%08x", i ? "," : "", offs[i], cols[i]);
    s += buf;
  }
  return s + " transform=" + transform(tr);
}

std::string style(const BLAmdStyle* st) {
  if (!st) return "style=null";
  char buf[64];
  switch (st->kind) {
    case BL_AMD_STYLE_RGBA32: snprintf(buf, sizeof(buf), "rgba32=%08x", st->rgba32); return buf;
    case BL_AMD_STYLE_RGBA64: snprintf(buf, sizeof(buf), "rgba64=%016llx", (unsigned long long)st->rgba64); return buf;
    case BL_AMD_STYLE_GRADIENT:
      return "gradient " + gradient(st->gradient_type, st->values, st->extend, st->stop_offsets, st->stop_rgba32,
                                    st->n_stops, st->transform);
    case BL_AMD_STYLE_PATTERN:
      return "pattern image=" + std::to_string(serial_of(st->image)) + " extend=" + std::to_string(st->extend) +
             " transform=" + transform(st->transform);
    default: return "kind=" + std::to_string(st->kind);
  }
}

// Element size of array geometries, struct size of the others (0: not read).
size_t geometry_bytes(uint32_t type, bool* is_view) {
  *is_view = type >= BL_AMD_GEOMETRY_POLYLINEI && type <= BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD;
  switch (type) {
    case BL_AMD_GEOMETRY_BOXI: case BL_AMD_GEOMETRY_RECTI: return 16;
    case BL_AMD_GEOMETRY_BOXD: case BL_AMD_GEOMETRY_RECTD: return 32;
    case BL_AMD_GEOMETRY_CIRCLE: return 24;
    case BL_AMD_GEOMETRY_ELLIPSE: case BL_AMD_GEOMETRY_LINE: return 32;
    case BL_AMD_GEOMETRY_ROUND_RECT: case BL_AMD_GEOMETRY_ARC: case BL_AMD_GEOMETRY_CHORD:
    case BL_AMD_GEOMETRY_PIE: case BL_AMD_GEOMETRY_TRIANGLE: return 48;
    case BL_AMD_GEOMETRY_POLYLINEI: case BL_AMD_GEOMETRY_POLYGONI: return 8;
    case BL_AMD_GEOMETRY_POLYLINED: case BL_AMD_GEOMETRY_POLYGOND: return 16;
    case BL_AMD_GEOMETRY_ARRAY_VIEW_BOXI: case BL_AMD_GEOMETRY_ARRAY_VIEW_RECTI: return 16;
    case BL_AMD_GEOMETRY_ARRAY_VIEW_BOXD: case BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD: return 32;
    default: return 0;
  }
}

std::string geometry(uint32_t type, const void* data) {
  char buf[96];
  if (!data) { snprintf(buf, sizeof(buf), "type=%u null", type); return buf; }
  bool view = false;
  const size_t size = geometry_bytes(type, &view);
  if (view) {
    const BLAmdArrayView* v = static_cast<const BLAmdArrayView*>(data);
    snprintf(buf, sizeof(buf), "type=%u view n=%zu h=%016llx", type, v->size,
             (unsigned long long)(v->size ? fnv(v->data, v->size * size) : 0));
    return buf;
  }
  snprintf(buf, sizeof(buf), "type=%u struct h=%016llx", type, (unsigned long long)(size ? fnv(data, size) : 0));
  return buf;
}

std::string path(const double* origin, const uint8_t* cmds, const double* vtx, size_t n) {
  char buf[128];
  snprintf(buf, sizeof(buf), "origin=%s n=%zu cmds=%016llx vtx=%016llx", origin ? dbl(origin, 2).c_str() : "none", n,
           (unsigned long long)fnv(cmds, n), (unsigned long long)fnv(vtx, n * 16));
  return buf;
}

std::string area_of(const int32_t* a) {
  return a ? std::to_string(a[0]) + " " + std::to_string(a[1]) + " " + std::to_string(a[2]) + " " + std::to_string(a[3])
           : std::string("none");
}

}  // namespace

extern "C" {

// Test hooks: the log so far, and a reset of log and handle serials.
const char* stub_log(void) { return g_log.c_str(); }
void stub_reset(void) { g_log.clear(); g_serial = 0; }

uint32_t bl_amd_version(void) { return 0; }
int bl_amd_device_count(void) { return 0; }

BLAmdResult bl_amd_image_create(BLAmdImage** out, int w, int h, uint32_t format) {
  StubImage* img = new StubImage{++g_serial, w, h, format};
  *out = reinterpret_cast<BLAmdImage*>(img);
  logf("image_create %d %d %u -> %u", w, h, format, img->serial);
  return 0;
}
BLAmdResult bl_amd_image_destroy(BLAmdImage* img) { delete reinterpret_cast<StubImage*>(img); return 0; }
BLAmdResult bl_amd_image_upload(BLAmdImage* img, const void* pixels, intptr_t stride) {
  const StubImage* s = reinterpret_cast<const StubImage*>(img);
  const size_t row = size_t(s->w) * (s->format == BL_AMD_FORMAT_A8 ? 1 : 4);
  uint64_t h = 0;
  for (int y = 0; y < s->h; y++) h = h * 31 + fnv(static_cast<const uint8_t*>(pixels) + y * stride, row);
  logf("image_upload %u %016llx", s->serial, (unsigned long long)h);
  return 0;
}
// Writes a marker instead of pixels: byte x of row y is (3 * y + x) & 0xFF,
// so a caller's test can see that the write-back reached its own buffer.
BLAmdResult bl_amd_image_download(const BLAmdImage* img, void* pixels, intptr_t stride) {
  const StubImage* s = reinterpret_cast<const StubImage*>(img);
  const size_t row = size_t(s->w) * (s->format == BL_AMD_FORMAT_A8 ? 1 : 4);
  for (int y = 0; y < s->h; y++) {
    uint8_t* p = static_cast<uint8_t*>(pixels) + y * stride;
    for (size_t x = 0; x < row; x++) p[x] = uint8_t(3 * y + x);
  }
  logf("image_download %u", s->serial);
  return 0;
}
BLAmdResult bl_amd_image_fill_raw(BLAmdImage* img, uint32_t v) { logf("fill_raw %u %08x", serial_of(img), v); return 0; }
BLAmdResult bl_amd_image_get_data(const BLAmdImage*, int*, int*, uint32_t*, void**, intptr_t*) { return 0; }

BLAmdResult bl_amd_context_create(BLAmdContext** out) {
  *out = reinterpret_cast<BLAmdContext*>(new StubContext{++g_serial});
  return 0;
}
BLAmdResult bl_amd_context_destroy(BLAmdContext* c) { delete reinterpret_cast<StubContext*>(c); return 0; }
BLAmdResult bl_amd_context_begin(BLAmdContext*, BLAmdImage* target) { logf("begin %u", serial_of(target)); return 0; }
BLAmdResult bl_amd_context_end(BLAmdContext*) { logf("end"); return 0; }
BLAmdResult bl_amd_context_reset(BLAmdContext*) { logf("reset"); return 0; }
BLAmdResult bl_amd_context_flush(BLAmdContext*, uint32_t) { logf("flush"); return 0; }
BLAmdResult bl_amd_context_get_error_flags(const BLAmdContext*, uint32_t* f) { *f = 0; return 0; }
BLAmdResult bl_amd_context_get_deferred_error(const BLAmdContext*, uint64_t* i, uint32_t* c) { *i = 0; *c = 0; return 0; }

BLAmdResult bl_amd_context_set_comp_op(BLAmdContext*, uint32_t v) { logf("set_comp_op %u", v); return 0; }
BLAmdResult bl_amd_context_set_fill_rule(BLAmdContext*, uint32_t v) { logf("set_fill_rule %u", v); return 0; }
BLAmdResult bl_amd_context_set_global_alpha(BLAmdContext*, double a) { logf("set_global_alpha %.17g", a); return 0; }
BLAmdResult bl_amd_context_set_transform(BLAmdContext*, const double m[6]) { logf("set_transform %s", dbl(m, 6).c_str()); return 0; }
BLAmdResult bl_amd_context_set_pattern_quality(BLAmdContext*, uint32_t q) { logf("set_pattern_quality %u", q); return 0; }
BLAmdResult bl_amd_context_set_fill_style_rgba32(BLAmdContext*, uint32_t v) { logf("style_rgba32 %08x", v); return 0; }
BLAmdResult bl_amd_context_set_fill_style_rgba64(BLAmdContext*, uint64_t v) { logf("style_rgba64 %016llx", (unsigned long long)v); return 0; }
BLAmdResult bl_amd_context_set_fill_style_rgba(BLAmdContext*, const float v[4]) {
  logf("style_rgba %.9g %.9g %.9g %.9g", v[0], v[1], v[2], v[3]);
  return 0;
}
BLAmdResult bl_amd_context_set_fill_alpha(BLAmdContext*, double a) { logf("set_fill_alpha %.17g", a); return 0; }
BLAmdResult bl_amd_context_set_fill_style_gradient(BLAmdContext*, uint32_t type, const double* values, uint32_t extend,
                                                   const double* offs, const uint32_t* cols, size_t n, const double* tr) {
  logf("style_gradient %s", gradient(type, values, extend, offs, cols, n, tr).c_str());
  return 0;
}
BLAmdResult bl_amd_context_set_fill_style_pattern(BLAmdContext*, const BLAmdImage* img, uint32_t extend, const double* tr) {
  logf("style_pattern image=%u extend=%u transform=%s", serial_of(img), extend, transform(tr).c_str());
  return 0;
}

BLAmdResult bl_amd_context_fill_path_d(BLAmdContext*, const double* o, const uint8_t* c, const double* v, size_t n) {
  logf("fill_path %s", path(o, c, v, n).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_polygon_d(BLAmdContext*, const double* v, size_t n) {
  logf("fill_polygon n=%zu %016llx", n, (unsigned long long)fnv(v, n * 16));
  return 0;
}
BLAmdResult bl_amd_context_fill_rect_d(BLAmdContext*, double x, double y, double w, double h) {
  logf("fill_rect_d %.17g %.17g %.17g %.17g", x, y, w, h);
  return 0;
}
BLAmdResult bl_amd_context_fill_rect_i(BLAmdContext*, int x, int y, int w, int h) { logf("fill_rect_i %d %d %d %d", x, y, w, h); return 0; }
BLAmdResult bl_amd_context_fill_all(BLAmdContext*) { logf("fill_all"); return 0; }
BLAmdResult bl_amd_context_fill_all_rgba32(BLAmdContext*, uint32_t c) { logf("fill_all rgba32=%08x", c); return 0; }
BLAmdResult bl_amd_context_fill_rect_i_rgba32(BLAmdContext*, int x, int y, int w, int h, uint32_t c) {
  logf("fill_rect_i %d %d %d %d rgba32=%08x", x, y, w, h, c);
  return 0;
}
BLAmdResult bl_amd_context_fill_rect_d_rgba32(BLAmdContext*, double x, double y, double w, double h, uint32_t c) {
  logf("fill_rect_d %.17g %.17g %.17g %.17g rgba32=%08x", x, y, w, h, c);
  return 0;
}
BLAmdResult bl_amd_context_fill_path_d_rgba32(BLAmdContext*, const double* o, const uint8_t* c, const double* v, size_t n, uint32_t col) {
  logf("fill_path %s rgba32=%08x", path(o, c, v, n).c_str(), col);
  return 0;
}
BLAmdResult bl_amd_context_fill_all_rgba64(BLAmdContext*, uint64_t c) { logf("fill_all rgba64=%016llx", (unsigned long long)c); return 0; }
BLAmdResult bl_amd_context_fill_rect_i_rgba64(BLAmdContext*, int x, int y, int w, int h, uint64_t c) {
  logf("fill_rect_i %d %d %d %d rgba64=%016llx", x, y, w, h, (unsigned long long)c);
  return 0;
}
BLAmdResult bl_amd_context_fill_rect_d_rgba64(BLAmdContext*, double x, double y, double w, double h, uint64_t c) {
  logf("fill_rect_d %.17g %.17g %.17g %.17g rgba64=%016llx", x, y, w, h, (unsigned long long)c);
  return 0;
}
BLAmdResult bl_amd_context_fill_path_d_rgba64(BLAmdContext*, const double* o, const uint8_t* c, const double* v, size_t n, uint64_t col) {
  logf("fill_path %s rgba64=%016llx", path(o, c, v, n).c_str(), (unsigned long long)col);
  return 0;
}
BLAmdResult bl_amd_context_fill_geometry_rgba64(BLAmdContext*, uint32_t type, const void* data, uint64_t c) {
  logf("fill_geometry %s rgba64=%016llx", geometry(type, data).c_str(), (unsigned long long)c);
  return 0;
}
BLAmdResult bl_amd_context_fill_all_ext(BLAmdContext*, const BLAmdStyle* st) { logf("fill_all %s", style(st).c_str()); return 0; }
BLAmdResult bl_amd_context_fill_rect_i_ext(BLAmdContext*, int x, int y, int w, int h, const BLAmdStyle* st) {
  logf("fill_rect_i %d %d %d %d %s", x, y, w, h, style(st).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_rect_d_ext(BLAmdContext*, double x, double y, double w, double h, const BLAmdStyle* st) {
  logf("fill_rect_d %.17g %.17g %.17g %.17g %s", x, y, w, h, style(st).c_str());
  return 0;
}

BLAmdResult bl_amd_context_set_gradient_quality(BLAmdContext*, uint32_t q) {
  logf("set_gradient_quality %u", q);
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_i(BLAmdContext*, int x, int y, const BLAmdImage* mask, const int32_t* area) {
  logf("fill_mask_i %d %d mask image=%u area=%s", x, y, serial_of(mask), area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_d(BLAmdContext*, double x, double y, const BLAmdImage* mask, const int32_t* area) {
  logf("fill_mask_d %.17g %.17g mask image=%u area=%s", x, y, serial_of(mask), area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_i_rgba32(BLAmdContext*, int x, int y, const BLAmdImage* mask, const int32_t* area,
                                              uint32_t c) {
  logf("fill_mask_i %d %d mask image=%u area=%s rgba32=%08x", x, y, serial_of(mask), area_of(area).c_str(), c);
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_d_rgba32(BLAmdContext*, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint32_t c) {
  logf("fill_mask_d %.17g %.17g mask image=%u area=%s rgba32=%08x", x, y, serial_of(mask), area_of(area).c_str(), c);
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_i_rgba64(BLAmdContext*, int x, int y, const BLAmdImage* mask, const int32_t* area,
                                              uint64_t c) {
  logf("fill_mask_i %d %d mask image=%u area=%s rgba64=%016llx", x, y, serial_of(mask), area_of(area).c_str(),
       (unsigned long long)c);
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_d_rgba64(BLAmdContext*, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint64_t c) {
  logf("fill_mask_d %.17g %.17g mask image=%u area=%s rgba64=%016llx", x, y, serial_of(mask), area_of(area).c_str(),
       (unsigned long long)c);
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_i_ext(BLAmdContext*, int x, int y, const BLAmdImage* mask, const int32_t* area,
                                           const BLAmdStyle* st) {
  logf("fill_mask_i %d %d mask image=%u area=%s %s", x, y, serial_of(mask), area_of(area).c_str(), style(st).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_mask_d_ext(BLAmdContext*, double x, double y, const BLAmdImage* mask,
                                           const int32_t* area, const BLAmdStyle* st) {
  logf("fill_mask_d %.17g %.17g mask image=%u area=%s %s", x, y, serial_of(mask), area_of(area).c_str(),
       style(st).c_str());
  return 0;
}
BLAmdResult bl_amd_context_blit_image_i(BLAmdContext*, int x, int y, const BLAmdImage* image, const int32_t* area) {
  logf("blit_image_i %d %d image=%u area=%s", x, y, serial_of(image), area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_blit_image_d(BLAmdContext*, double x, double y, const BLAmdImage* image, const int32_t* area) {
  logf("blit_image_d %.17g %.17g image=%u area=%s", x, y, serial_of(image), area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_blit_scaled_image_i(BLAmdContext*, int x, int y, int w, int h, const BLAmdImage* image,
                                               const int32_t* area) {
  logf("blit_scaled_image_i %d %d %d %d image=%u area=%s", x, y, w, h, serial_of(image), area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_blit_scaled_image_d(BLAmdContext*, double x, double y, double w, double h,
                                               const BLAmdImage* image, const int32_t* area) {
  logf("blit_scaled_image_d %.17g %.17g %.17g %.17g image=%u area=%s", x, y, w, h, serial_of(image),
       area_of(area).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_path_d_ext(BLAmdContext*, const double* o, const uint8_t* c, const double* v, size_t n, const BLAmdStyle* st) {
  logf("fill_path %s %s", path(o, c, v, n).c_str(), style(st).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_geometry_ext(BLAmdContext*, uint32_t type, const void* data, const BLAmdStyle* st) {
  logf("fill_geometry %s %s", geometry(type, data).c_str(), style(st).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_geometry(BLAmdContext*, uint32_t type, const void* data) {
  logf("fill_geometry %s", geometry(type, data).c_str());
  return 0;
}
BLAmdResult bl_amd_context_fill_geometry_rgba32(BLAmdContext*, uint32_t type, const void* data, uint32_t c) {
  logf("fill_geometry %s rgba32=%08x", geometry(type, data).c_str(), c);
  return 0;
}
BLAmdResult bl_amd_context_clear_all(BLAmdContext*) { logf("clear_all"); return 0; }
BLAmdResult bl_amd_context_clear_rect_i(BLAmdContext*, int x, int y, int w, int h) { logf("clear_rect_i %d %d %d %d", x, y, w, h); return 0; }
BLAmdResult bl_amd_context_clear_rect_d(BLAmdContext*, double x, double y, double w, double h) {
  logf("clear_rect_d %.17g %.17g %.17g %.17g", x, y, w, h);
  return 0;
}

BLAmdResult bl_amd_context_get_stats(const BLAmdContext*, BLAmdStats* out) { std::memset(out, 0, sizeof(*out)); return 0; }
BLAmdResult bl_amd_context_replay(BLAmdContext*) { return 0; }
BLAmdResult bl_amd_context_replay_noalloc(BLAmdContext*) { return 0; }

BLAmdResult bl_amd_batch_create(BLAmdBatch** out) { *out = reinterpret_cast<BLAmdBatch*>(new StubContext{++g_serial}); return 0; }
BLAmdResult bl_amd_batch_destroy(BLAmdBatch* b) { delete reinterpret_cast<StubContext*>(b); return 0; }
BLAmdResult bl_amd_batch_add(BLAmdBatch*, BLAmdContext*) { logf("batch_add"); return 0; }
BLAmdResult bl_amd_batch_clear(BLAmdBatch*) { return 0; }
BLAmdResult bl_amd_batch_flush(BLAmdBatch*) { logf("flush"); return 0; }
BLAmdResult bl_amd_batch_replay(BLAmdBatch*) { return 0; }
BLAmdResult bl_amd_batch_get_stats(const BLAmdBatch*, BLAmdStats* out) { std::memset(out, 0, sizeof(*out)); return 0; }

BLAmdResult bl_amd_set_sizing(uint32_t) { return 0; }
BLAmdResult bl_amd_cost_model_set(const float*) { return 0; }
BLAmdResult bl_amd_context_collect_costs(BLAmdContext*, float*, float*, size_t, size_t* n) { *n = 0; return 0; }
BLAmdResult bl_amd_cost_model_forward(const float*, size_t, const float*, float*) { return 0; }
BLAmdResult bl_amd_cost_model_grad(const float*, const float*, size_t, const float*, float*, float*) { return 0; }
BLAmdResult bl_amd_cost_model_train(const float*, const float*, size_t, float*, uint32_t, uint32_t, float, float*) { return 0; }
BLAmdResult bl_amd_set_render_schedule(uint32_t, uint32_t) { return 0; }
BLAmdResult bl_amd_tile_model_set(const float*) { return 0; }
BLAmdResult bl_amd_context_collect_tile_costs(BLAmdContext*, float*, float*, size_t, size_t* n) { if (n) *n = 0; return 0; }
BLAmdResult bl_amd_batch_collect_tile_costs(BLAmdBatch*, float*, float*, size_t, size_t* n) { if (n) *n = 0; return 0; }
BLAmdResult bl_amd_context_autotune(BLAmdContext*, uint32_t, BLAmdTune* out) { if (out) std::memset(out, 0, sizeof(*out)); return 0; }
BLAmdResult bl_amd_batch_autotune(BLAmdBatch*, uint32_t, BLAmdTune* out) { if (out) std::memset(out, 0, sizeof(*out)); return 0; }
size_t bl_amd_autotune_log(char* buf, size_t cap) { if (buf && cap) buf[0] = 0; return 0; }
BLAmdResult bl_amd_autotune_reset(void) { return 0; }
BLAmdResult bl_amd_set_graph_replay(uint32_t) { return 0; }

}  // extern "C"

#include "../../blend2d_amd/csrc/scene_player.cpp"
