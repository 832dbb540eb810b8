// Host frontend of the MI355X fill path: the BLContext state machine
// (raster/rastercontext.cpp), command recording and the flush that runs the
// three kernel stages.  Exported through the C ABI in include/blend2d_amd.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <new>
#include <string>
#include <vector>

#include "../../include/blend2d_amd.h"
#include "device_types.h"
#include "cost_model.h"
#include "kernels.h"
#include "style.h"

using namespace b2amd;

namespace {

constexpr uint32_t kErrOOM = BL_AMD_ERROR_OUT_OF_MEMORY;
constexpr uint32_t kErrInvalid = BL_AMD_ERROR_INVALID_VALUE;
constexpr uint32_t kErrState = BL_AMD_ERROR_INVALID_STATE;
constexpr uint32_t kErrNotImpl = BL_AMD_ERROR_NOT_IMPLEMENTED;
constexpr uint32_t kErrDevice = BL_AMD_ERROR_DEVICE;
constexpr uint32_t kErrTooLarge = BL_AMD_ERROR_DATA_TOO_LARGE;

// A failing HIP call is reported once on stderr (the BLResult alone does not
// say which call failed or why) and turned into BL_AMD_ERROR_DEVICE.
void report_hip(const char* what, hipError_t e, int line) {
  static int reported = 0;
  if (reported++ < 4) fprintf(stderr, "blend2d_amd: %s failed at context.cpp:%d: %s\n", what, line, hipGetErrorString(e));
}
#define HIP_OK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) { report_hip(#expr, e_, __LINE__); return kErrDevice; } } while (0)

// Flattening tolerance: approximation_options().flatten_tolerance (0.2,
// core/path_p.h:89-98) in 24.8 units, squared (rastercontext.cpp:208-211).
const double kToleranceSq = (0.2 * 256.0) * (0.2 * 256.0);

// Band height the reference picks for a canvas width with thread_count == 0
// (rastercontext.cpp:4154-4193); the conic fetcher's row origin depends on it.
uint32_t ref_band_height(uint32_t width) {
  uint32_t bh = 64;
  size_t pixels = (size_t)width * bh;
  do {
    if (pixels * 4 <= 1024 * 256) break;
    bh >>= 1;
    pixels >>= 1;
  } while (bh > 8);
  return bh;
}

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  ~DevBuf() { if (ptr) (void)hipFree(ptr); }
  bool reserve(size_t bytes) {
    if (bytes <= cap) return true;
    if (ptr) { (void)hipFree(ptr); ptr = nullptr; cap = 0; }
    size_t n = std::max<size_t>(bytes + bytes / 4, 4096);
    if (hipMalloc(&ptr, n) != hipSuccess) { ptr = nullptr; return false; }
    cap = n;
    return true;
  }
  template<typename T> T* as() const { return static_cast<T*>(ptr); }
};

// Host-recorded commands of one flush, possibly merged from many contexts.
struct FrameHost {
  std::vector<DevCanvas> canvases;
  std::vector<DevCmd> cmds;
  std::vector<CmdXform> xforms;
  std::vector<SrcSeg> src;
  std::vector<double> vtx;
  std::vector<DevStyle> styles;
  // gradient tables, filled on the device from their stops at upload
  std::vector<lut::Job> lut_jobs;
  std::vector<Stop> lut_stops;
  uint64_t lut_entries = 0;
  bool dither = false;   // some style samples a dithered gradient
  // upper bound of the edge work items of `src`: 1 per line, 3 per quad, 8 per cubic
  uint64_t items = 0;
  std::vector<DevImage> images;
  // Host only: per command, the rendering call (counted from begin()) that
  // recorded it, so errors found on the device name their call.
  std::vector<uint64_t> cmd_call;
  void clear() {
    canvases.clear(); cmds.clear(); xforms.clear(); src.clear(); vtx.clear();
    styles.clear(); lut_jobs.clear(); lut_stops.clear(); lut_entries = 0; dither = false; items = 0; images.clear();
    cmd_call.clear();
  }
};

struct Stats {
  uint64_t cmds = 0, src = 0, segs = 0, binned = 0, tiles = 0, pixels = 0;
  uint64_t frame_bytes = 0, lut_bytes = 0, image_bytes = 0;
  uint64_t pred_segs = 0, pred_bins = 0;
  uint32_t sizing = 0, fallback = 0, syncs = 0;
  uint32_t render_kernels = 0;   // bit 0 k_render, 1 k_render_box_solid, 2 k_render_box, 3 k_render_box_fast
  double render_ms = 0.0, total_ms = 0.0;
  double merge_ms = 0.0, upload_ms = 0.0, pack_ms = 0.0;
  uint32_t schedule = BL_AMD_SCHEDULE_GRID, queue_wgs = 0;
  uint32_t graph = 0;   // 1: a replay launched as a captured HIP graph
};

// How a frame sizes its segment and bin buffers.
enum Sizing : uint32_t {
  kSizeExact = 0,     // counting pass of the edge builder + host read-back
  kSizeLearned = 1,   // the cost model's predictions (pool mode)
  kSizeHistory = 2,   // the previous frame's exact totals (pool mode)
};

// Process-wide sizing mode and cost-model parameters.
uint32_t g_sizing = kSizeExact;
std::vector<float> g_model;
uint64_t g_model_version = 0;

// Process-wide tile schedule, tile cost model and the autotuner's table
// (frame shape -> schedule) with its log.
uint32_t g_schedule = BL_AMD_SCHEDULE_AUTO, g_queue_wgs = 4;
std::vector<float> g_tile_model;
uint64_t g_tile_model_version = 0;
struct TunedSchedule { uint32_t schedule, wgs; };
// Replays run as captured HIP graphs (bl_amd_set_graph_replay).
uint32_t g_graph_replay = 0;
std::map<uint64_t, TunedSchedule> g_tuned;
std::string g_tune_log;

uint32_t ilog2_u64(uint64_t v) { uint32_t k = 0; while (v > 1) { v >>= 1; k++; } return k; }

// The device's work queue, shared by every context of the process: one HIP
// stream, the timing events and a pinned read-back area.  Frames of
// different contexts are executed one after another (the mutex also makes
// flushes from different threads safe).  Created on first use and kept
// until the process ends.
struct DeviceQueue {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t* host_readback = nullptr;
  uint32_t cus = 0;   // compute units of the device (persistent grids)
  // Pinned staging memory of frame uploads (grown on demand, kept).
  uint8_t* staging = nullptr;
  size_t staging_cap = 0;
  bool ready = false;
  bool staging_reserve(size_t bytes) {
    if (bytes <= staging_cap) return true;
    if (staging) { (void)hipHostFree(staging); staging = nullptr; staging_cap = 0; }
    const size_t n = std::max<size_t>(bytes + bytes / 4, (size_t)1 << 20);
    if (hipHostMalloc((void**)&staging, n, hipHostMallocDefault) != hipSuccess) { staging = nullptr; return false; }
    staging_cap = n;
    return true;
  }
  bool init() {
    if (ready) return true;
    if (!stream && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) { stream = nullptr; return false; }
    if (!ev0 && hipEventCreate(&ev0) != hipSuccess) { ev0 = nullptr; return false; }
    if (!ev1 && hipEventCreate(&ev1) != hipSuccess) { ev1 = nullptr; return false; }
    if (!host_readback && hipHostMalloc((void**)&host_readback, 256, hipHostMallocDefault) != hipSuccess) { host_readback = nullptr; return false; }
    int dev = 0, n_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
      return false;
    cus = (uint32_t)n_cu;
    ready = true;
    return true;
  }
};

DeviceQueue& device_queue() {
  static DeviceQueue* q = new DeviceQueue();
  return *q;
}

// Slices of one caller-owned device region, 256-byte aligned, handed out in
// order; used instead of the runner's grow-only buffers by the execution
// that allocates nothing.
struct ScratchArena {
  uint8_t* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > cap || bytes > cap - at) return nullptr;
    used = at + bytes;
    return base + at;
  }
};

// Device-side state of the pipeline: the resident copy of the last uploaded
// frame and the scratch of its stages.  A frame stays resident until the next
// upload, so it can be executed again (replayed) with no host-side work.
struct Runner {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf in;
  DevBuf lut_buf;   // gradient tables of the resident frame
  DevBuf src_count, src_off, bounds, segs, seg_cmd, mat_cnt, mat_off, mat_pres, list_off, list, binned, cursor, scan_tmp, counters, cmd_tiles;
  DevBuf tile_hdr, group_hdr, tile_px, list_style, band_buf, item_cnt, item_off, items;
  DevBuf cm_params, cm_acc, cm_feat, cm_pred, cm_act;
  DevBuf cmd_err;
  // Tile work queue: queue, bucket counters + head, keys, features,
  // predictions, logged ticks; tile cost-model parameters.
  DevBuf q_queue, q_misc, q_key, q_feat, q_pred, q_ticks, tm_params;
  uint64_t tm_version = 0;
  uint32_t cus = 0;
  // Schedule forced by the autotuner (BL_AMD_SCHEDULE_AUTO: none).
  uint32_t force_schedule = BL_AMD_SCHEDULE_AUTO, force_wgs = 0;
  // Tile cost logging: per analytic tile, features and targets.
  bool collect_tiles = false;
  std::vector<float> tile_x, tile_t;

  // Upload generation: bumped by every upload (a new resident frame).
  uint64_t gen = 0;
  // Counts of the last exact execution of the resident frame: what a graph
  // replay bakes in instead of reading them back.
  struct ExactPlan {
    bool valid = false;
    uint64_t gen = 0;
    uint32_t n_segs = 0, some_failed = 0, n_binned = 0, n_list = 0;
    uint32_t n_atiles = 0, n_fgroups = 0, n_sgroups = 0, n_ggroups = 0;
  } plan;
  // The captured replay and what it was captured with (buffers, frame,
  // schedule, tile model): any difference captures it again.
  hipGraphExec_t graph_exec = nullptr;
  FrameScratch graph_s{};
  void* graph_q[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  uint32_t* graph_cursor = nullptr;
  uint64_t graph_gen = ~0ull, graph_model = ~0ull;
  uint32_t graph_sched = ~0u, graph_wgs = ~0u;
  ~Runner() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  }
  // Commands of the last execution that failed on the device: (command
  // index within the frame, kFailInvalidGeometry / kFailTooManySegments bits).
  std::vector<std::pair<uint32_t, uint32_t>> failed;
  uint64_t cm_version = 0;
  uint32_t* host_readback = nullptr;

  // Resident frame.
  bool resident = false;
  FrameDev f{};
  uint64_t n_src = 0;
  uint32_t item_cap = 0;   // upper bound of edge work items: lines + 3 per quad + 8 per cubic
  uint32_t bands_max = 0;
  uint32_t max_bands() const { return bands_max; }
  uint64_t frame_bytes = 0, lut_bytes = 0, image_bytes = 0;
  double pack_ms = 0.0;   // host time the last upload spent packing

  // History sizing: exact totals of the last exactly counted frame.
  uint64_t hist_segs = 0, hist_bins = 0;
  // Cost-model logging: per analytic command, features and log2(1 + exact count).
  bool collect = false;
  std::vector<float> collected_x, collected_t;

  // Binds the process-wide queue (caller holds its mutex).
  bool init() {
    DeviceQueue& q = device_queue();
    if (!q.init()) return false;
    stream = q.stream;
    ev0 = q.ev0;
    ev1 = q.ev1;
    host_readback = q.host_readback;
    cus = q.cus;
    return true;
  }

  // Frame shape the autotuner keys its choices by: tile, command and canvas
  // counts, each by its power of two.
  uint64_t shape_key() const {
    return ((uint64_t)ilog2_u64((uint64_t)f.n_tiles + 1) << 40) | ((uint64_t)ilog2_u64((uint64_t)f.n_cmds + 1) << 20) |
           (uint64_t)ilog2_u64((uint64_t)f.n_canvases + 1);
  }
  // Launches the render stage: the analytic tiles in the schedule in force
  // (`allow_queue` false: GRID), the box tile groups one workgroup each.
  uint32_t render_stage(FrameScratch& s, uint32_t n_a, uint32_t n_f, uint32_t n_s, uint32_t n_g, bool allow_queue,
                        Stats& stats);

  uint32_t upload(const std::vector<FrameHost*>& parts);
  uint32_t execute(Stats& stats);
  // After the frame's last sync: reads the per-command failure bits back
  // when the device flagged any (`any` = the flag word read back).
  uint32_t read_failures(const FrameScratch& s, uint32_t any);
  // `arena` null: the runner's own buffers (grown as needed); otherwise
  // every intermediate buffer is a slice of the arena and nothing is
  // allocated (BL_AMD_ERROR_DATA_TOO_LARGE when the frame does not fit).
  // `known`: a graph replay -- the launches are captured into a HIP graph
  // with the counts of `known` in place of the read-backs (or the graph
  // captured before is launched when nothing changed).
  uint32_t execute_exact(Stats& stats, ScratchArena* arena = nullptr, const ExactPlan* known = nullptr);
  // The schedule the analytic tiles of this frame run with.
  void resolve_schedule(uint32_t n_a, bool allow_queue, uint32_t& sched, uint32_t& wgs) const;
  // 0 done, 1 pool overflow (redo exactly), otherwise an error
  uint32_t execute_pool(Stats& stats, uint32_t mode);
  bool model_views(cm::FrameModel& m);
  // `parts`: the recorded frames of one context, or of every context of a
  // batch (one canvas each), packed into one device frame.
  uint32_t run(const std::vector<FrameHost*>& parts, Stats& stats, double merge_ms = 0.0) {
    std::lock_guard<std::mutex> lock(device_queue().mu);
    stats = Stats{};
    const auto t0 = std::chrono::steady_clock::now();
    uint32_t r = upload(parts);
    const double up = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (r) return r;
    r = execute(stats);
    stats.merge_ms = merge_ms;
    stats.upload_ms = up;
    stats.pack_ms = pack_ms;
    return r;
  }
  uint32_t replay(Stats& stats) {
    std::lock_guard<std::mutex> lock(device_queue().mu);
    return execute(stats);
  }
  // The same with exact sizing and every intermediate buffer in `arena`.
  uint32_t replay_in(Stats& stats, ScratchArena& arena) {
    std::lock_guard<std::mutex> lock(device_queue().mu);
    stats = Stats{};
    if (!resident) return BL_AMD_SUCCESS;
    return execute_exact(stats, &arena);
  }
};

// Section sizes of a packed frame: per part (context) and in total.
struct PartBase {
  size_t cmds, xforms, src, vtx, styles, luts, images, jobs, stops;   // luts: table entries
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Sections of a packed frame, in layout order.
enum Section : uint32_t { kSecCanvases, kSecCmds, kSecXforms, kSecSrc, kSecVtx, kSecStyles, kSecJobs, kSecImages,
                          kSecStops, kSections };

// Elements [lo, hi) of section `sec` of part `k` into the packed staging
// layout, rebasing their indices (canvas, command, vertex, style, LUT, stop
// and image bases).
static void pack_range(uint8_t* st, const FrameHost& fh, uint32_t k, const PartBase& b, const size_t off[kSections],
                       uint32_t sec, size_t lo, size_t hi) {
  switch (sec) {
    case kSecCmds: {
      DevCmd* out = (DevCmd*)(st + off[kSecCmds]) + b.cmds;
      for (size_t i = lo; i < hi; i++) {
        DevCmd c = fh.cmds[i];
        c.style += (uint32_t)b.styles;
        if (c.kind == kCmdBoxM) c.box_mask[0] += (uint32_t)b.images;
        c.canvas = k;
        out[i] = c;
      }
      break;
    }
    case kSecXforms:
      memcpy((CmdXform*)(st + off[kSecXforms]) + b.xforms + lo, fh.xforms.data() + lo, (hi - lo) * sizeof(CmdXform));
      break;
    case kSecSrc: {
      SrcSeg* out = (SrcSeg*)(st + off[kSecSrc]) + b.src;
      const uint32_t vb = (uint32_t)(b.vtx / 2), cb = (uint32_t)b.cmds;
      for (size_t i = lo; i < hi; i++) {
        SrcSeg g = fh.src[i];
        g.cmd += cb;
        g.i0 += vb;
        g.i1 += vb;
        out[i] = g;
      }
      break;
    }
    case kSecVtx:
      memcpy((double*)(st + off[kSecVtx]) + b.vtx + lo, fh.vtx.data() + lo, (hi - lo) * sizeof(double));
      break;
    case kSecStyles: {
      DevStyle* out = (DevStyle*)(st + off[kSecStyles]) + b.styles;
      for (size_t i = lo; i < hi; i++) {
        DevStyle d = fh.styles[i];
        if (d.kind != kFetchSolid) d.lut += (uint32_t)b.luts;
        d.img += (uint32_t)b.images;
        out[i] = d;
      }
      break;
    }
    case kSecJobs: {
      lut::Job* out = (lut::Job*)(st + off[kSecJobs]) + b.jobs;
      for (size_t i = lo; i < hi; i++) {
        lut::Job j = fh.lut_jobs[i];
        j.lut += (uint32_t)b.luts;
        j.stop0 += (uint32_t)b.stops;
        out[i] = j;
      }
      break;
    }
    case kSecImages:
      memcpy((DevImage*)(st + off[kSecImages]) + b.images + lo, fh.images.data() + lo, (hi - lo) * sizeof(DevImage));
      break;
    case kSecStops:
      memcpy((Stop*)(st + off[kSecStops]) + b.stops + lo, fh.lut_stops.data() + lo, (hi - lo) * sizeof(Stop));
      break;
    default:
      break;
  }
}

static size_t section_base(const PartBase& b, uint32_t sec) {
  switch (sec) {
    case kSecCmds: return b.cmds;
    case kSecXforms: return b.xforms;
    case kSecSrc: return b.src;
    case kSecVtx: return b.vtx;
    case kSecStyles: return b.styles;
    case kSecJobs: return b.jobs;
    case kSecImages: return b.images;
    case kSecStops: return b.stops;
    default: return 0;
  }
}

static size_t section_elem(uint32_t sec) {
  switch (sec) {
    case kSecCanvases: return sizeof(DevCanvas);
    case kSecCmds: return sizeof(DevCmd);
    case kSecXforms: return sizeof(CmdXform);
    case kSecSrc: return sizeof(SrcSeg);
    case kSecVtx: return sizeof(double);
    case kSecStyles: return sizeof(DevStyle);
    case kSecJobs: return sizeof(lut::Job);
    case kSecImages: return sizeof(DevImage);
    case kSecStops: return sizeof(Stop);
    default: return 0;
  }
}

static size_t section_count(const FrameHost& fh, uint32_t sec) {
  switch (sec) {
    case kSecCmds: return fh.cmds.size();
    case kSecXforms: return fh.xforms.size();
    case kSecSrc: return fh.src.size();
    case kSecVtx: return fh.vtx.size();
    case kSecStyles: return fh.styles.size();
    case kSecJobs: return fh.lut_jobs.size();
    case kSecImages: return fh.images.size();
    case kSecStops: return fh.lut_stops.size();
    default: return 0;
  }
}

// Lays out tiles, bands and the (band, command) matrix of every canvas, packs
// the parts into pinned staging memory (several host threads for a large
// batch) and copies the frame to the device in one transfer.
uint32_t Runner::upload(const std::vector<FrameHost*>& parts) {
  resident = false;
  pack_ms = 0.0;
  const auto t_start = std::chrono::steady_clock::now();
  size_t n_cmds_total = 0;
  for (const FrameHost* p : parts) n_cmds_total += p->cmds.size();
  if (n_cmds_total == 0) return BL_AMD_SUCCESS;
  if (!init()) return kErrDevice;

  // canvases (one per part) and the per-part section bases
  std::vector<DevCanvas> canvases;
  canvases.reserve(parts.size());
  std::vector<PartBase> bases(parts.size());
  PartBase tot{0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t n_src_total = 0;
  uint64_t n_items_max = 0;
  bool dither = false;
  for (size_t k = 0; k < parts.size(); k++) {
    const FrameHost& fh = *parts[k];
    dither |= fh.dither;
    bases[k] = tot;
    DevCanvas cv = fh.canvases[0];
    cv.cmd_base = (uint32_t)tot.cmds;
    canvases.push_back(cv);
    tot.cmds += fh.cmds.size();
    tot.xforms += fh.xforms.size();
    tot.src += fh.src.size();
    tot.vtx += fh.vtx.size();
    tot.styles += fh.styles.size();
    tot.luts += fh.lut_entries;
    tot.jobs += fh.lut_jobs.size();
    tot.stops += fh.lut_stops.size();
    tot.images += fh.images.size();
    n_src_total += fh.src.size();
    n_items_max += fh.items;
  }

  uint64_t n_mat = 0;
  uint32_t n_tiles = 0, n_groups = 0, band_base = 0;
  for (DevCanvas& c : canvases) {
    c.tiles_x = (c.w + kTileW - 1) / kTileW;
    c.bands = (c.h + kTileH - 1) / kTileH;
    c.groups_x = (c.tiles_x + kGroupTiles - 1) / kGroupTiles;
    c.ref_band_h = ref_band_height(c.w);
    c.band_base = band_base;
    c.tile_base = n_tiles;
    c.group_base = n_groups;
    c.mat_base = n_mat;
    band_base += c.bands;
    n_tiles += c.tiles_x * c.bands;
    n_groups += c.groups_x * c.bands;
    n_mat += (uint64_t)c.bands * c.n_cmds;
  }
  if (n_mat >= 0xFFFFFFF0ull || n_src_total >= 0xFFFFFFF0ull || n_items_max >= 0xFFFFFFF0ull || tot.vtx / 2 >= 0xFFFFFFF0ull ||
      tot.styles >= 0xFFFFFFF0ull || tot.luts >= 0xFFFFFFF0ull || tot.stops >= 0xFFFFFFF0ull)
    return kErrOOM;

  // section offsets
  size_t off[kSections];
  size_t at = 0;
  const size_t sizes[kSections] = {canvases.size() * sizeof(DevCanvas), tot.cmds * sizeof(DevCmd), tot.xforms * sizeof(CmdXform),
                           tot.src * sizeof(SrcSeg), tot.vtx * sizeof(double), tot.styles * sizeof(DevStyle),
                           tot.jobs * sizeof(lut::Job), tot.images * sizeof(DevImage), tot.stops * sizeof(Stop)};
  for (uint32_t i = 0; i < kSections; i++) { off[i] = at; at = align256(at + sizes[i]); }
  const size_t blob_size = at;

  DeviceQueue& q = device_queue();
  if (!q.staging_reserve(blob_size)) return kErrOOM;
  uint8_t* st = q.staging;
  memcpy(st + off[kSecCanvases], canvases.data(), sizes[kSecCanvases]);
  if (!in.reserve(blob_size + 256)) return kErrOOM;
  if (!lut_buf.reserve(tot.luts * 4 + 256)) return kErrOOM;

  // Packing tasks: every section of every part, in chunks of at most kChunk
  // elements, in the order of their place in the packed layout.  They are
  // grouped into batches of about kBatch bytes of that layout; as soon as
  // every task of the next batch is done its bytes are copied to the device,
  // so the copies overlap the packing of the batches after them.
  struct Task { uint32_t part, sec; size_t lo, hi, dst; uint32_t batch; };
  constexpr size_t kChunk = (size_t)1 << 18;
  constexpr size_t kBatch = (size_t)32 << 20;
  std::vector<Task> tasks;
  for (size_t k = 0; k < parts.size(); k++)
    for (uint32_t sec = kSecCmds; sec < kSections; sec++) {
      const size_t n = section_count(*parts[k], sec);
      const size_t at0 = off[sec] + section_base(bases[k], sec) * section_elem(sec);
      for (size_t lo = 0; lo < n; lo += kChunk)
        tasks.push_back(Task{(uint32_t)k, sec, lo, std::min(n, lo + kChunk), at0 + lo * section_elem(sec), 0});
    }
  std::sort(tasks.begin(), tasks.end(), [](const Task& x, const Task& y) { return x.dst < y.dst; });
  std::vector<size_t> batch_start{0};   // layout byte where each batch starts
  std::vector<uint32_t> batch_tasks{0};
  for (Task& t : tasks) {
    if (t.dst - batch_start.back() >= kBatch) {
      batch_start.push_back(t.dst);
      batch_tasks.push_back(0);
    }
    t.batch = (uint32_t)(batch_start.size() - 1);
    batch_tasks.back()++;
  }
  const size_t n_batches = batch_start.size();
  std::unique_ptr<std::atomic<uint32_t>[]> remaining(new std::atomic<uint32_t>[n_batches]);
  for (size_t b = 0; b < n_batches; b++) remaining[b].store(batch_tasks[b], std::memory_order_relaxed);
  size_t issued = 0;
  hipError_t copy_err = hipSuccess;
  // (main thread only) copies every finished batch not yet copied, in order
  auto issue_ready = [&](bool all) {
    while (issued < n_batches && (all || remaining[issued].load(std::memory_order_acquire) == 0)) {
      const size_t lo = batch_start[issued], hi = issued + 1 < n_batches ? batch_start[issued + 1] : blob_size;
      if (copy_err == hipSuccess)
        copy_err = hipMemcpyAsync(in.as<uint8_t>() + lo, st + lo, hi - lo, hipMemcpyHostToDevice, stream);
      issued++;
    }
  };
  auto run_task = [&](const Task& t) {
    pack_range(st, *parts[t.part], t.part, bases[t.part], off, t.sec, t.lo, t.hi);
    remaining[t.batch].fetch_sub(1, std::memory_order_release);
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t n_threads = std::min<size_t>({tasks.size(), (size_t)(hw ? hw : 1), (size_t)16,
                                            blob_size / ((size_t)8 << 20) + 1});
  std::atomic<size_t> next{0};
  auto worker = [&](bool main) {
    for (size_t i = next.fetch_add(1, std::memory_order_relaxed); i < tasks.size();
         i = next.fetch_add(1, std::memory_order_relaxed)) {
      run_task(tasks[i]);
      if (main) issue_ready(false);
    }
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < n_threads; t++) pool.emplace_back(worker, false);
  worker(true);
  for (std::thread& th : pool) th.join();
  pack_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  issue_ready(true);
  if (copy_err != hipSuccess) {
    report_hip("hipMemcpyAsync", copy_err, __LINE__);
    (void)hipStreamSynchronize(stream);   // copies already issued read the staging memory
    return kErrDevice;
  }

  // gradient tables from their stops, on the device
  if (tot.jobs)
    launch_gradient_luts((const lut::Job*)(in.as<uint8_t>() + off[kSecJobs]), (const Stop*)(in.as<uint8_t>() + off[kSecStops]),
                         (uint32_t)tot.jobs, lut_buf.as<uint32_t>(), stream);
  // the staging memory is reused by the next upload
  HIP_OK(hipStreamSynchronize(stream));

  const uint8_t* base = in.as<uint8_t>();
  f = FrameDev{};
  f.canvases = (const DevCanvas*)(base + off[kSecCanvases]); f.n_canvases = (uint32_t)canvases.size();
  f.cmds = (const DevCmd*)(base + off[kSecCmds]); f.n_cmds = (uint32_t)tot.cmds;
  f.xforms = (const CmdXform*)(base + off[kSecXforms]);
  f.src = (const SrcSeg*)(base + off[kSecSrc]); f.n_src = (uint32_t)tot.src;
  f.vtx = (const double*)(base + off[kSecVtx]);
  f.styles = (const DevStyle*)(base + off[kSecStyles]);
  f.luts = lut_buf.as<uint32_t>();
  f.dither = dither ? 1u : 0u;
  f.images = (const DevImage*)(base + off[kSecImages]);
  f.n_tiles = n_tiles;
  f.n_groups = n_groups;
  f.n_bands = band_base;
  f.n_mat = n_mat;
  f.tol_sq = kToleranceSq;
  n_src = n_src_total;
  item_cap = (uint32_t)n_items_max;
  bands_max = 0;
  for (const DevCanvas& c : canvases) bands_max = std::max(bands_max, c.bands);
  lut_bytes = tot.luts * 4;
  frame_bytes = blob_size + lut_bytes;
  image_bytes = 0;
  for (const FrameHost* p : parts)
    for (const DevImage& im : p->images) image_bytes += (uint64_t)im.w * im.h * (im.format == kFmtA8 ? 1u : 4u);
  resident = true;
  gen++;
  plan.valid = false;
  return BL_AMD_SUCCESS;
}

// Device buffers of the cost model for the resident frame.
bool Runner::model_views(cm::FrameModel& m) {
  const size_t nc = std::max<uint32_t>(f.n_cmds, 1);
  if (!cm_acc.reserve(nc * sizeof(cm::CmdAcc)) || !cm_feat.reserve(nc * cm::kIn * 4) || !cm_pred.reserve(nc * cm::kOut * 4) ||
      !cm_act.reserve(nc * 8 + 64) || !cm_params.reserve(cm::kParams * 4))
    return false;
  if (!g_model.empty() && cm_version != g_model_version) {
    if (hipMemcpyAsync(cm_params.ptr, g_model.data(), cm::kParams * 4, hipMemcpyHostToDevice, stream) != hipSuccess) return false;
    if (hipStreamSynchronize(stream) != hipSuccess) return false;
    cm_version = g_model_version;
  }
  m.params = cm_params.as<float>();
  m.acc = cm_acc.as<cm::CmdAcc>();
  m.features = cm_feat.as<float>();
  m.pred = cm_pred.as<float>();
  m.actual_segs = cm_act.as<uint32_t>();
  m.actual_bins = cm_act.as<uint32_t>() + nc;
  m.totals = (unsigned long long*)(cm_act.as<uint8_t>() + nc * 8);
  return true;
}

uint32_t Runner::read_failures(const FrameScratch& s, uint32_t any) {
  failed.clear();
  if (!any) return BL_AMD_SUCCESS;
  std::vector<uint32_t> bits(f.n_cmds);
  HIP_OK(hipMemcpyAsync(bits.data(), s.cmd_err, (size_t)f.n_cmds * 4, hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  for (uint32_t c = 0; c < f.n_cmds; c++)
    if (bits[c]) failed.emplace_back(c, bits[c]);
  return BL_AMD_SUCCESS;
}

uint32_t Runner::execute(Stats& stats) {
  stats = Stats{};
  failed.clear();
  if (!resident) return BL_AMD_SUCCESS;
  uint32_t mode = g_sizing;
  if (collect || collect_tiles || (mode == kSizeLearned && g_model.empty())) mode = kSizeExact;
  // Graph replays: the first execution of a resident frame runs exactly and
  // records its counts; later ones launch the captured graph.
  if (g_graph_replay && !collect && !collect_tiles) {
    if (plan.valid && plan.gen == gen) return execute_exact(stats, nullptr, &plan);
    return execute_exact(stats);
  }
  if (mode != kSizeExact) {
    uint32_t r = execute_pool(stats, mode);
    if (r != 1) return r;
    // pool overflow: the frame was not composited; redo it exactly
    Stats exact;
    r = execute_exact(exact);
    exact.sizing = mode;
    exact.fallback = 1;
    exact.syncs += stats.syncs;
    exact.pred_segs = stats.pred_segs;
    exact.pred_bins = stats.pred_bins;
    exact.total_ms += stats.total_ms;
    stats = exact;
    return r;
  }
  return execute_exact(stats);
}

// Pool mode: segment and bin buffers are sized ahead of time (from the cost
// model's predictions or from the last exact frame) and the edge builder
// runs once, without the counting pass and without host round trips in
// between; any overflow is detected at the end of the frame.
uint32_t Runner::execute_pool(Stats& stats, uint32_t mode) {
  auto t0 = std::chrono::steady_clock::now();
  const uint64_t n_mat = f.n_mat;
  size_t scan_words = scan_tmp_words(std::max<size_t>({(size_t)n_src, (size_t)n_mat, (size_t)f.n_bands}) + 1);
  if (!item_cnt.reserve((n_src + 1) * 4) || !item_off.reserve((n_src + 1) * 4) ||
      !items.reserve(std::max<size_t>(item_cap, 1) * sizeof(EdgeItem)) ||
      !bounds.reserve((size_t)f.n_cmds * sizeof(CmdBounds)) || !mat_cnt.reserve((n_mat + 1) * 4) ||
      !mat_off.reserve((n_mat + 1) * 4) || !mat_pres.reserve((n_mat + 1) * 4) || !list_off.reserve((n_mat + 1) * 4) ||
      !list.reserve((n_mat + 1) * 16) || !list_style.reserve((n_mat + 1) * 4) || !cursor.reserve((n_mat + 1) * 4) ||
      !scan_tmp.reserve(scan_words * 4) ||
      !counters.reserve(64) || !cmd_tiles.reserve((size_t)f.n_cmds * 4) ||
      !tile_hdr.reserve((size_t)f.n_tiles * 16) || !group_hdr.reserve((size_t)f.n_groups * 32) ||
      !band_buf.reserve(((size_t)f.n_bands + 1) * 40) || !tile_px.reserve((size_t)f.n_tiles * 8) ||
      !cmd_err.reserve((size_t)f.n_cmds * 4))
    return kErrOOM;

  FrameScratch s;
  memset(&s, 0, sizeof(s));
  s.tile_hdr = tile_hdr.as<uint4>();
  s.group_hdr = group_hdr.as<uint4>();
  s.band_nonsolid = band_buf.as<uint32_t>();
  s.band_tiles = s.band_nonsolid + (f.n_bands + 1);
  s.band_sgroups = s.band_tiles + (f.n_bands + 1);
  s.band_ggroups = s.band_sgroups + (f.n_bands + 1);
  s.tile_off = s.band_ggroups + (f.n_bands + 1);
  s.sgroup_off = s.tile_off + (f.n_bands + 1);
  s.ggroup_off = s.sgroup_off + (f.n_bands + 1);
  s.band_slow = s.ggroup_off + (f.n_bands + 1);
  s.band_fgroups = s.band_slow + (f.n_bands + 1);
  s.fgroup_off = s.band_fgroups + (f.n_bands + 1);
  s.tile_px = tile_px.as<unsigned long long>();
  s.item_cnt = item_cnt.as<uint32_t>();
  s.item_off = item_off.as<uint32_t>();
  s.items = items.as<EdgeItem>();
  s.cmd_err = cmd_err.as<uint32_t>();
  s.err_any = counters.as<uint32_t>() + 4;
  s.bounds = bounds.as<CmdBounds>();
  s.mat_cnt = mat_cnt.as<uint32_t>();
  s.mat_off = mat_off.as<uint32_t>();
  s.mat_pres = mat_pres.as<uint32_t>();
  s.list_off = list_off.as<uint32_t>();
  s.list_ent = list.as<uint4>();
  s.list_style = list_style.as<uint32_t>();
  s.scan_tmp = scan_tmp.as<uint32_t>();
  s.pixels = counters.as<unsigned long long>();
  s.seg_cursor = counters.as<uint32_t>() + 2;
  s.flags = counters.as<uint32_t>() + 3;
  s.cmd_tiles = cmd_tiles.as<uint32_t>();

  launch_fill_u32(counters.as<uint32_t>(), 0u, 5, stream);
  launch_fill_u32(s.cmd_err, 0u, f.n_cmds, stream);
  launch_init_bounds(s.bounds, f.n_cmds, stream);

  launch_edge_items(f, s, stream);

  // Pool sizes: the emitter writes segments densely (exact slots per item).
  uint64_t want_segs, want_bins;
  if (mode == kSizeLearned) {
    cm::FrameModel m;
    if (!model_views(m)) return kErrOOM;
    cm::launch_features(f, m, stream);
    cm::launch_infer_frame(f, m, stream);
    HIP_OK(hipMemcpyAsync(host_readback + 4, m.totals, 16, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    stats.syncs++;
    uint64_t tot[2];
    memcpy(tot, host_readback + 4, 16);
    stats.pred_segs = tot[0];
    stats.pred_bins = tot[1];
    want_segs = tot[0] + tot[0] / 4 + 1024;
    want_bins = tot[1] + tot[1] / 4 + 1024;
  }
  else {
    want_segs = hist_segs + hist_segs / 4 + 1024;
    want_bins = hist_bins + hist_bins / 4 + 1024;
  }
  // (the cursor may run past the capacity by up to the slots of the
  // allocations that fail, so the capacity stays below 2^31)
  if (want_segs > 0x7FFFF000ull || want_bins > 0xFFFFF000ull) return 1;
  if (!segs.reserve(want_segs * sizeof(DevSeg)) || !seg_cmd.reserve(want_segs * 4) || !binned.reserve(want_bins * sizeof(DevSeg)))
    return kErrOOM;
  const uint32_t cap_segs = (uint32_t)std::min<size_t>(std::min(segs.cap / sizeof(DevSeg), seg_cmd.cap / 4), 0x7FFFF000ull);
  const uint32_t cap_bins = (uint32_t)std::min<size_t>(binned.cap / sizeof(DevSeg), 0xFFFFF000ull);
  s.segs = segs.as<DevSeg>();
  s.seg_cmd = seg_cmd.as<uint32_t>();
  s.binned = binned.as<DevSeg>();
  s.binned_cap = cap_bins;

  launch_edges_pool(f, s, cap_segs, item_cap, stream);
#ifdef B2AMD_EMU_TRACE
  fprintf(stderr, "pool: cap_segs %u cursor %u flags %u want %llu\n", cap_segs, counters.as<uint32_t>()[2], counters.as<uint32_t>()[3], (unsigned long long)want_segs);
#endif
  launch_cmd_tiles(f, s, stream);
  launch_fill_u32(s.mat_cnt, 0u, n_mat + 1, stream);
  launch_bin_count(f, s, cap_segs, stream);
  launch_presence(f, s, stream);
  scan_exclusive(s.mat_cnt, s.mat_off, n_mat, s.scan_tmp, stream);
  scan_exclusive(s.mat_pres, s.list_off, n_mat, s.scan_tmp, stream);
  launch_band_class(f, s, stream);
  HIP_OK(hipMemcpyAsync(cursor.ptr, s.mat_off, (n_mat + 1) * 4, hipMemcpyDeviceToDevice, stream));
  launch_bin_scatter(f, s, cap_segs, cursor.as<uint32_t>(), stream);
  launch_list_build(f, s, stream);

  // One read-back before the render stage: a pool overflow ends the frame
  // here (nothing composited; the caller redoes it exactly), otherwise the
  // tile and tile-group counts size the render grids exactly (upper-bound
  // grids of a frame with a million tiles spend milliseconds launching
  // workgroups that exit at once).
  uint32_t n_atiles = 0, n_fgroups = 0, n_sgroups = 0, n_ggroups = 0;
  HIP_OK(hipMemcpyAsync(host_readback + 3, s.flags, 4, hipMemcpyDeviceToHost, stream));
  if (f.n_tiles) {
    HIP_OK(hipMemcpyAsync(host_readback + 12, s.tile_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 13, s.sgroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 14, s.ggroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 15, s.fgroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
  }
  HIP_OK(hipStreamSynchronize(stream));
  stats.syncs++;
  if (host_readback[3]) {
    stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats.sizing = mode;
    return 1;
  }
  if (f.n_tiles) {
    n_atiles = host_readback[12];
    n_sgroups = host_readback[13];
    n_ggroups = host_readback[14];
    n_fgroups = host_readback[15];
    if (n_atiles > f.n_tiles || (uint64_t)n_fgroups + n_sgroups + n_ggroups > f.n_groups) return kErrDevice;
  }

  launch_render_setup(f, s, stream);
  HIP_OK(hipEventRecord(ev0, stream));
  if (uint32_t r = render_stage(s, n_atiles, n_fgroups, n_sgroups, n_ggroups, true, stats)) return r;
  stats.render_kernels = (n_atiles ? 1u : 0u) | (n_sgroups ? 2u : 0u) | (n_ggroups ? 4u : 0u) | (n_fgroups ? 8u : 0u);
  HIP_OK(hipEventRecord(ev1, stream));
  launch_render_stats(f, s, stream);
  HIP_OK(hipMemcpyAsync(host_readback, counters.ptr, 20, hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(host_readback + 8, s.mat_off + n_mat, 4, hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  stats.syncs++;
  if (hipError_t e_ = hipGetLastError(); e_ != hipSuccess) { report_hip("kernel launch", e_, __LINE__); return kErrDevice; }
  auto t1 = std::chrono::steady_clock::now();
  stats.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  stats.sizing = mode;
  if (host_readback[3]) return 1;
  if (uint32_t r = read_failures(s, host_readback[4])) return r;

  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  uint64_t pixels;
  memcpy(&pixels, host_readback, 8);
  stats.cmds = f.n_cmds;
  stats.src = n_src;
  stats.segs = std::min(host_readback[2], cap_segs);   // slots taken, unused chunk tails included
  stats.binned = host_readback[8];
  stats.tiles = f.n_tiles;
  stats.pixels = pixels;
  stats.frame_bytes = frame_bytes;
  stats.lut_bytes = lut_bytes;
  stats.image_bytes = image_bytes;
  stats.render_ms = ms;
  return BL_AMD_SUCCESS;
}

void Runner::resolve_schedule(uint32_t n_a, bool allow_queue, uint32_t& sched, uint32_t& wgs) const {
  sched = BL_AMD_SCHEDULE_GRID;
  wgs = 0;
  if (allow_queue && n_a) {
    if (collect_tiles) { sched = BL_AMD_SCHEDULE_QUEUE; wgs = 4; }
    else if (force_schedule != BL_AMD_SCHEDULE_AUTO) { sched = force_schedule; wgs = force_wgs; }
    else if (g_schedule != BL_AMD_SCHEDULE_AUTO) { sched = g_schedule; wgs = g_queue_wgs; }
    else {
      auto it = g_tuned.find(shape_key());
      if (it != g_tuned.end()) { sched = it->second.schedule; wgs = it->second.wgs; }
    }
  }
  if (sched == BL_AMD_SCHEDULE_QUEUE_LEARNED && g_tile_model.empty()) sched = BL_AMD_SCHEDULE_QUEUE_HEURISTIC;
  if (wgs < 1 || wgs > 4) wgs = 4;
}

uint32_t Runner::render_stage(FrameScratch& s, uint32_t n_a, uint32_t n_f, uint32_t n_s, uint32_t n_g, bool allow_queue,
                              Stats& stats) {
  uint32_t sched, wgs;
  resolve_schedule(n_a, allow_queue, sched, wgs);
  stats.schedule = sched;
  stats.queue_wgs = sched == BL_AMD_SCHEDULE_GRID ? 0u : wgs;
  if (sched == BL_AMD_SCHEDULE_GRID) {
    launch_render(f, s, n_a, n_f, n_s, n_g, stream);
    return BL_AMD_SUCCESS;
  }
  const bool ordered = sched != BL_AMD_SCHEDULE_QUEUE;
  const bool learned = sched == BL_AMD_SCHEDULE_QUEUE_LEARNED;
  const bool feats = ordered || collect_tiles;
  if (!q_misc.reserve(1024 + 64) || (ordered && (!q_queue.reserve((size_t)n_a * 4) || !q_key.reserve((size_t)n_a * 4))) ||
      (feats && !q_feat.reserve((size_t)n_a * 64)) || (learned && (!q_pred.reserve((size_t)n_a * 8) || !tm_params.reserve(cm::kParams * 4))) ||
      (collect_tiles && !q_ticks.reserve((size_t)n_a * 8)))
    return kErrOOM;
  s.queue_hist = q_misc.as<uint32_t>();
  s.queue_head = q_misc.as<uint32_t>() + 256;
  s.queue = ordered ? q_queue.as<uint32_t>() : nullptr;
  s.tile_key = ordered ? q_key.as<uint32_t>() : nullptr;
  s.tile_feat = feats ? q_feat.as<float>() : nullptr;
  s.tile_cycles = collect_tiles ? q_ticks.as<uint32_t>() : nullptr;
  if (feats) launch_tile_features(f, s, n_a, stream);
  if (learned) {
    if (tm_version != g_tile_model_version) {
      HIP_OK(hipMemcpyAsync(tm_params.ptr, g_tile_model.data(), cm::kParams * 4, hipMemcpyHostToDevice, stream));
      HIP_OK(hipStreamSynchronize(stream));
      tm_version = g_tile_model_version;
    }
    cm::launch_mlp_forward(s.tile_feat, n_a, tm_params.as<float>(), q_pred.as<float>(), stream);
  }
  if (ordered) launch_tile_queue(f, s, n_a, learned ? q_pred.as<float>() : nullptr, stream);
  const uint32_t grid = std::min<uint32_t>(n_a, std::max<uint32_t>(cus, 1) * wgs);
  launch_render_queue(f, s, grid, wgs, stream);
  launch_render(f, s, 0, n_f, n_s, n_g, stream);
  return BL_AMD_SUCCESS;
}

// Exact sizing: counting pass of the edge builder, scan, read-back, emit;
// bin count, scans, read-back, scatter.  Two host round trips read back the
// sizes the next stage allocates (segment count, binned count).
uint32_t Runner::execute_exact(Stats& stats, ScratchArena* arena, const ExactPlan* known) {
  stats = Stats{};
  auto t0 = std::chrono::steady_clock::now();
  const bool graph = known != nullptr && !arena;
  const uint64_t n_mat = f.n_mat;
  const uint32_t kErrFit = arena ? kErrTooLarge : kErrOOM;
  // A device buffer of `bytes`: the runner's own, grown as needed, or the
  // next slice of the arena.
  auto take = [&](DevBuf& own, size_t bytes) -> void* {
    if (arena) return arena->take(bytes);
    return own.reserve(bytes) ? own.ptr : nullptr;
  };

  size_t scan_words = scan_tmp_words(std::max<size_t>({(size_t)n_src, (size_t)item_cap, (size_t)n_mat, (size_t)f.n_bands}) + 1);
  FrameScratch s;
  memset(&s, 0, sizeof(s));
  s.item_cnt = (uint32_t*)take(item_cnt, (n_src + 1) * 4);
  s.item_off = (uint32_t*)take(item_off, (n_src + 1) * 4);
  s.items = (EdgeItem*)take(items, std::max<size_t>(item_cap, 1) * sizeof(EdgeItem));
  s.src_count = (uint32_t*)take(src_count, ((size_t)item_cap + 1) * 4);
  s.src_off = (uint32_t*)take(src_off, ((size_t)item_cap + 1) * 4);
  s.bounds = (CmdBounds*)take(bounds, (size_t)f.n_cmds * sizeof(CmdBounds));
  s.mat_cnt = (uint32_t*)take(mat_cnt, (n_mat + 1) * 4);
  s.mat_off = (uint32_t*)take(mat_off, (n_mat + 1) * 4);
  s.mat_pres = (uint32_t*)take(mat_pres, (n_mat + 1) * 4);
  s.list_off = (uint32_t*)take(list_off, (n_mat + 1) * 4);
  uint32_t* cursor_p = (uint32_t*)take(cursor, (n_mat + 1) * 4);
  s.scan_tmp = (uint32_t*)take(scan_tmp, scan_words * 4);
  uint32_t* counters_p = (uint32_t*)take(counters, 64);
  s.cmd_tiles = (uint32_t*)take(cmd_tiles, (size_t)f.n_cmds * 4);
  s.tile_hdr = (uint4*)take(tile_hdr, (size_t)f.n_tiles * 16);
  s.group_hdr = (uint4*)take(group_hdr, (size_t)f.n_groups * 32);
  s.band_nonsolid = (uint32_t*)take(band_buf, ((size_t)f.n_bands + 1) * 40);
  s.tile_px = (unsigned long long*)take(tile_px, (size_t)f.n_tiles * 8);
  s.cmd_err = (uint32_t*)take(cmd_err, (size_t)f.n_cmds * 4);
  if (!s.item_cnt || !s.item_off || !s.items || !s.src_count || !s.src_off || !s.bounds || !s.mat_cnt || !s.mat_off ||
      !s.mat_pres || !s.list_off || !cursor_p || !s.scan_tmp || !counters_p || !s.cmd_tiles || !s.tile_hdr ||
      !s.group_hdr || !s.band_nonsolid || !s.tile_px || !s.cmd_err)
    return kErrFit;
  s.band_tiles = s.band_nonsolid + (f.n_bands + 1);
  s.band_sgroups = s.band_tiles + (f.n_bands + 1);
  s.band_ggroups = s.band_sgroups + (f.n_bands + 1);
  s.tile_off = s.band_ggroups + (f.n_bands + 1);
  s.sgroup_off = s.tile_off + (f.n_bands + 1);
  s.ggroup_off = s.sgroup_off + (f.n_bands + 1);
  s.band_slow = s.ggroup_off + (f.n_bands + 1);
  s.band_fgroups = s.band_slow + (f.n_bands + 1);
  s.fgroup_off = s.band_fgroups + (f.n_bands + 1);
  s.pixels = (unsigned long long*)counters_p;
  s.err_any = counters_p + 4;

  // Graph replay: every buffer is taken now (the counts are known), the
  // render stage's queue buffers and the tile model are made ready, and the
  // graph captured before is launched if it was captured with the same
  // buffers, frame and schedule; otherwise the launches below are captured.
  uint32_t g_sched = 0, g_wgs = 0;
  void* q_ptrs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  struct CaptureGuard {
    hipStream_t st;
    bool on = false;
    ~CaptureGuard() {
      if (!on) return;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
    }
  } capture{stream};
  auto run_graph = [&]() -> uint32_t {
    HIP_OK(hipEventRecord(ev0, stream));
    HIP_OK(hipGraphLaunch(graph_exec, stream));
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 2, s.pixels, 8, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 4, s.err_any, 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    stats.syncs++;
    if (hipError_t e_ = hipGetLastError(); e_ != hipSuccess) { report_hip("graph launch", e_, __LINE__); return kErrDevice; }
    if (uint32_t r = read_failures(s, host_readback[4])) return r;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    uint64_t pixels;
    memcpy(&pixels, host_readback + 2, 8);
    stats.cmds = f.n_cmds;
    stats.src = n_src;
    stats.segs = known->n_segs;
    stats.binned = known->n_binned;
    stats.tiles = f.n_tiles;
    stats.pixels = pixels;
    stats.frame_bytes = frame_bytes;
    stats.lut_bytes = lut_bytes;
    stats.image_bytes = image_bytes;
    stats.render_ms = ms;   // the whole graph: every device stage
    stats.render_kernels = (known->n_atiles ? 1u : 0u) | (known->n_sgroups ? 2u : 0u) | (known->n_ggroups ? 4u : 0u) |
                           (known->n_fgroups ? 8u : 0u);
    stats.schedule = graph_sched;
    stats.queue_wgs = graph_sched == BL_AMD_SCHEDULE_GRID ? 0u : graph_wgs;
    stats.graph = 1;
    stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BL_AMD_SUCCESS;
  };
  if (graph) {
    s.segs = (DevSeg*)take(segs, (size_t)std::max<uint32_t>(known->n_segs, 1) * sizeof(DevSeg));
    s.seg_cmd = (uint32_t*)take(seg_cmd, (size_t)std::max<uint32_t>(known->n_segs, 1) * 4);
    s.binned = (DevSeg*)take(binned, (size_t)std::max<uint32_t>(known->n_binned, 1) * sizeof(DevSeg));
    s.list_ent = (uint4*)take(list, (size_t)std::max<uint32_t>(known->n_list, 1) * 16);
    s.list_style = (uint32_t*)take(list_style, (size_t)std::max<uint32_t>(known->n_list, 1) * 4);
    if (!s.segs || !s.seg_cmd || !s.binned || !s.list_ent || !s.list_style) return kErrOOM;
    resolve_schedule(known->n_atiles, true, g_sched, g_wgs);
    if (g_sched != BL_AMD_SCHEDULE_GRID) {
      const size_t n_a = known->n_atiles;
      if (!q_misc.reserve(1024 + 64) || !q_queue.reserve(n_a * 4) || !q_key.reserve(n_a * 4) || !q_feat.reserve(n_a * 64) ||
          !q_pred.reserve(n_a * 8) || !tm_params.reserve(cm::kParams * 4))
        return kErrOOM;
      if (g_sched == BL_AMD_SCHEDULE_QUEUE_LEARNED && tm_version != g_tile_model_version) {
        HIP_OK(hipMemcpyAsync(tm_params.ptr, g_tile_model.data(), cm::kParams * 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        tm_version = g_tile_model_version;
      }
      q_ptrs[0] = q_misc.ptr; q_ptrs[1] = q_queue.ptr; q_ptrs[2] = q_key.ptr;
      q_ptrs[3] = q_feat.ptr; q_ptrs[4] = q_pred.ptr; q_ptrs[5] = tm_params.ptr;
    }
    const bool same = graph_exec && graph_gen == gen && graph_sched == g_sched && graph_wgs == g_wgs &&
                      graph_model == g_tile_model_version && graph_cursor == cursor_p &&
                      memcmp(&graph_s, &s, sizeof(s)) == 0 && memcmp(graph_q, q_ptrs, sizeof(q_ptrs)) == 0;
    if (same) return run_graph();
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    graph_s = s;
    graph_cursor = cursor_p;
    memcpy(graph_q, q_ptrs, sizeof(q_ptrs));
    HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    capture.on = true;
  }

  cm::FrameModel model{};
  if (collect && !arena) {
    if (!model_views(model)) return kErrOOM;
    cm::launch_features(f, model, stream);
  }

  // Stage 1: edges (work items; count, scan, emit).
  launch_fill_u32(counters_p, 0u, 5, stream);
  launch_fill_u32(s.cmd_err, 0u, f.n_cmds, stream);
  launch_init_bounds(s.bounds, f.n_cmds, stream);
  launch_edge_items(f, s, stream);
  launch_edges(f, s, true, item_cap, stream);
  launch_cmd_tiles(f, s, stream);
  // A frame of box fills only has no edges: no counting round trip.
  uint32_t n_segs = 0;
  bool some_failed = false;
  if (graph) {
    if (n_src) scan_exclusive(s.src_count, s.src_off, item_cap, s.scan_tmp, stream);
    n_segs = known->n_segs;
    some_failed = known->some_failed != 0;
  }
  else {
    if (n_src) {
      scan_exclusive(s.src_count, s.src_off, item_cap, s.scan_tmp, stream);
      HIP_OK(hipMemcpyAsync(host_readback, s.src_off + item_cap, 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipMemcpyAsync(host_readback + 5, s.err_any, 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipStreamSynchronize(stream));
      stats.syncs++;
      n_segs = host_readback[0];
    }
    // segments of failed commands are not binned
    some_failed = n_src && host_readback[5];
    if (n_segs >= 0x80000000u) return kErrDevice;
    s.segs = (DevSeg*)take(segs, (size_t)std::max<uint32_t>(n_segs, 1) * sizeof(DevSeg));
    s.seg_cmd = (uint32_t*)take(seg_cmd, (size_t)std::max<uint32_t>(n_segs, 1) * 4);
    if (!s.segs || !s.seg_cmd) return kErrFit;
  }
  launch_edges(f, s, false, item_cap, stream);

  // Stage 2: binning.
  launch_fill_u32(s.mat_cnt, 0u, n_mat + 1, stream);
  launch_bin_count(f, s, n_segs, stream);
  launch_presence(f, s, stream);
  scan_exclusive(s.mat_cnt, s.mat_off, n_mat, s.scan_tmp, stream);
  scan_exclusive(s.mat_pres, s.list_off, n_mat, s.scan_tmp, stream);
  launch_band_class(f, s, stream);
  uint32_t n_binned, n_list, n_atiles, n_sgroups, n_ggroups, n_fgroups;
  if (graph) {
    n_binned = known->n_binned;
    n_list = known->n_list;
    n_atiles = known->n_atiles;
    n_sgroups = known->n_sgroups;
    n_ggroups = known->n_ggroups;
    n_fgroups = known->n_fgroups;
  }
  else {
    HIP_OK(hipMemcpyAsync(host_readback, s.mat_off + n_mat, 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(host_readback + 1, s.list_off + n_mat, 4, hipMemcpyDeviceToHost, stream));
    if (f.n_tiles) {
      HIP_OK(hipMemcpyAsync(host_readback + 12, s.tile_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipMemcpyAsync(host_readback + 13, s.sgroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipMemcpyAsync(host_readback + 14, s.ggroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipMemcpyAsync(host_readback + 15, s.fgroup_off + f.n_bands, 4, hipMemcpyDeviceToHost, stream));
    }
    HIP_OK(hipStreamSynchronize(stream));
    stats.syncs++;
    n_binned = host_readback[0];
    n_list = host_readback[1];
    n_atiles = f.n_tiles ? host_readback[12] : 0u;
    n_sgroups = f.n_tiles ? host_readback[13] : 0u;
    n_ggroups = f.n_tiles ? host_readback[14] : 0u;
    n_fgroups = f.n_tiles ? host_readback[15] : 0u;
    if (n_atiles > f.n_tiles || (uint64_t)n_fgroups + n_sgroups + n_ggroups > f.n_groups) return kErrDevice;
    // every edge segment crosses at least one band and at most all of them;
    // the command lists hold at most one entry per (band, command) cell
    if ((n_binned < n_segs && !some_failed) || (uint64_t)n_binned > (uint64_t)n_segs * std::max<uint32_t>(max_bands(), 1) ||
        n_list > n_mat)
      return kErrDevice;
    s.binned = (DevSeg*)take(binned, (size_t)std::max<uint32_t>(n_binned, 1) * sizeof(DevSeg));
    s.list_ent = (uint4*)take(list, (size_t)std::max<uint32_t>(n_list, 1) * 16);
    s.list_style = (uint32_t*)take(list_style, (size_t)std::max<uint32_t>(n_list, 1) * 4);
    if (!s.binned || !s.list_ent || !s.list_style) return kErrFit;
  }
  launch_copy_u32(cursor_p, s.mat_off, n_mat + 1, stream);
  launch_bin_scatter(f, s, n_segs, cursor_p, stream);
  launch_list_build(f, s, stream);

  // Stage 3: render.
  launch_render_setup(f, s, stream);
  if (!graph) HIP_OK(hipEventRecord(ev0, stream));
  if (uint32_t r = render_stage(s, n_atiles, n_fgroups, n_sgroups, n_ggroups, !arena, stats)) return r;
  stats.render_kernels = (n_atiles ? 1u : 0u) | (n_sgroups ? 2u : 0u) | (n_ggroups ? 4u : 0u) | (n_fgroups ? 8u : 0u);
  if (!graph) HIP_OK(hipEventRecord(ev1, stream));
  launch_render_stats(f, s, stream);
  if (graph) {
    hipGraph_t g = nullptr;
    capture.on = false;
    HIP_OK(hipStreamEndCapture(stream, &g));
    const hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { graph_exec = nullptr; report_hip("hipGraphInstantiate", e, __LINE__); return kErrDevice; }
    graph_gen = gen;
    graph_sched = g_sched;
    graph_wgs = g_wgs;
    graph_model = g_tile_model_version;
    return run_graph();
  }
  HIP_OK(hipMemcpyAsync(host_readback + 2, s.pixels, 8, hipMemcpyDeviceToHost, stream));
  HIP_OK(hipMemcpyAsync(host_readback + 4, s.err_any, 4, hipMemcpyDeviceToHost, stream));
  if (collect && !arena) {
    cm::launch_actuals(f, s, model, n_segs, stream);
    const size_t nc = f.n_cmds;
    std::vector<float> feat(nc * cm::kIn);
    std::vector<uint32_t> act(2 * std::max<size_t>(nc, 1));
    HIP_OK(hipMemcpyAsync(feat.data(), model.features, feat.size() * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(act.data(), model.actual_segs, act.size() * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    collected_x.clear();
    collected_t.clear();
    for (size_t c = 0; c < nc; c++) {
      const float* x = &feat[c * cm::kIn];
      if (x[0] + x[1] + x[2] <= 0.0f) continue;   // no source segments (or not an analytic fill)
      collected_x.insert(collected_x.end(), x, x + cm::kIn);
      collected_t.push_back((float)log2(1.0 + act[c]));
      collected_t.push_back((float)log2(1.0 + act[std::max<size_t>(nc, 1) + c]));
    }
  }
  HIP_OK(hipStreamSynchronize(stream));
  stats.syncs++;
  if (hipError_t e_ = hipGetLastError(); e_ != hipSuccess) { report_hip("kernel launch", e_, __LINE__); return kErrDevice; }
  // (the arena execution is a replay: its failures were reported by the
  // flush that uploaded the frame, and it keeps no host-side list)
  if (!arena)
    if (uint32_t r = read_failures(s, host_readback[4])) return r;
  if (collect_tiles && !arena) {
    tile_x.assign((size_t)n_atiles * 16, 0.0f);
    std::vector<uint32_t> ticks((size_t)n_atiles * 2);
    if (n_atiles) {
      HIP_OK(hipMemcpyAsync(tile_x.data(), q_feat.ptr, tile_x.size() * 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipMemcpyAsync(ticks.data(), q_ticks.ptr, ticks.size() * 4, hipMemcpyDeviceToHost, stream));
      HIP_OK(hipStreamSynchronize(stream));
    }
    tile_t.resize((size_t)n_atiles * 2);
    for (size_t i = 0; i < n_atiles; i++) {
      tile_t[2 * i] = (float)log2(1.0 + ticks[i]);
      tile_t[2 * i + 1] = (float)log2(1.0 + ticks[n_atiles + i]);
    }
  }
  hist_segs = n_segs;
  hist_bins = n_binned;
  if (!arena) {
    plan.valid = true;
    plan.gen = gen;
    plan.n_segs = n_segs;
    plan.some_failed = some_failed ? 1u : 0u;
    plan.n_binned = n_binned;
    plan.n_list = n_list;
    plan.n_atiles = n_atiles;
    plan.n_fgroups = n_fgroups;
    plan.n_sgroups = n_sgroups;
    plan.n_ggroups = n_ggroups;
  }
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  auto t1 = std::chrono::steady_clock::now();

  uint64_t pixels;
  memcpy(&pixels, host_readback + 2, 8);
  stats.cmds = f.n_cmds;
  stats.src = n_src;
  stats.segs = n_segs;
  stats.binned = n_binned;
  stats.tiles = f.n_tiles;
  stats.pixels = pixels;
  stats.frame_bytes = frame_bytes;
  stats.lut_bytes = lut_bytes;
  stats.image_bytes = image_bytes;
  stats.render_ms = ms;
  stats.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return BL_AMD_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------------------
// Public objects
// ---------------------------------------------------------------------------

struct BLAmdImage {
  void* dptr = nullptr;
  int64_t stride = 0;
  int w = 0, h = 0;
  uint32_t format = 0;
  // The caller's reference (bl_amd_image_create / bl_amd_image_destroy)
  // plus one per context that renders into the image, per fill style that
  // samples it and per recorded or resident (replayable) frame that uses it:
  // the pixels are freed when the last reference goes, as the reference
  // defers the free of an image that still has users (core/image.cpp:171-175).
  std::atomic<int> refs{1};
  ~BLAmdImage() { if (dptr) (void)hipFree(dptr); }
};

namespace {

void image_retain(const BLAmdImage* img) {
  if (img) const_cast<BLAmdImage*>(img)->refs.fetch_add(1, std::memory_order_relaxed);
}
void image_release(const BLAmdImage* img) {
  if (img && const_cast<BLAmdImage*>(img)->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete img;
}

// A counted reference to an image.
class ImageRef {
 public:
  ImageRef() = default;
  explicit ImageRef(const BLAmdImage* p) : p_(p) { image_retain(p_); }
  ImageRef(const ImageRef& o) : p_(o.p_) { image_retain(p_); }
  ImageRef& operator=(const ImageRef& o) {
    image_retain(o.p_);
    image_release(p_);
    p_ = o.p_;
    return *this;
  }
  ~ImageRef() { image_release(p_); }
  const BLAmdImage* get() const { return p_; }
  const BLAmdImage* operator->() const { return p_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  const BLAmdImage* p_ = nullptr;
};

enum StyleType { kStyleSolid = 0, kStyleGradient = 1, kStylePattern = 2 };

// Fill style captured at set time with the transform in effect then
// (rastercontext.cpp:653-695 initializes the fetch data in set_style).
struct StyleState {
  uint32_t type = kStyleSolid;
  uint32_t fmt = kFxFRGB32;     // FormatExt seen by comp-op simplification
  uint32_t status = 0;          // 0 renders, 1 NOP, otherwise an error code
  DevStyle dev{};
  // gradient: its stops and table size (the table is built on the device)
  std::vector<Stop> stops;
  uint32_t lut_size = 0;
  ImageRef image;
  // blits: the pattern samples the rectangle view[0..3] = x, y, w, h of the image
  bool has_view = false;
  int32_t view[4] = {0, 0, 0, 0};
  uint64_t version = 0;
};

}  // namespace

struct BLAmdContext {
  BLAmdImage* target = nullptr;
  uint32_t comp_op = kCompSrcOver;
  uint32_t fill_rule = 0;
  uint32_t pattern_quality = 1;
  uint32_t gradient_quality = 0;
  Mat user{1, 0, 0, 1, 0, 0};
  double global_alpha = 1.0;
  double fill_alpha = 1.0;
  StyleState style;
  uint64_t style_counter = 0;
  // style version -> index in frame.styles (single cached entry)
  uint64_t frame_style_version = ~0ull;
  uint32_t frame_style_index = 0;

  FrameHost frame;          // commands since the last flush (one canvas)
  std::vector<ImageRef> frame_images;     // pattern images of `frame`
  std::vector<ImageRef> resident_images;  // target and pattern images of the resident frame
  Runner runner;
  Stats stats;

  // Rendering calls since begin(); errors found on the device while
  // rendering (BLContextErrorFlags bits, context.h:156-189) and the first
  // failing call of the last flush with its error.
  uint64_t call_seq = 0;
  uint32_t error_flags = 0;
  uint64_t deferred_call = ~0ull;
  uint32_t deferred_code = 0;

  void reset_state() {
    comp_op = kCompSrcOver;
    fill_rule = 0;
    pattern_quality = 1;
    gradient_quality = 0;
    user = Mat{1, 0, 0, 1, 0, 0};
    global_alpha = 1.0;
    fill_alpha = 1.0;
    style = StyleState{};
    style.dev.kind = kFetchSolid;
    style.dev.solid = premultiply(0xFF000000u);
    style.version = ++style_counter;
    frame_style_version = ~0ull;
  }
};

struct BLAmdBatch {
  std::vector<BLAmdContext*> contexts;
  Runner runner;
  Stats stats;
  std::vector<ImageRef> resident_images;  // targets and pattern images of the resident frame
};

namespace {

// styleAlphaI of the fill slot: round(global_alpha * 255 * fill_alpha)
// (rastercontext.cpp:1920-1939, :2000-2029).
uint32_t alpha_i(const BLAmdContext* c) { return (uint32_t)round_to_int(c->global_alpha * 255.0 * c->fill_alpha); }

// FormatExt of a solid colour (formatFromRgba32, rastercontext.cpp:90-93).
uint32_t solid_format(uint32_t rgba32) { return rgba32 == 0 ? kFxZERO32 : (rgba32 >= 0xFF000000u ? kFxFRGB32 : kFxPRGB32); }

StyleState make_solid_style(BLAmdContext* c, uint32_t rgba32) {
  StyleState st;
  st.type = kStyleSolid;
  st.fmt = solid_format(rgba32);
  st.dev.kind = kFetchSolid;
  st.dev.solid = premultiply(rgba32);
  st.version = ++c->style_counter;
  return st;
}

uint32_t image_index(BLAmdContext* c, const BLAmdImage* img) {
  for (size_t i = 0; i < c->frame_images.size(); i++) if (c->frame_images[i].get() == img) return (uint32_t)i;
  c->frame_images.push_back(ImageRef(img));
  DevImage d{};
  d.ptr = (uint64_t)(uintptr_t)img->dptr;
  d.stride = img->stride;
  d.w = (uint32_t)img->w;
  d.h = (uint32_t)img->h;
  d.format = img->format;
  c->frame.images.push_back(d);
  return (uint32_t)(c->frame_images.size() - 1);
}

// The rectangle x, y, w, h of `img` as an image of its own (a blit's source
// area, rastercontext.cpp:3944-4017: init_image_source).
uint32_t image_view_index(BLAmdContext* c, const BLAmdImage* img, const int32_t v[4]) {
  const int64_t bpp = img->format == kFmtA8 ? 1 : 4;
  DevImage d{};
  d.ptr = (uint64_t)(uintptr_t)img->dptr + (uint64_t)((int64_t)v[1] * img->stride + (int64_t)v[0] * bpp);
  d.stride = img->stride;
  d.w = (uint32_t)v[2];
  d.h = (uint32_t)v[3];
  d.format = img->format;
  for (size_t i = 0; i < c->frame.images.size(); i++) {
    const DevImage& e = c->frame.images[i];
    if (c->frame_images[i].get() == img && e.ptr == d.ptr && e.w == d.w && e.h == d.h) return (uint32_t)i;
  }
  c->frame_images.push_back(ImageRef(img));
  c->frame.images.push_back(d);
  return (uint32_t)(c->frame_images.size() - 1);
}

// A resolved render call: the command template and the error the call
// reports if its geometry turns out to be non-empty (the reference looks the
// pipeline up only then, rastercontext.cpp:2734-2749).
struct Resolved {
  DevCmd cmd;
  uint32_t nop = 0;       // 1: the call renders nothing
  uint32_t status = 0;    // deferred error
  uint32_t overridden = 0;   // the comp op replaced the style by a solid colour
};

uint32_t exec_op(uint32_t comp) {
  switch (comp) {
    case kCompSrcCopy: return kOpSrcCopy;
    case kCompSrcOver: return kOpSrcOver;
    case kCompPlus: return kOpPlus;
    case kCompMultiply: return kOpMultiply;
    case kCompScreen: return kOpScreen;
    default: return 0xFFFFFFFFu;
  }
}

// Comp op + style -> command (rastercontext.cpp:1231-1342 and
// compopsimplifyimpl_p.h).  `nop_transform` marks calls that a non-invertible
// user transform turns into NOPs (geometry fills, not fill_all).
Resolved resolve_with(BLAmdContext* c, StyleState& st, uint32_t alpha, bool nop_transform) {
  Resolved r;
  memset(&r.cmd, 0, sizeof(r.cmd));
  if (alpha == 0 || st.status == 1 || (nop_transform && transform_type(c->user) == kTTInvalid)) { r.nop = 1; return r; }
  const uint32_t dst_fmt = c->target->format;
  Simplified s = simplify(c->comp_op, dst_fmt, st.fmt);
  if (s.op == kCompNop) { r.nop = 1; return r; }
  if (st.status) { r.status = st.status; return r; }
  uint32_t op = exec_op(s.op);
  if (op == 0xFFFFFFFFu ||
      ((op == kOpMultiply || op == kOpScreen) && s.dst != kFxPRGB32) ||
      (s.src == kFxA8 && s.dst != kFxA8 && op != kOpSrcCopy && op != kOpSrcOver && op != kOpPlus)) {
    r.status = kErrNotImpl;
    return r;
  }
  DevCmd& cmd = r.cmd;
  cmd.op = op;
  cmd.alpha = alpha;
  cmd.src_fmt = s.src;
  cmd.canvas = 0;

  if (s.solid != kSolidNone) {
    DevStyle d{};
    d.kind = kFetchSolid;
    d.solid = s.solid == kSolidTransparent ? 0u : s.solid == kSolidOpaqueBlack ? 0xFF000000u : 0xFFFFFFFFu;
    cmd.style = (uint32_t)c->frame.styles.size();
    c->frame.styles.push_back(d);
    r.overridden = 1;
    return r;
  }
  if (c->frame_style_version != st.version) {
    DevStyle d = st.dev;
    if (st.type == kStyleGradient && d.kind != kFetchSolid) {
      FrameHost& fr = c->frame;
      d.lut = (uint32_t)fr.lut_entries;
      fr.lut_jobs.push_back(lut::Job{(uint32_t)fr.lut_entries, st.lut_size, (uint32_t)fr.lut_stops.size(),
                                     (uint32_t)st.stops.size(), d.dither});
      fr.lut_stops.insert(fr.lut_stops.end(), st.stops.begin(), st.stops.end());
      fr.lut_entries += d.dither ? 2u * st.lut_size : st.lut_size;
      fr.dither |= d.dither != 0;
    }
    if (st.type == kStylePattern) d.img = st.has_view ? image_view_index(c, st.image.get(), st.view) : image_index(c, st.image.get());
    c->frame_style_index = (uint32_t)c->frame.styles.size();
    c->frame.styles.push_back(d);
    c->frame_style_version = st.version;
  }
  cmd.style = c->frame_style_index;
  return r;
}

Resolved resolve_fill(BLAmdContext* c, bool nop_transform = true) { return resolve_with(c, c->style, alpha_i(c), nop_transform); }

// Explicit solid colour (BL_CONTEXT_RESOLVE_EXPLICIT_SOLID_OP,
// rastercontext.cpp:1258-1275): the context's style is not touched.
Resolved resolve_solid(BLAmdContext* c, uint32_t rgba32, bool nop_transform = true) {
  StyleState st = make_solid_style(c, rgba32);
  return resolve_with(c, st, alpha_i(c), nop_transform);
}

// clear_* (resolve_clear_op, rastercontext.cpp:1232-1242): SRC_COPY of the
// solid-override colour (transparent unless the simplification picks one)
// at full alpha; global alpha, fill alpha and style do not apply.
Resolved resolve_clear(BLAmdContext* c, bool nop_transform) {
  Resolved r;
  memset(&r.cmd, 0, sizeof(r.cmd));
  if (nop_transform && transform_type(c->user) == kTTInvalid) { r.nop = 1; return r; }
  Simplified s = simplify(kCompSrcCopy, c->target->format, kFxPRGB32);
  if (s.op == kCompNop) { r.nop = 1; return r; }
  uint32_t op = exec_op(s.op);
  if (op == 0xFFFFFFFFu) { r.status = kErrNotImpl; return r; }
  DevStyle d{};
  d.kind = kFetchSolid;
  d.solid = s.solid == kSolidOpaqueBlack ? 0xFF000000u : s.solid == kSolidOpaqueWhite ? 0xFFFFFFFFu : 0u;
  r.cmd.op = op;
  r.cmd.alpha = 255;
  r.cmd.src_fmt = s.src;
  r.cmd.style = (uint32_t)c->frame.styles.size();
  c->frame.styles.push_back(d);
  return r;
}

CmdXform fixed_xform(const BLAmdContext* c) {
  CmdXform x;
  x.m00 = c->user.m00 * 256.0; x.m01 = c->user.m01 * 256.0;
  x.m10 = c->user.m10 * 256.0; x.m11 = c->user.m11 * 256.0;
  x.m20 = c->user.m20 * 256.0; x.m21 = c->user.m21 * 256.0;
  x.scale_only = transform_type(c->user) <= kTTScale ? 1u : 0u;
  x.pad = 0;
  return x;
}

uint32_t push_cmd(BLAmdContext* c, const DevCmd& cmd, const CmdXform* xf) {
  c->frame.cmds.push_back(cmd);
  c->frame.cmd_call.push_back(c->call_seq ? c->call_seq - 1 : 0);
  if (xf) c->frame.xforms.push_back(*xf);
  else c->frame.xforms.push_back(CmdXform{});
  return (uint32_t)(c->frame.cmds.size() - 1);
}

enum { kCmdMove = 0, kCmdOn = 1, kCmdQuad = 2, kCmdCubic = 4 };

// Splits a path into source segments the way EdgeSourcePath iterates it
// (edgebuilder_p.h:160-285, 1032-1070): figures start at move-to, quads take
// two vertices, cubics three, anything else closes the figure.
void record_path_segments(BLAmdContext* c, uint32_t cmd_i, const uint8_t* cmds, const double* vtx, size_t n) {
  const uint32_t vbase = (uint32_t)(c->frame.vtx.size() / 2);
  c->frame.vtx.insert(c->frame.vtx.end(), vtx, vtx + 2 * n);
  size_t i = 0;
  for (;;) {
    while (i < n && cmds[i] != kCmdMove) i++;
    if (i >= n) break;
    uint32_t start = vbase + (uint32_t)i, prev = start;
    i++;
    for (;;) {
      if (i < n && cmds[i] == kCmdOn) {
        c->frame.src.push_back(SrcSeg{cmd_i, kSrcLine, prev, vbase + (uint32_t)i});
        c->frame.items += 1;
        prev = vbase + (uint32_t)i;
        i++;
      }
      else if (i + 2 <= n && cmds[i] == kCmdQuad) {
        c->frame.src.push_back(SrcSeg{cmd_i, kSrcQuad, prev, vbase + (uint32_t)i});
        c->frame.items += 3;
        prev = vbase + (uint32_t)i + 1;
        i += 2;
      }
      else if (i + 3 <= n && cmds[i] == kCmdCubic) {
        c->frame.src.push_back(SrcSeg{cmd_i, kSrcCubic, prev, vbase + (uint32_t)i});
        c->frame.items += 8;
        prev = vbase + (uint32_t)i + 2;
        i += 3;
      }
      else {
        c->frame.src.push_back(SrcSeg{cmd_i, kSrcLine, prev, start});
        c->frame.items += 1;
        break;
      }
    }
  }
}

void record_poly_segments(BLAmdContext* c, uint32_t cmd_i, const double* vtx, size_t n) {
  const uint32_t vbase = (uint32_t)(c->frame.vtx.size() / 2);
  c->frame.vtx.insert(c->frame.vtx.end(), vtx, vtx + 2 * n);
  for (size_t i = 1; i < n; i++) c->frame.src.push_back(SrcSeg{cmd_i, kSrcLine, vbase + (uint32_t)i - 1, vbase + (uint32_t)i});
  c->frame.src.push_back(SrcSeg{cmd_i, kSrcLine, vbase + (uint32_t)n - 1, vbase});
  c->frame.items += n;
}

// Whether the vertices' hull reaches the canvas rows: the reference builds
// edges first and reports a pipeline error only for non-empty edges; the
// hull is the host-side stand-in for that test.
bool hull_reaches_rows(const BLAmdContext* c, const CmdXform& xf, const double* vtx, size_t n) {
  double lo = INFINITY, hi = -INFINITY;
  for (size_t i = 0; i < n; i++) {
    double y = vtx[2 * i] * xf.m01 + vtx[2 * i + 1] * xf.m11 + xf.m21;
    if (y != y) continue;
    lo = y < lo ? y : lo;
    hi = y > hi ? y : hi;
  }
  return lo < hi && hi > 0.0 && lo < (double)c->target->h * 256.0;
}

uint32_t fill_poly_rule(BLAmdContext* c, const Resolved& r, const double* vtx, size_t n, uint32_t rule) {
  if (n < 3 || r.nop) return BL_AMD_SUCCESS;
  CmdXform xf = fixed_xform(c);
  if (r.status) return hull_reaches_rows(c, xf, vtx, n) ? r.status : BL_AMD_SUCCESS;
  DevCmd cmd = r.cmd;
  cmd.kind = kCmdAnalytic;
  cmd.rule_mask = rule == 0 ? 0xFFFFFFFFu : 0x1FFu;
  uint32_t ci = push_cmd(c, cmd, &xf);
  record_poly_segments(c, ci, vtx, n);
  return BL_AMD_SUCCESS;
}

uint32_t fill_path_rule(BLAmdContext* c, const Resolved& r, const double* origin, const uint8_t* cmds, const double* vtx, size_t n,
                        uint32_t rule) {
  if (n == 0 || r.nop) return BL_AMD_SUCCESS;
  CmdXform xf = fixed_xform(c);
  if (origin) {
    // fill_unclipped_path_with_origin (rastercontext.cpp:2812-2821)
    double ox = origin[0] * xf.m00 + origin[1] * xf.m10 + xf.m20;
    double oy = origin[0] * xf.m01 + origin[1] * xf.m11 + xf.m21;
    xf.m20 = ox;
    xf.m21 = oy;
  }
  if (r.status) return hull_reaches_rows(c, xf, vtx, n) ? r.status : BL_AMD_SUCCESS;
  DevCmd cmd = r.cmd;
  cmd.kind = kCmdAnalytic;
  cmd.rule_mask = rule == 0 ? 0xFFFFFFFFu : 0x1FFu;
  uint32_t ci = push_cmd(c, cmd, &xf);
  record_path_segments(c, ci, cmds, vtx, n);
  return BL_AMD_SUCCESS;
}

// FillBoxA (fillgeneric_p.h:22-65)
uint32_t fill_box_a(BLAmdContext* c, const Resolved& r, int x0, int y0, int x1, int y1) {
  if (r.nop) return BL_AMD_SUCCESS;
  if (r.status) return r.status;
  DevCmd cmd = r.cmd;
  cmd.kind = kCmdBoxA;
  cmd.box[0] = x0; cmd.box[1] = y0; cmd.box[2] = x1; cmd.box[3] = y1;
  cmd.fetch_row0 = (uint32_t)y0;
  push_cmd(c, cmd, nullptr);
  return BL_AMD_SUCCESS;
}

// FillData::init_box_u_8bpc_24x8 (pipedefs_p.h:644-815) as per-pixel masks.
uint32_t fill_box_u(BLAmdContext* c, const Resolved& r, int x0, int y0, int x1, int y1) {
  if (r.nop) return BL_AMD_SUCCESS;
  if (r.status) return r.status;
  DevCmd cmd = r.cmd;
  const uint32_t alpha = cmd.alpha;
  uint32_t ax0 = (uint32_t)x0 >> 8, ay0 = (uint32_t)y0 >> 8;
  uint32_t ax1 = (uint32_t)(x1 + 0xFF) >> 8, ay1 = (uint32_t)(y1 + 0xFF) >> 8;
  uint32_t fx0 = (uint32_t)x0 & 0xFFu, fy0 = (uint32_t)y0 & 0xFFu;
  uint32_t fx1 = ((uint32_t)(x1 - 1) & 0xFFu) + 1u, fy1 = ((uint32_t)(y1 - 1) & 0xFFu) + 1u;
  uint32_t w = ax1 - ax0, h = ay1 - ay0;
  fy0 = (h == 1 ? fy1 : 256u) - fy0;
  uint32_t fy0_a = fy0 * alpha, fy1_a = fy1 * alpha;
  uint32_t m[3][3];
  if (w == 1) {
    uint32_t fx = fx1 - fx0;
    m[0][0] = m[0][1] = m[0][2] = (fx * fy0_a) >> 16;
    m[1][0] = m[1][1] = m[1][2] = (fx * alpha) >> 8;
    m[2][0] = m[2][1] = m[2][2] = (fx * fy1_a) >> 16;
  }
  else {
    uint32_t lx = 256 - fx0;
    m[0][0] = (lx * fy0_a) >> 16; m[0][1] = fy0_a >> 8; m[0][2] = (fx1 * fy0_a) >> 16;
    m[1][0] = (lx * alpha) >> 8;  m[1][1] = alpha;      m[1][2] = (fx1 * alpha) >> 8;
    m[2][0] = (lx * fy1_a) >> 16; m[2][1] = fy1_a >> 8; m[2][2] = (fx1 * fy1_a) >> 16;
  }
  uint32_t by0 = ay0, by1 = ay0 + h;
  uint32_t top_inner = w == 1 ? m[0][0] : m[0][1];
  uint32_t mid_inner = w == 1 ? m[1][0] : m[1][1];
  uint32_t bot_inner = w == 1 ? m[2][0] : m[2][1];
  if (h == 1) {
    if (top_inner == 0) return BL_AMD_SUCCESS;
  }
  else {
    by0 += (uint32_t)(top_inner == 0);
    by1 -= (uint32_t)(bot_inner == 0);
    if (!(by0 < by1)) return BL_AMD_SUCCESS;
    if (w == 1 && mid_inner == 0) return BL_AMD_SUCCESS;
  }
  cmd.kind = kCmdBoxU;
  cmd.box[0] = (int32_t)ax0; cmd.box[1] = (int32_t)by0; cmd.box[2] = (int32_t)ax1; cmd.box[3] = (int32_t)by1;
  cmd.box_row_first = (int32_t)ay0;
  cmd.box_row_last = (int32_t)(ay0 + h - 1);
  for (int k = 0; k < 3; k++) cmd.box_mask[k] = m[k][0] | (m[k][1] << 8) | (m[k][2] << 16);
  cmd.fetch_row0 = by0;
  push_cmd(c, cmd, nullptr);
  return BL_AMD_SUCCESS;
}

// fill_unclipped_box_d (rastercontext.cpp:2871-2891)
uint32_t fill_box_d(BLAmdContext* c, const Resolved& r, double bx0, double by0, double bx1, double by1) {
  if (r.nop) return BL_AMD_SUCCESS;
  CmdXform m = fixed_xform(c);
  int tt = std::max(transform_type(c->user), (int)kTTScale);
  double cx1 = (double)c->target->w * 256.0, cy1 = (double)c->target->h * 256.0;
  if (tt <= kTTSwap) {
    double x0 = bx0 * m.m00 + by0 * m.m10 + m.m20;
    double y0 = bx0 * m.m01 + by0 * m.m11 + m.m21;
    double x1 = bx1 * m.m00 + by1 * m.m10 + m.m20;
    double y1 = bx1 * m.m01 + by1 * m.m11 + m.m21;
    double bxa = x1 < x0 ? x1 : x0, bya = y1 < y0 ? y1 : y0;
    double bxb = x0 < x1 ? x1 : x0, byb = y0 < y1 ? y1 : y0;
    double dx0 = 0.0 < bxa ? bxa : 0.0, dy0 = 0.0 < bya ? bya : 0.0;
    double dx1 = cx1 < bxb ? cx1 : bxb, dy1 = cy1 < byb ? cy1 : byb;
    if (!((dx0 < dx1) & (dy0 < dy1))) return BL_AMD_SUCCESS;
    int ix0 = (int)dx0, iy0 = (int)dy0, ix1 = (int)dx1, iy1 = (int)dy1;
    if (ix0 >= ix1 || iy0 >= iy1) return BL_AMD_SUCCESS;
    if (((ix0 | iy0 | ix1 | iy1) & 0xFF) == 0) return fill_box_a(c, r, ix0 >> 8, iy0 >> 8, ix1 >> 8, iy1 >> 8);
    return fill_box_u(c, r, ix0, iy0, ix1, iy1);
  }
  double poly[8] = { bx0, by0, bx1, by0, bx1, by1, bx0, by1 };
  return fill_poly_rule(c, r, poly, 4, 1);
}

// ContextFlags::kInfoIntegralTranslation (rastercontext.cpp:773-789): the user
// transform is a translation by whole pixels within the safe fixed-point range
// (fp_min/max_safe_coord_d, rastercontext.cpp:4357-4358).
bool integral_translation(const BLAmdContext* c, int& tx, int& ty) {
  if (transform_type(c->user) > kTTTranslate) return false;
  double fx = c->user.m20 * 256.0, fy = c->user.m21 * 256.0;
  const double lo = floor((double)(INT32_MIN + 1) * 256.0);
  const double hi = floor((double)(INT32_MAX - 1 - std::max(c->target->w, c->target->h)) * 256.0);
  if (!(fx >= lo && fx <= hi && fy >= lo && fy <= hi)) return false;
  int64_t tx64 = floor_to_int64(fx), ty64 = floor_to_int64(fy);
  if (((tx64 | ty64) & 0xFF) != 0) return false;
  tx = (int)(tx64 >> 8);
  ty = (int)(ty64 >> 8);
  return true;
}

// fillUnclippedRectI (rastercontext.cpp:2898-2917)
uint32_t fill_rect_i(BLAmdContext* c, const Resolved& r, int x, int y, int w, int h) {
  if (r.nop) return BL_AMD_SUCCESS;
  int tx = 0, ty = 0;
  const bool integral = integral_translation(c, tx, ty);
  if (!integral) {
    if (w <= 0 || h <= 0) return BL_AMD_SUCCESS;
    return fill_box_d(c, r, (double)x, (double)y, (double)x + (double)w, (double)y + (double)h);
  }
  int64_t x0 = (int64_t)x + tx, y0 = (int64_t)y + ty;
  int64_t x1 = (int64_t)w + x0, y1 = (int64_t)h + y0;
  x0 = std::max<int64_t>(x0, 0); y0 = std::max<int64_t>(y0, 0);
  x1 = std::min<int64_t>(x1, c->target->w); y1 = std::min<int64_t>(y1, c->target->h);
  if (x0 >= x1 || y0 >= y1) return BL_AMD_SUCCESS;
  return fill_box_a(c, r, (int)x0, (int)y0, (int)x1, (int)y1);
}

// ---- blits (rastercontext.cpp:3938-4150) ----------------------------------
//
// A blit is a box fill whose style is the image's source area as a pattern
// (extend REFLECT, BL_RASTER_CONTEXT_PREFERRED_BLIT_EXTEND, rastercontext_p.h:51)
// under global alpha alone (BL_CONTEXT_RESOLVE_BLIT_OP, rastercontext.cpp:
// 1332-1342, 1445-1453); comp-op simplification sees the image's format.

constexpr uint32_t kBlitExtend = 2;   // REFLECT

// check_image_area (rastercontext.cpp:2305-2320): the whole image, or `area`
// when it lies inside the image.
bool blit_area(const BLAmdImage* img, const int32_t* area, int32_t out[4]) {
  out[0] = 0; out[1] = 0; out[2] = img->w; out[3] = img->h;
  if (area) {
    const unsigned max_w = (unsigned)img->w - (unsigned)area[0];
    const unsigned max_h = (unsigned)img->h - (unsigned)area[1];
    if ((max_w > (unsigned)img->w) | ((unsigned)area[2] > max_w) | (max_h > (unsigned)img->h) | ((unsigned)area[3] > max_h))
      return false;
    for (int k = 0; k < 4; k++) out[k] = area[k];
  }
  return true;
}

// The comp op replaces the image by a solid colour (CompOpSolidId of the
// simplification with the image's format): the blit renders its destination
// box without a pattern (prepare_overridden_fetch).
bool blit_overridden(const BLAmdContext* c, const BLAmdImage* img) {
  const Simplified s = simplify(c->comp_op, c->target->format, img->format);
  return s.op != kCompNop && s.solid != kSolidNone;
}

Resolved resolve_blit(BLAmdContext* c, const BLAmdImage* img, const int32_t view[4], const Mat& t) {
  StyleState st;
  st.type = kStylePattern;
  st.version = ++c->style_counter;
  st.fmt = img->format;
  st.image = ImageRef(img);
  st.has_view = true;
  for (int k = 0; k < 4; k++) st.view[k] = view[k];
  // The pattern is set up only when the comp op keeps the image; a pattern
  // that cannot be set up makes the blit a NOP (setup_pattern_affine
  // failing returns BL_SUCCESS, rastercontext.cpp:4019-4020).
  if (!blit_overridden(c, img) &&
      init_pattern(st.dev, (uint32_t)view[2], (uint32_t)view[3], kBlitExtend, c->pattern_quality, t) == 0)
    st.status = 1;
  return resolve_with(c, st, (uint32_t)round_to_int(c->global_alpha * 255.0), true);
}

double bl_max_d(double a, double b) { return a < b ? b : a; }
double bl_min_d(double a, double b) { return b < a ? b : a; }

// blit_image_d_impl (rastercontext.cpp:3940-4031).
uint32_t blit_d(BLAmdContext* c, double x, double y, const BLAmdImage* img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(img, area, src)) return kErrInvalid;
  const Mat& u = c->user;
  const int tt = transform_type(u);
  if (tt == kTTInvalid) return BL_AMD_SUCCESS;
  if (tt <= kTTTranslate) {
    // the pixel box of the blit, from the 24.8 start and the source size
    const CmdXform m = fixed_xform(c);
    const double sx = x * m.m00 + m.m20, sy = y * m.m11 + m.m21;
    const double dx0 = bl_max_d(sx, 0.0), dy0 = bl_max_d(sy, 0.0);
    const double dx1 = bl_min_d(sx + (double)src[2] * m.m00, (double)c->target->w * 256.0);
    const double dy1 = bl_min_d(sy + (double)src[3] * m.m11, (double)c->target->h * 256.0);
    // clipped out, invalid coordinates or an empty area: only a solid
    // override (which fills the unclipped box) can still render
    int ix0 = 0, iy0 = 0, ix1 = 0, iy1 = 0;
    bool empty = !((dx0 < dx1) & (dy0 < dy1));
    if (!empty) {
      ix0 = (int)dx0; iy0 = (int)dy0; ix1 = (int)dx1; iy1 = (int)dy1;
      empty = !((ix0 < ix1) & (iy0 < iy1));
    }
    if (empty && !blit_overridden(c, img)) return BL_AMD_SUCCESS;
    Mat t{1.0, 0.0, 0.0, 1.0, sx / 256.0, sy / 256.0};
    Resolved r = resolve_blit(c, img, src, t);
    if (r.nop) return BL_AMD_SUCCESS;
    if (r.overridden) return fill_box_d(c, r, x, y, x + (double)src[2], y + (double)src[3]);
    // fill_clipped_box_f: a pixel-aligned box composites with a constant mask
    if (((ix0 | iy0 | ix1 | iy1) & 0xFF) == 0) return fill_box_a(c, r, ix0 >> 8, iy0 >> 8, ix1 >> 8, iy1 >> 8);
    return fill_box_u(c, r, ix0, iy0, ix1, iy1);
  }
  // any other transform: the source translated to the origin, through it
  const Mat tr{1.0, 0.0, 0.0, 1.0, x, y};
  Resolved r = resolve_blit(c, img, src, multiply(tr, u));
  if (r.nop) return BL_AMD_SUCCESS;
  return fill_box_d(c, r, x, y, x + (double)src[2], y + (double)src[3]);
}

// blit_image_i_impl (rastercontext.cpp:4033-4066): with an integral
// translation the destination box is clipped in integers; otherwise as
// blit_image_d.
uint32_t blit_i(BLAmdContext* c, int x, int y, const BLAmdImage* img, const int32_t* area) {
  int tx = 0, ty = 0;
  if (!integral_translation(c, tx, ty)) return blit_d(c, (double)x, (double)y, img, area);
  int32_t src[4];
  if (!blit_area(img, area, src)) return kErrInvalid;
  const int64_t dx = (int64_t)x + tx, dy = (int64_t)y + ty;
  const int64_t x0 = std::max<int64_t>(dx, 0), y0 = std::max<int64_t>(dy, 0);
  const int64_t x1 = std::min<int64_t>(dx + src[2], c->target->w), y1 = std::min<int64_t>(dy + src[3], c->target->h);
  // clipped out: a NOP whatever the comp op (translate_and_clip_rect_to_blit_i)
  if (x0 >= x1 || y0 >= y1) return BL_AMD_SUCCESS;
  // the source pixel of destination pixel X is X - dx (+ the area's origin)
  Mat t{1.0, 0.0, 0.0, 1.0, (double)dx, (double)dy};
  Resolved r = resolve_blit(c, img, src, t);
  if (r.nop) return BL_AMD_SUCCESS;
  return fill_box_a(c, r, (int)x0, (int)y0, (int)x1, (int)y1);
}

// blit_scaled_image_d_impl (rastercontext.cpp:4072-4110).
uint32_t blit_scaled_d(BLAmdContext* c, double x, double y, double w, double h, const BLAmdImage* img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(img, area, src)) return kErrInvalid;
  if (w == (double)src[2] && h == (double)src[3]) return blit_d(c, x, y, img, area);
  const Mat sm{w / (double)src[2], 0.0, 0.0, h / (double)src[3], x, y};
  Resolved r = resolve_blit(c, img, src, multiply(sm, c->user));
  if (r.nop) return BL_AMD_SUCCESS;
  return fill_box_d(c, r, x, y, x + w, y + h);
}

// blit_scaled_image_i_impl (rastercontext.cpp:4112-4150).
uint32_t blit_scaled_i(BLAmdContext* c, int x, int y, int w, int h, const BLAmdImage* img, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(img, area, src)) return kErrInvalid;
  if (w == src[2] && h == src[3]) return blit_i(c, x, y, img, area);
  const Mat sm{(double)w / (double)src[2], 0.0, 0.0, (double)h / (double)src[3], (double)x, (double)y};
  Resolved r = resolve_blit(c, img, src, multiply(sm, c->user));
  if (r.nop) return BL_AMD_SUCCESS;
  return fill_box_d(c, r, (double)x, (double)y, (double)x + (double)w, (double)y + (double)h);
}

// ---- mask fills (rastercontext.cpp:3590-3780) ----------------------------
//
// The reference renders a mask only pixel-aligned under a translation: a box
// whose pixels take their coverage from bytes of the mask area's rows
// (fill_clipped_box_masked_a); any other placement is NOT_IMPLEMENTED there
// (fill_unclipped_mask_d), and so here.

// The box x0..x1 x y0..y1 with the mask area at (mx, my) on its top-left
// pixel: the area's view is exactly the box.
uint32_t fill_box_masked(BLAmdContext* c, const Resolved& r, int x0, int y0, int x1, int y1, const BLAmdImage* mask,
                         int mx, int my) {
  if (r.nop) return BL_AMD_SUCCESS;
  if (r.status) return r.status;
  DevCmd cmd = r.cmd;
  cmd.kind = kCmdBoxM;
  cmd.box[0] = x0; cmd.box[1] = y0; cmd.box[2] = x1; cmd.box[3] = y1;
  cmd.fetch_row0 = (uint32_t)y0;
  const int32_t view[4] = { mx, my, x1 - x0, y1 - y0 };
  cmd.box_mask[0] = image_view_index(c, mask, view);
  push_cmd(c, cmd, nullptr);
  return BL_AMD_SUCCESS;
}

// fill_mask_d_impl + fill_unclipped_mask_d (rastercontext.cpp:3594-3690).
uint32_t fill_mask_d(BLAmdContext* c, const Resolved& r, double x, double y, const BLAmdImage* mask, const int32_t* area) {
  int32_t src[4];
  if (!blit_area(mask, area, src)) return kErrInvalid;
  if (r.nop) return BL_AMD_SUCCESS;
  if (transform_type(c->user) > kTTTranslate) return kErrNotImpl;
  const CmdXform m = fixed_xform(c);
  const double sx = x * m.m00 + m.m20, sy = y * m.m11 + m.m21;
  const double dx0 = bl_max_d(sx, 0.0), dy0 = bl_max_d(sy, 0.0);
  const double dx1 = bl_min_d(sx + (double)src[2] * m.m00, (double)c->target->w * 256.0);
  const double dy1 = bl_min_d(sy + (double)src[3] * m.m11, (double)c->target->h * 256.0);
  // clipped out, invalid coordinates or an empty area
  if (!((dx0 < dx1) & (dy0 < dy1))) return BL_AMD_SUCCESS;
  const int64_t fx = floor_to_int64(sx), fy = floor_to_int64(sy);
  if ((fx | fy) & 0xFF) return kErrNotImpl;
  const int ix0 = (int)dx0, iy0 = (int)dy0, ix1 = (int)dx1, iy1 = (int)dy1;
  const int x0 = ix0 >> 8, y0 = iy0 >> 8, x1 = (ix1 + 0xFF) >> 8, y1 = (iy1 + 0xFF) >> 8;
  if (((ix0 | iy0 | ix1 | iy1) & 0xFF) != 0) return kErrNotImpl;
  const int tx = (int)(fx >> 8), ty = (int)(fy >> 8);
  return fill_box_masked(c, r, x0, y0, x1, y1, mask, src[0] + x0 - tx, src[1] + y0 - ty);
}

// fill_mask_i_impl (rastercontext.cpp:3723-3743): an integral translation
// clips in integers (translate_and_clip_rect_to_blit_i), else fill_mask_d.
uint32_t fill_mask_i(BLAmdContext* c, const Resolved& r, int x, int y, const BLAmdImage* mask, const int32_t* area) {
  int tx = 0, ty = 0;
  if (!integral_translation(c, tx, ty)) return fill_mask_d(c, r, (double)x, (double)y, mask, area);
  int32_t src[4];
  if (!blit_area(mask, area, src)) return kErrInvalid;
  const int64_t dx = (int64_t)x + tx, dy = (int64_t)y + ty;
  const int64_t x0 = std::max<int64_t>(dx, 0), y0 = std::max<int64_t>(dy, 0);
  const int64_t x1 = std::min<int64_t>(dx + src[2], c->target->w), y1 = std::min<int64_t>(dy + src[3], c->target->h);
  if (x0 >= x1 || y0 >= y1) return BL_AMD_SUCCESS;
  return fill_box_masked(c, r, (int)x0, (int)y0, (int)x1, (int)y1, mask, src[0] + (int)(x0 - dx), src[1] + (int)(y0 - dy));
}

// fill_all (rastercontext.cpp:2720-2725) with the canvas as the clip box.
uint32_t fill_all_box(BLAmdContext* c, const Resolved& r) {
  return fill_box_a(c, r, 0, 0, c->target->w, c->target->h);
}

// The path BLPath::add_geometry builds for a geometry (core/path.cpp:1396-1640),
// clockwise direction.
const double kKappa = 0.55228474983079339840;   // 4/3 (sqrt(2) - 1)

struct TmpPath {
  std::vector<uint8_t> cmd;
  std::vector<double> vtx;
  void move_to(double x, double y) { cmd.push_back(kCmdMove); vtx.push_back(x); vtx.push_back(y); }
  void line_to(double x, double y) { cmd.push_back(kCmdOn); vtx.push_back(x); vtx.push_back(y); }
  void cubic_to(double x1, double y1, double x2, double y2, double x3, double y3) {
    cmd.push_back(kCmdCubic); cmd.push_back(kCmdCubic); cmd.push_back(kCmdOn);
    vtx.insert(vtx.end(), {x1, y1, x2, y2, x3, y3});
  }
  void close() { cmd.push_back(5); vtx.push_back(NAN); vtx.push_back(NAN); }
  void box(double x0, double y0, double x1, double y1) {
    // append_box_internal, clockwise
    move_to(x0, y0); line_to(x1, y0); line_to(x1, y1); line_to(x0, y1); close();
  }
};

uint32_t check_target(const BLAmdContext* c) { return c && c->target ? BL_AMD_SUCCESS : kErrState; }

// Entry of every rendering call (fill_*, clear_*): counts it, then checks the target.
uint32_t render_call(BLAmdContext* c) {
  if (c) c->call_seq++;
  return check_target(c);
}

// BLContextErrorFlags (context.h:156-189) and result codes of a command that
// failed on the device.
constexpr uint32_t kFlagInvalidGeometry = 0x00000004u, kFlagOutOfMemory = 0x40000000u;
constexpr uint32_t kErrInvalidGeometry = BL_AMD_ERROR_INVALID_GEOMETRY;

// Hands the commands that failed in `runner`'s last execution to the
// contexts that recorded them (`parts` in frame order, context k recorded
// parts[k]): accumulated error flags, and the first failing call of each.
void attribute_failures(const Runner& runner, const std::vector<BLAmdContext*>& ctxs) {
  std::vector<uint32_t> base(ctxs.size() + 1, 0);
  for (size_t k = 0; k < ctxs.size(); k++) {
    base[k + 1] = base[k] + (uint32_t)ctxs[k]->frame.cmds.size();
    ctxs[k]->deferred_call = ~0ull;
    ctxs[k]->deferred_code = 0;
  }
  for (const auto& fc : runner.failed) {
    const size_t k = (size_t)(std::upper_bound(base.begin(), base.end(), fc.first) - base.begin()) - 1;
    if (k >= ctxs.size()) continue;
    BLAmdContext* c = ctxs[k];
    const uint32_t local = fc.first - base[k];
    const uint64_t call = local < c->frame.cmd_call.size() ? c->frame.cmd_call[local] : 0;
    const bool geom = (fc.second & kFailInvalidGeometry) != 0;
    c->error_flags |= geom ? kFlagInvalidGeometry : kFlagOutOfMemory;
    if (call < c->deferred_call) {
      c->deferred_call = call;
      c->deferred_code = geom ? kErrInvalidGeometry : kErrOOM;
    }
  }
}

// Prepares a context's frame for merging: canvas record.
void finalize_frame(BLAmdContext* c) {
  c->frame.canvases.clear();
  DevCanvas cv{};
  cv.ptr = (uint64_t)(uintptr_t)c->target->dptr;
  cv.stride = c->target->stride;
  cv.w = (uint32_t)c->target->w;
  cv.h = (uint32_t)c->target->h;
  cv.format = c->target->format;
  cv.cmd_base = 0;
  cv.n_cmds = (uint32_t)c->frame.cmds.size();
  c->frame.canvases.push_back(cv);
}

void reset_frame(BLAmdContext* c) {
  c->frame.clear();
  c->frame_images.clear();
  c->frame_style_version = ~0ull;
}

// The images a frame just uploaded uses stay referenced while it is resident
// (replayable): the target and the pattern images.
void keep_resident(std::vector<ImageRef>& keep, const BLAmdContext* c) {
  keep.push_back(ImageRef(c->target));
  keep.insert(keep.end(), c->frame_images.begin(), c->frame_images.end());
}

uint32_t flush_one(BLAmdContext* c) {
  finalize_frame(c);
  uint32_t r = c->runner.run(std::vector<FrameHost*>{&c->frame}, c->stats);
  attribute_failures(c->runner, std::vector<BLAmdContext*>{c});
  if (!c->frame.cmds.empty()) {
    c->resident_images.clear();
    keep_resident(c->resident_images, c);
  }
  reset_frame(c);
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

uint32_t bl_amd_version(void) { return (0u << 16) | (1u << 8) | 0u; }

int bl_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

BLAmdResult bl_amd_image_create(BLAmdImage** out, int w, int h, uint32_t format) {
  if (!out) return kErrInvalid;
  *out = nullptr;
  if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return kErrInvalid;
  if (format != BL_AMD_FORMAT_PRGB32 && format != BL_AMD_FORMAT_XRGB32 && format != BL_AMD_FORMAT_A8) return kErrInvalid;
  std::unique_ptr<BLAmdImage> img(new (std::nothrow) BLAmdImage());
  if (!img) return kErrOOM;
  const int bpp = format == BL_AMD_FORMAT_A8 ? 1 : 4;
  img->stride = ((int64_t)w * bpp + 15) & ~(int64_t)15;
  img->w = w; img->h = h; img->format = format;
  if (hipMalloc(&img->dptr, (size_t)img->stride * (size_t)h) != hipSuccess) { img->dptr = nullptr; return kErrOOM; }
  if (hipMemset(img->dptr, 0, (size_t)img->stride * (size_t)h) != hipSuccess) return kErrDevice;
  *out = img.release();
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_image_destroy(BLAmdImage* img) {
  image_release(img);
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_image_upload(BLAmdImage* img, const void* pixels, intptr_t stride) {
  if (!img || !pixels) return kErrInvalid;
  const size_t row = (size_t)img->w * (img->format == BL_AMD_FORMAT_A8 ? 1 : 4);
  if (hipMemcpy2D(img->dptr, (size_t)img->stride, pixels, (size_t)stride, row, (size_t)img->h, hipMemcpyHostToDevice) != hipSuccess)
    return kErrDevice;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_image_download(const BLAmdImage* img, void* pixels, intptr_t stride) {
  if (!img || !pixels) return kErrInvalid;
  const size_t row = (size_t)img->w * (img->format == BL_AMD_FORMAT_A8 ? 1 : 4);
  if (hipMemcpy2D(pixels, (size_t)stride, img->dptr, (size_t)img->stride, row, (size_t)img->h, hipMemcpyDeviceToHost) != hipSuccess)
    return kErrDevice;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_image_fill_raw(BLAmdImage* img, uint32_t value) {
  if (!img) return kErrInvalid;
  if (img->format == BL_AMD_FORMAT_A8) {
    if (hipMemset(img->dptr, (int)(value & 0xFFu), (size_t)img->stride * img->h) != hipSuccess) return kErrDevice;
    return BL_AMD_SUCCESS;
  }
  // the row padding (stride is a multiple of 4) takes the value too
  if (hipMemsetD32((hipDeviceptr_t)img->dptr, (int)value, (size_t)img->stride / 4 * img->h) != hipSuccess)
    return kErrDevice;
  return hipDeviceSynchronize() == hipSuccess ? BL_AMD_SUCCESS : kErrDevice;
}

BLAmdResult bl_amd_image_get_data(const BLAmdImage* img, int* w, int* h, uint32_t* format, void** device_pixels, intptr_t* stride) {
  if (!img) return kErrInvalid;
  if (w) *w = img->w;
  if (h) *h = img->h;
  if (format) *format = img->format;
  if (device_pixels) *device_pixels = img->dptr;
  if (stride) *stride = (intptr_t)img->stride;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_create(BLAmdContext** out) {
  if (!out) return kErrInvalid;
  *out = new (std::nothrow) BLAmdContext();
  if (!*out) return kErrOOM;
  (*out)->reset_state();
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_destroy(BLAmdContext* ctx) {
  if (ctx && ctx->target) {
    image_release(ctx->target);
    ctx->target = nullptr;
  }
  delete ctx;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_begin(BLAmdContext* ctx, BLAmdImage* target) {
  if (!ctx || !target) return kErrInvalid;
  if (ctx->target) {
    uint32_t r = bl_amd_context_end(ctx);
    if (r) return r;
  }
  image_retain(target);
  ctx->target = target;
  ctx->reset_state();
  reset_frame(ctx);
  ctx->call_seq = 0;
  ctx->error_flags = 0;
  ctx->deferred_call = ~0ull;
  ctx->deferred_code = 0;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_flush(BLAmdContext* ctx, uint32_t flags) {
  (void)flags;
  if (uint32_t r = check_target(ctx)) return r;
  return flush_one(ctx);
}

BLAmdResult bl_amd_context_end(BLAmdContext* ctx) {
  if (uint32_t r = check_target(ctx)) return r;
  uint32_t r = flush_one(ctx);
  image_release(ctx->target);
  ctx->target = nullptr;
  return r;
}

BLAmdResult bl_amd_context_set_comp_op(BLAmdContext* ctx, uint32_t comp_op) {
  if (uint32_t r = check_target(ctx)) return r;
  if (comp_op > 28) return kErrInvalid;
  ctx->comp_op = comp_op;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_set_fill_rule(BLAmdContext* ctx, uint32_t fill_rule) {
  if (uint32_t r = check_target(ctx)) return r;
  if (fill_rule > 1) return kErrInvalid;
  ctx->fill_rule = fill_rule;
  return BL_AMD_SUCCESS;
}

// set_global_alpha_impl (rastercontext.cpp:2000-2029)
BLAmdResult bl_amd_context_set_global_alpha(BLAmdContext* ctx, double alpha) {
  if (uint32_t r = check_target(ctx)) return r;
  if (isnan(alpha)) return kErrInvalid;
  ctx->global_alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_set_transform(BLAmdContext* ctx, const double m[6]) {
  if (uint32_t r = check_target(ctx)) return r;
  if (!m) return kErrInvalid;
  ctx->user = Mat{m[0], m[1], m[2], m[3], m[4], m[5]};
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_set_pattern_quality(BLAmdContext* ctx, uint32_t quality) {
  if (uint32_t r = check_target(ctx)) return r;
  if (quality > 1) return kErrInvalid;
  ctx->pattern_quality = quality;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_set_gradient_quality(BLAmdContext* ctx, uint32_t quality) {
  if (uint32_t r = check_target(ctx)) return r;
  if (quality > 2) return kErrInvalid;
  ctx->gradient_quality = quality;
  return BL_AMD_SUCCESS;
}

// resolve_explicit_solid_op (rastercontext.cpp:1259-1275, :90-93)
BLAmdResult bl_amd_context_set_fill_style_rgba32(BLAmdContext* ctx, uint32_t rgba32) {
  if (uint32_t r = check_target(ctx)) return r;
  StyleState& st = ctx->style;
  st = StyleState{};
  st.type = kStyleSolid;
  st.fmt = rgba32 == 0 ? kFxZERO32 : (rgba32 >= 0xFF000000u ? kFxFRGB32 : kFxPRGB32);
  st.dev.kind = kFetchSolid;
  st.dev.solid = premultiply(rgba32);
  st.version = ++ctx->style_counter;
  return BL_AMD_SUCCESS;
}

}  // extern "C"

namespace {

// init_non_solid_fetch_data, gradient branch (rastercontext.cpp:416-481):
// the style takes the user transform of the moment.
uint32_t build_gradient_style(BLAmdContext* ctx, StyleState& st, uint32_t type, const double* values, uint32_t extend,
                              const double* stop_offsets, const uint32_t* stop_rgba32, size_t n_stops,
                              const double* transform) {
  if (type > 2 || extend > 2 || !values || (n_stops && (!stop_offsets || !stop_rgba32))) return kErrInvalid;
  st = StyleState{};
  st.type = kStyleGradient;
  st.version = ++ctx->style_counter;

  std::vector<Stop> stops;
  for (size_t i = 0; i < n_stops; i++) add_stop(stops, stop_offsets[i], stop_rgba32[i]);
  if (stops.empty()) { st.status = 1; return BL_AMD_SUCCESS; }

  bool solid;
  uint32_t fmt;
  uint32_t lut_size = gradient_info(stops, solid, fmt);
  st.fmt = fmt;
  if (solid) {
    uint64_t last = stops.back().rgba64;
    uint32_t c32 = (uint32_t)(((last >> 56) & 0xFFu) << 24 | ((last >> 40) & 0xFFu) << 16 | ((last >> 24) & 0xFFu) << 8 | ((last >> 8) & 0xFFu));
    st.dev.kind = kFetchSolid;
    st.dev.solid = premultiply(c32);
    return BL_AMD_SUCCESS;
  }

  Mat t = ctx->user;
  if (transform) {
    Mat gm{transform[0], transform[1], transform[2], transform[3], transform[4], transform[5]};
    if (transform_type(gm) != kTTIdentity) t = multiply(gm, ctx->user);
  }
  // DITHER samples a 64-bit table of twice the size; A8 targets never dither
  // (rastercontext.cpp:456-467, GradientInfo::lut_size, core/gradient_p.h:113-120)
  if (ctx->gradient_quality >= 2 && ctx->target && ctx->target->format != kFmtA8) {
    st.dev.dither = 1;
    lut_size = std::min<uint32_t>(lut_size * 2u, 1024u);
  }
  st.lut_size = lut_size;
  st.stops = std::move(stops);
  bool ok = type == 0 ? init_linear(st.dev, lut_size, values, extend, t)
          : type == 1 ? init_radial(st.dev, lut_size, values, extend, t)
                      : init_conic(st.dev, lut_size, values, t);
  if (!ok) st.status = 1;
  return BL_AMD_SUCCESS;
}

// init_non_solid_fetch_data, pattern branch (rastercontext.cpp:377-414).
uint32_t build_pattern_style(BLAmdContext* ctx, StyleState& st, const BLAmdImage* image, uint32_t extend,
                             const double* transform) {
  if (!image || extend > 2) return kErrInvalid;
  st = StyleState{};
  st.type = kStylePattern;
  st.version = ++ctx->style_counter;
  st.fmt = image->format;
  st.image = ImageRef(image);
  Mat t = ctx->user;
  if (transform) {
    Mat pm{transform[0], transform[1], transform[2], transform[3], transform[4], transform[5]};
    if (transform_type(pm) != kTTIdentity) t = multiply(pm, ctx->user);
  }
  int r = init_pattern(st.dev, (uint32_t)image->w, (uint32_t)image->h, extend, ctx->pattern_quality, t);
  if (r == 0) st.status = 1;
  return BL_AMD_SUCCESS;
}

// rgba32FromRgba64 (core/rgba_p.h:67-74): the high byte of each channel.
uint32_t rgba32_from_rgba64(uint64_t v) {
  return (uint32_t)(((v >> 56) & 0xFFu) << 24 | ((v >> 40) & 0xFFu) << 16 | ((v >> 24) & 0xFFu) << 8 | ((v >> 8) & 0xFFu));
}

// resolve_explicit_style_op (rastercontext.cpp:1277-1328): the call's own
// style, built with the user transform of the moment; the context's style is
// not touched.  Sets `err` for an invalid style description.
Resolved resolve_explicit(BLAmdContext* ctx, const BLAmdStyle* style, bool nop_transform, uint32_t& err) {
  err = BL_AMD_SUCCESS;
  Resolved nop;
  memset(&nop.cmd, 0, sizeof(nop.cmd));
  nop.nop = 1;
  if (!style) { err = kErrInvalid; return nop; }
  switch (style->kind) {
    case BL_AMD_STYLE_RGBA32: return resolve_solid(ctx, style->rgba32, nop_transform);
    case BL_AMD_STYLE_RGBA64: return resolve_solid(ctx, rgba32_from_rgba64(style->rgba64), nop_transform);
    case BL_AMD_STYLE_GRADIENT:
    case BL_AMD_STYLE_PATTERN: {
      StyleState st;
      err = style->kind == BL_AMD_STYLE_GRADIENT
                ? build_gradient_style(ctx, st, style->gradient_type, style->values, style->extend, style->stop_offsets,
                                       style->stop_rgba32, style->n_stops, style->transform)
                : build_pattern_style(ctx, st, style->image, style->extend, style->transform);
      if (err) return nop;
      return resolve_with(ctx, st, alpha_i(ctx), nop_transform);
    }
    default:
      err = kErrInvalid;
      return nop;
  }
}

}  // namespace

extern "C" {

BLAmdResult bl_amd_context_set_fill_style_gradient(BLAmdContext* ctx, uint32_t type, const double* values, uint32_t extend,
                                                   const double* stop_offsets, const uint32_t* stop_rgba32, size_t n_stops,
                                                   const double* transform) {
  if (uint32_t r = check_target(ctx)) return r;
  StyleState st;
  if (uint32_t r = build_gradient_style(ctx, st, type, values, extend, stop_offsets, stop_rgba32, n_stops, transform)) return r;
  ctx->style = st;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_set_fill_style_pattern(BLAmdContext* ctx, const BLAmdImage* image, uint32_t extend, const double* transform) {
  if (uint32_t r = check_target(ctx)) return r;
  StyleState st;
  if (uint32_t r = build_pattern_style(ctx, st, image, extend, transform)) return r;
  ctx->style = st;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_fill_path_d(BLAmdContext* ctx, const double* origin, const uint8_t* cmds, const double* vertices, size_t n) {
  if (uint32_t r = render_call(ctx)) return r;
  if (n == 0) return BL_AMD_SUCCESS;
  if (!cmds || !vertices) return kErrInvalid;
  return fill_path_rule(ctx, resolve_fill(ctx), origin, cmds, vertices, n, ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_path_d_rgba32(BLAmdContext* ctx, const double* origin, const uint8_t* cmds, const double* vertices,
                                              size_t n, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  if (n == 0) return BL_AMD_SUCCESS;
  if (!cmds || !vertices) return kErrInvalid;
  return fill_path_rule(ctx, resolve_solid(ctx, rgba32), origin, cmds, vertices, n, ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_polygon_d(BLAmdContext* ctx, const double* vertices, size_t n) {
  if (uint32_t r = render_call(ctx)) return r;
  if (n && !vertices) return kErrInvalid;
  if (n < 3) return BL_AMD_SUCCESS;
  return fill_poly_rule(ctx, resolve_fill(ctx), vertices, n, ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_rect_d(BLAmdContext* ctx, double x, double y, double w, double h) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_box_d(ctx, resolve_fill(ctx), x, y, x + w, y + h);
}

BLAmdResult bl_amd_context_fill_rect_d_rgba32(BLAmdContext* ctx, double x, double y, double w, double h, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_box_d(ctx, resolve_solid(ctx, rgba32), x, y, x + w, y + h);
}

BLAmdResult bl_amd_context_fill_mask_i(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask, const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_i(ctx, resolve_fill(ctx), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_d(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask, const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_d(ctx, resolve_fill(ctx), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_i_rgba32(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                              const int32_t* area, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_i(ctx, resolve_solid(ctx, rgba32), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_d_rgba32(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_d(ctx, resolve_solid(ctx, rgba32), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_i_rgba64(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                              const int32_t* area, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_i(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_d_rgba64(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                              const int32_t* area, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  return fill_mask_d(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_i_ext(BLAmdContext* ctx, int x, int y, const BLAmdImage* mask,
                                           const int32_t* area, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_mask_i(ctx, res, x, y, mask, area);
}

BLAmdResult bl_amd_context_fill_mask_d_ext(BLAmdContext* ctx, double x, double y, const BLAmdImage* mask,
                                           const int32_t* area, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!mask) return kErrInvalid;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_mask_d(ctx, res, x, y, mask, area);
}

BLAmdResult bl_amd_context_blit_image_i(BLAmdContext* ctx, int x, int y, const BLAmdImage* image, const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!image) return kErrInvalid;
  return blit_i(ctx, x, y, image, area);
}

BLAmdResult bl_amd_context_blit_image_d(BLAmdContext* ctx, double x, double y, const BLAmdImage* image, const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!image) return kErrInvalid;
  return blit_d(ctx, x, y, image, area);
}

BLAmdResult bl_amd_context_blit_scaled_image_i(BLAmdContext* ctx, int x, int y, int w, int h, const BLAmdImage* image,
                                               const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!image) return kErrInvalid;
  return blit_scaled_i(ctx, x, y, w, h, image, area);
}

BLAmdResult bl_amd_context_blit_scaled_image_d(BLAmdContext* ctx, double x, double y, double w, double h,
                                               const BLAmdImage* image, const int32_t* area) {
  if (uint32_t r = render_call(ctx)) return r;
  if (!image) return kErrInvalid;
  return blit_scaled_d(ctx, x, y, w, h, image, area);
}

BLAmdResult bl_amd_context_fill_rect_i(BLAmdContext* ctx, int x, int y, int w, int h) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_rect_i(ctx, resolve_fill(ctx), x, y, w, h);
}

BLAmdResult bl_amd_context_fill_rect_i_rgba32(BLAmdContext* ctx, int x, int y, int w, int h, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_rect_i(ctx, resolve_solid(ctx, rgba32), x, y, w, h);
}

BLAmdResult bl_amd_context_fill_all(BLAmdContext* ctx) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_all_box(ctx, resolve_fill(ctx, false));
}

BLAmdResult bl_amd_context_fill_all_rgba32(BLAmdContext* ctx, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_all_box(ctx, resolve_solid(ctx, rgba32, false));
}

BLAmdResult bl_amd_context_clear_all(BLAmdContext* ctx) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_all_box(ctx, resolve_clear(ctx, false));
}

BLAmdResult bl_amd_context_clear_rect_i(BLAmdContext* ctx, int x, int y, int w, int h) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_rect_i(ctx, resolve_clear(ctx, true), x, y, w, h);
}

BLAmdResult bl_amd_context_clear_rect_d(BLAmdContext* ctx, double x, double y, double w, double h) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_box_d(ctx, resolve_clear(ctx, true), x, y, x + w, y + h);
}

// set_style_alpha_impl (rastercontext.cpp:1920-1939)
BLAmdResult bl_amd_context_set_fill_alpha(BLAmdContext* ctx, double alpha) {
  if (uint32_t r = check_target(ctx)) return r;
  if (isnan(alpha)) return kErrInvalid;
  ctx->fill_alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
  return BL_AMD_SUCCESS;
}

// set_style_rgba64_impl (rastercontext.cpp:588-614): the high byte of each
// 16-bit channel (RgbaInternal::rgba32FromRgba64, core/rgba_p.h:67-74).
BLAmdResult bl_amd_context_set_fill_style_rgba64(BLAmdContext* ctx, uint64_t rgba64) {
  return bl_amd_context_set_fill_style_rgba32(ctx, rgba32_from_rgba64(rgba64));
}

// set_style_rgba_impl (rastercontext.cpp:616-655): NaN disables the style;
// channels are clamped, premultiplied in float and rounded half-to-even.
BLAmdResult bl_amd_context_set_fill_style_rgba(BLAmdContext* ctx, const float rgba[4]) {
  if (uint32_t r = check_target(ctx)) return r;
  if (!rgba) return kErrInvalid;
  StyleState& st = ctx->style;
  st = StyleState{};
  st.version = ++ctx->style_counter;
  st.type = kStyleSolid;
  st.dev.kind = kFetchSolid;
  if (isnan(rgba[0]) || isnan(rgba[1]) || isnan(rgba[2]) || isnan(rgba[3])) { st.status = 1; return BL_AMD_SUCCESS; }
  float n[4];
  for (int i = 0; i < 4; i++) n[i] = rgba[i] < 0.0f ? 0.0f : (rgba[i] > 1.0f ? 1.0f : rgba[i]);
  float a_scale = n[3] * 255.0f;
  uint32_t r = (uint32_t)nearbyintf(n[0] * a_scale);
  uint32_t g = (uint32_t)nearbyintf(n[1] * a_scale);
  uint32_t b = (uint32_t)nearbyintf(n[2] * a_scale);
  uint32_t a = (uint32_t)nearbyintf(a_scale);
  uint32_t p = (a << 24) | (r << 16) | (g << 8) | b;
  st.fmt = solid_format(p);
  st.dev.solid = p;
  return BL_AMD_SUCCESS;
}

// fill_unclipped_geometry (rastercontext.cpp:2925-2980) + BLPath::add_geometry
// (core/path.cpp:1396-1640) for the geometry types that do not need the
// arc approximation.
static uint32_t fill_geometry_resolved(BLAmdContext* ctx, const Resolved& r, uint32_t type, const void* data) {
  if (type != BL_AMD_GEOMETRY_NONE && !data) return kErrInvalid;
  switch (type) {
    case BL_AMD_GEOMETRY_NONE:
      return BL_AMD_SUCCESS;
    case BL_AMD_GEOMETRY_BOXI: {
      const int32_t* b = (const int32_t*)data;
      return fill_box_d(ctx, r, (double)b[0], (double)b[1], (double)b[2], (double)b[3]);
    }
    case BL_AMD_GEOMETRY_BOXD: {
      const double* b = (const double*)data;
      return fill_box_d(ctx, r, b[0], b[1], b[2], b[3]);
    }
    case BL_AMD_GEOMETRY_RECTI: {
      const int32_t* b = (const int32_t*)data;
      return fill_rect_i(ctx, r, b[0], b[1], b[2], b[3]);
    }
    case BL_AMD_GEOMETRY_RECTD: {
      const double* b = (const double*)data;
      return fill_box_d(ctx, r, b[0], b[1], b[0] + b[2], b[1] + b[3]);
    }
    case BL_AMD_GEOMETRY_POLYGONI:
    case BL_AMD_GEOMETRY_POLYLINEI:
    case BL_AMD_GEOMETRY_POLYGOND:
    case BL_AMD_GEOMETRY_POLYLINED: {
      const BLAmdArrayView* v = (const BLAmdArrayView*)data;
      if (v->size < 3) return BL_AMD_SUCCESS;
      if (!v->data) return kErrInvalid;
      if (type == BL_AMD_GEOMETRY_POLYGOND || type == BL_AMD_GEOMETRY_POLYLINED)
        return fill_poly_rule(ctx, r, (const double*)v->data, v->size, ctx->fill_rule);
      std::vector<double> pts(2 * v->size);
      const int32_t* src = (const int32_t*)v->data;
      for (size_t i = 0; i < 2 * v->size; i++) pts[i] = (double)src[i];
      return fill_poly_rule(ctx, r, pts.data(), v->size, ctx->fill_rule);
    }
    default:
      break;
  }

  TmpPath p;
  switch (type) {
    case BL_AMD_GEOMETRY_CIRCLE:
    case BL_AMD_GEOMETRY_ELLIPSE: {
      const double* e = (const double*)data;
      double x0 = e[0], y0 = e[1], rx, ry;
      if (type == BL_AMD_GEOMETRY_CIRCLE) { rx = e[2]; ry = fabs(rx); }
      else { rx = e[2]; ry = e[3]; }
      double kx = rx * kKappa, ky = ry * kKappa;
      p.move_to(x0 + rx, y0);
      p.cubic_to(x0 + rx, y0 + ky, x0 + kx, y0 + ry, x0, y0 + ry);
      p.cubic_to(x0 - kx, y0 + ry, x0 - rx, y0 + ky, x0 - rx, y0);
      p.cubic_to(x0 - rx, y0 - ky, x0 - kx, y0 - ry, x0, y0 - ry);
      p.cubic_to(x0 + kx, y0 - ry, x0 + rx, y0 - ky, x0 + rx, y0);
      p.close();
      break;
    }
    case BL_AMD_GEOMETRY_ROUND_RECT: {
      const double* q = (const double*)data;
      double x0 = q[0], y0 = q[1], x1 = q[0] + q[2], y1 = q[1] + q[3];
      double rx = std::min(fabs(q[4]), q[2] * 0.5);
      double ry = std::min(fabs(q[5]), q[3] * 0.5);
      if (!(rx > 1e-14 && ry > 1e-14)) { p.box(x0, y0, x1, y1); break; }
      double kx = rx * (1.0 - kKappa), ky = ry * (1.0 - kKappa);
      p.move_to(x0 + rx, y0);
      p.line_to(x1 - rx, y0);
      p.cubic_to(x1 - kx, y0, x1, y0 + ky, x1, y0 + ry);
      p.line_to(x1, y1 - ry);
      p.cubic_to(x1, y1 - ky, x1 - kx, y1, x1 - rx, y1);
      p.line_to(x0 + rx, y1);
      p.cubic_to(x0 + kx, y1, x0, y1 - ky, x0, y1 - ry);
      p.line_to(x0, y0 + ry);
      p.cubic_to(x0, y0 + ky, x0 + kx, y0, x0 + rx, y0);
      p.close();
      break;
    }
    case BL_AMD_GEOMETRY_LINE: {
      const double* l = (const double*)data;
      p.move_to(l[0], l[1]);
      p.line_to(l[2], l[3]);
      break;
    }
    case BL_AMD_GEOMETRY_TRIANGLE: {
      const double* t = (const double*)data;
      p.move_to(t[0], t[1]);
      p.line_to(t[2], t[3]);
      p.line_to(t[4], t[5]);
      p.close();
      break;
    }
    case BL_AMD_GEOMETRY_ARRAY_VIEW_BOXI:
    case BL_AMD_GEOMETRY_ARRAY_VIEW_BOXD:
    case BL_AMD_GEOMETRY_ARRAY_VIEW_RECTI:
    case BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD: {
      const BLAmdArrayView* v = (const BLAmdArrayView*)data;
      if (!v->size) return BL_AMD_SUCCESS;
      if (!v->data) return kErrInvalid;
      // items that are not valid boxes / rects are skipped (geometry/commons_p.h:42-56)
      for (size_t i = 0; i < v->size; i++) {
        if (type == BL_AMD_GEOMETRY_ARRAY_VIEW_BOXI) {
          const int32_t* b = (const int32_t*)v->data + 4 * i;
          if (b[0] < b[2] && b[1] < b[3]) p.box((double)b[0], (double)b[1], (double)b[2], (double)b[3]);
        }
        else if (type == BL_AMD_GEOMETRY_ARRAY_VIEW_BOXD) {
          const double* b = (const double*)v->data + 4 * i;
          if (b[0] < b[2] && b[1] < b[3]) p.box(b[0], b[1], b[2], b[3]);
        }
        else if (type == BL_AMD_GEOMETRY_ARRAY_VIEW_RECTI) {
          const int32_t* b = (const int32_t*)v->data + 4 * i;
          int64_t x1 = (int64_t)b[0] + b[2], y1 = (int64_t)b[1] + b[3];
          bool overflow = x1 != (int32_t)x1 || y1 != (int32_t)y1;
          if (overflow || !(b[0] < x1 && b[1] < y1)) continue;
          double x0 = (double)b[0], y0 = (double)b[1];
          p.box(x0, y0, (double)b[2] + x0, (double)b[3] + y0);
        }
        else {
          const double* b = (const double*)v->data + 4 * i;
          double x1 = b[0] + b[2], y1 = b[1] + b[3];
          if (b[0] < x1 && b[1] < y1) p.box(b[0], b[1], b[2] + b[0], b[3] + b[1]);
        }
      }
      break;
    }
    case BL_AMD_GEOMETRY_ARC:
    case BL_AMD_GEOMETRY_CHORD:
    case BL_AMD_GEOMETRY_PIE:
      return r.nop ? BL_AMD_SUCCESS : kErrNotImpl;
    default:
      return kErrInvalid;
  }
  return fill_path_rule(ctx, r, nullptr, p.cmd.data(), p.vtx.data(), p.cmd.size(), ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_geometry(BLAmdContext* ctx, uint32_t type, const void* data) {
  if (uint32_t r = render_call(ctx)) return r;
  if (type > BL_AMD_GEOMETRY_MAX_VALUE) return kErrInvalid;
  return fill_geometry_resolved(ctx, resolve_fill(ctx), type, data);
}

BLAmdResult bl_amd_context_fill_geometry_rgba32(BLAmdContext* ctx, uint32_t type, const void* data, uint32_t rgba32) {
  if (uint32_t r = render_call(ctx)) return r;
  if (type > BL_AMD_GEOMETRY_MAX_VALUE) return kErrInvalid;
  return fill_geometry_resolved(ctx, resolve_solid(ctx, rgba32), type, data);
}

// _rgba64 forms: an inline rgba64 style (core/context.cpp:853-859 builds it
// and calls the _ext entry), converted to rgba32 by its high bytes.
BLAmdResult bl_amd_context_fill_all_rgba64(BLAmdContext* ctx, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_all_box(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64), false));
}

BLAmdResult bl_amd_context_fill_rect_i_rgba64(BLAmdContext* ctx, int x, int y, int w, int h, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_rect_i(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), x, y, w, h);
}

BLAmdResult bl_amd_context_fill_rect_d_rgba64(BLAmdContext* ctx, double x, double y, double w, double h, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  return fill_box_d(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), x, y, x + w, y + h);
}

BLAmdResult bl_amd_context_fill_path_d_rgba64(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                              const double* vertices, size_t n, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  if (n == 0) return BL_AMD_SUCCESS;
  if (!cmds || !vertices) return kErrInvalid;
  return fill_path_rule(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), origin, cmds, vertices, n, ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_geometry_rgba64(BLAmdContext* ctx, uint32_t type, const void* data, uint64_t rgba64) {
  if (uint32_t r = render_call(ctx)) return r;
  if (type > BL_AMD_GEOMETRY_MAX_VALUE) return kErrInvalid;
  return fill_geometry_resolved(ctx, resolve_solid(ctx, rgba32_from_rgba64(rgba64)), type, data);
}

// _ext forms: the call's own style (rastercontext.cpp:1277-1328).
BLAmdResult bl_amd_context_fill_all_ext(BLAmdContext* ctx, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, false, err);
  return err ? err : fill_all_box(ctx, res);
}

BLAmdResult bl_amd_context_fill_rect_i_ext(BLAmdContext* ctx, int x, int y, int w, int h, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_rect_i(ctx, res, x, y, w, h);
}

BLAmdResult bl_amd_context_fill_rect_d_ext(BLAmdContext* ctx, double x, double y, double w, double h, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_box_d(ctx, res, x, y, x + w, y + h);
}

BLAmdResult bl_amd_context_fill_path_d_ext(BLAmdContext* ctx, const double* origin, const uint8_t* cmds,
                                           const double* vertices, size_t n, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  if (n == 0) return BL_AMD_SUCCESS;
  if (!cmds || !vertices) return kErrInvalid;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_path_rule(ctx, res, origin, cmds, vertices, n, ctx->fill_rule);
}

BLAmdResult bl_amd_context_fill_geometry_ext(BLAmdContext* ctx, uint32_t type, const void* data, const BLAmdStyle* style) {
  if (uint32_t r = render_call(ctx)) return r;
  if (type > BL_AMD_GEOMETRY_MAX_VALUE) return kErrInvalid;
  uint32_t err;
  const Resolved res = resolve_explicit(ctx, style, true, err);
  return err ? err : fill_geometry_resolved(ctx, res, type, data);
}

// bl_context_reset (core/context.cpp:263-267): the context becomes a default
// one; like end() (which the reference maps to reset, :329-332) the pending
// calls are rendered first when a target is attached.
BLAmdResult bl_amd_context_reset(BLAmdContext* ctx) {
  if (!ctx) return kErrInvalid;
  uint32_t r = BL_AMD_SUCCESS;
  if (ctx->target) r = bl_amd_context_end(ctx);
  ctx->reset_state();
  reset_frame(ctx);
  ctx->call_seq = 0;
  ctx->error_flags = 0;
  ctx->deferred_call = ~0ull;
  ctx->deferred_code = 0;
  return r;
}

static void copy_stats(const Stats& st, BLAmdStats* out) {
  out->commands = st.cmds;
  out->sources = st.src;
  out->segments = st.segs;
  out->binned = st.binned;
  out->tiles = st.tiles;
  out->pixels = st.pixels;
  out->frame_bytes = st.frame_bytes;
  out->lut_bytes = st.lut_bytes;
  out->image_bytes = st.image_bytes;
  out->pred_segs = st.pred_segs;
  out->pred_bins = st.pred_bins;
  out->sizing = st.sizing;
  out->fallback = st.fallback;
  out->syncs = st.syncs;
  out->render_kernels = st.render_kernels;
  out->merge_ms = st.merge_ms;
  out->upload_ms = st.upload_ms;
  out->pack_ms = st.pack_ms;
  out->render_ms = st.render_ms;
  out->total_ms = st.total_ms;
  out->schedule = st.schedule;
  out->queue_wgs = st.queue_wgs;
  out->graph = st.graph;
}

BLAmdResult bl_amd_context_get_error_flags(const BLAmdContext* ctx, uint32_t* flags) {
  if (!ctx || !flags) return kErrInvalid;
  *flags = ctx->error_flags;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_get_deferred_error(const BLAmdContext* ctx, uint64_t* call_index, uint32_t* code) {
  if (!ctx || !call_index || !code) return kErrInvalid;
  *call_index = ctx->deferred_code ? ctx->deferred_call : 0;
  *code = ctx->deferred_code;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_get_stats(const BLAmdContext* ctx, BLAmdStats* out) {
  if (!ctx || !out) return kErrInvalid;
  copy_stats(ctx->stats, out);
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_replay(BLAmdContext* ctx) {
  if (uint32_t r = check_target(ctx)) return r;
  return ctx->runner.replay(ctx->stats);
}

BLAmdResult bl_amd_context_replay_noalloc(BLAmdContext* ctx) {
  if (uint32_t r = check_target(ctx)) return r;
  if (!ctx->runner.resident) return BL_AMD_SUCCESS;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kErrDevice;
  ScratchArena arena;
  size_t bytes = 0;
  if (noalloc_scratch(dev, (void**)&arena.base, &bytes) != hipSuccess) return kErrDevice;
  arena.cap = bytes;
  return ctx->runner.replay_in(ctx->stats, arena);
}

BLAmdResult bl_amd_batch_create(BLAmdBatch** out) {
  if (!out) return kErrInvalid;
  *out = new (std::nothrow) BLAmdBatch();
  return *out ? BL_AMD_SUCCESS : kErrOOM;
}

BLAmdResult bl_amd_batch_destroy(BLAmdBatch* b) {
  delete b;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_batch_add(BLAmdBatch* b, BLAmdContext* ctx) {
  if (!b || !ctx || !ctx->target) return kErrInvalid;
  // Every context of a batch is one canvas of one frame, and the canvases
  // are composited concurrently: a context added twice, or two contexts
  // rendering into the same image, would race on its pixels (the reference
  // renders each context's calls in order).  Such a context is rejected; it
  // can be flushed on its own.
  for (const BLAmdContext* c : b->contexts)
    if (c == ctx || c->target == ctx->target) return kErrInvalid;
  b->contexts.push_back(ctx);
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_batch_clear(BLAmdBatch* b) {
  if (!b) return kErrInvalid;
  b->contexts.clear();
  return BL_AMD_SUCCESS;
}

// Merges the pending frames of every context into one frame (indices
// rebased) and renders them with one pass of each kernel stage.
BLAmdResult bl_amd_batch_flush(BLAmdBatch* b) {
  if (!b) return kErrInvalid;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<FrameHost*> parts;
  parts.reserve(b->contexts.size());
  for (BLAmdContext* c : b->contexts) {
    if (!c->target) return kErrState;
    finalize_frame(c);
    parts.push_back(&c->frame);
  }
  const double merge_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  uint32_t r = b->runner.run(parts, b->stats, merge_ms);
  attribute_failures(b->runner, b->contexts);
  size_t n_cmds = 0;
  for (const BLAmdContext* c : b->contexts) n_cmds += c->frame.cmds.size();
  if (n_cmds) {
    b->resident_images.clear();
    for (const BLAmdContext* c : b->contexts) keep_resident(b->resident_images, c);
  }
  for (BLAmdContext* c : b->contexts) reset_frame(c);
  return r;
}

BLAmdResult bl_amd_batch_replay(BLAmdBatch* b) {
  if (!b) return kErrInvalid;
  return b->runner.replay(b->stats);
}

// ---------------------------------------------------------------------------
// Learned sizing: cost model
// ---------------------------------------------------------------------------

// Both settings are read by flushes under the device queue's mutex.
BLAmdResult bl_amd_set_sizing(uint32_t mode) {
  if (mode > kSizeHistory) return kErrInvalid;
  std::lock_guard<std::mutex> lock(device_queue().mu);
  g_sizing = mode;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_cost_model_set(const float* params) {
  std::lock_guard<std::mutex> lock(device_queue().mu);
  if (params) g_model.assign(params, params + cm::kParams);
  else g_model.clear();
  g_model_version++;
  return BL_AMD_SUCCESS;
}

// ---------------------------------------------------------------------------
// Tile scheduling, tile cost model, autotuner
// ---------------------------------------------------------------------------

BLAmdResult bl_amd_set_graph_replay(uint32_t on) {
  std::lock_guard<std::mutex> lock(device_queue().mu);
  g_graph_replay = on ? 1u : 0u;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_set_render_schedule(uint32_t schedule, uint32_t wgs_per_cu) {
  if (schedule > BL_AMD_SCHEDULE_AUTO || wgs_per_cu > 4) return kErrInvalid;
  std::lock_guard<std::mutex> lock(device_queue().mu);
  g_schedule = schedule;
  g_queue_wgs = wgs_per_cu ? wgs_per_cu : 4;
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_tile_model_set(const float* params) {
  std::lock_guard<std::mutex> lock(device_queue().mu);
  if (params) g_tile_model.assign(params, params + cm::kParams);
  else g_tile_model.clear();
  g_tile_model_version++;
  return BL_AMD_SUCCESS;
}

}  // extern "C"

namespace {

uint32_t collect_tiles_of(Runner& rn, Stats& stats, float* features, float* targets, size_t cap, size_t* n_out) {
  if (!n_out) return kErrInvalid;
  rn.collect_tiles = true;
  uint32_t r = rn.replay(stats);
  rn.collect_tiles = false;
  if (r) return r;
  const size_t rows = rn.tile_t.size() / 2;
  const size_t k = std::min(rows, cap);
  if (k && (!features || !targets)) return kErrInvalid;
  if (k) {
    memcpy(features, rn.tile_x.data(), k * 16 * sizeof(float));
    memcpy(targets, rn.tile_t.data(), k * 2 * sizeof(float));
  }
  *n_out = rows;
  return BL_AMD_SUCCESS;
}

const char* schedule_name(uint32_t s) {
  switch (s) {
    case BL_AMD_SCHEDULE_GRID: return "grid";
    case BL_AMD_SCHEDULE_QUEUE: return "queue";
    case BL_AMD_SCHEDULE_QUEUE_HEURISTIC: return "queue_heuristic";
    case BL_AMD_SCHEDULE_QUEUE_LEARNED: return "queue_learned";
    default: return "auto";
  }
}

// Times the render stage of the runner's resident frame under every
// candidate schedule (best of `reps` replays), records the fastest for the
// frame's shape and logs every candidate's time.
uint32_t autotune_runner(Runner& rn, Stats& last, uint32_t reps, BLAmdTune* out) {
  if (!out) return kErrInvalid;
  if (!rn.resident) return kErrInvalid;
  reps = std::max<uint32_t>(reps, 1);
  bool have_model;
  {
    std::lock_guard<std::mutex> lock(device_queue().mu);
    have_model = !g_tile_model.empty();
  }
  std::vector<TunedSchedule> cands = {{BL_AMD_SCHEDULE_GRID, 0}, {BL_AMD_SCHEDULE_QUEUE, 4}};
  for (uint32_t w = 2; w <= 4; w++) cands.push_back({BL_AMD_SCHEDULE_QUEUE_HEURISTIC, w});
  if (have_model)
    for (uint32_t w = 2; w <= 4; w++) cands.push_back({BL_AMD_SCHEDULE_QUEUE_LEARNED, w});
  std::vector<double> times(cands.size(), 1e30);
  uint32_t r = BL_AMD_SUCCESS;
  for (size_t c = 0; c < cands.size() && !r; c++) {
    rn.force_schedule = cands[c].schedule;
    rn.force_wgs = cands[c].wgs;
    for (uint32_t k = 0; k < reps && !r; k++) {
      Stats st;
      r = rn.replay(st);
      if (!r) times[c] = std::min(times[c], st.render_ms);
      last = st;
    }
  }
  rn.force_schedule = BL_AMD_SCHEDULE_AUTO;
  rn.force_wgs = 0;
  if (r) return r;
  size_t best = 0;
  for (size_t c = 1; c < cands.size(); c++)
    if (times[c] < times[best]) best = c;
  {
    std::lock_guard<std::mutex> lock(device_queue().mu);
    const uint64_t key = rn.shape_key();
    g_tuned[key] = cands[best];
    char line[160];
    snprintf(line, sizeof(line), "shape tiles~2^%u cmds~2^%u canvases~2^%u:", (unsigned)(key >> 40),
             (unsigned)((key >> 20) & 0xFFFFF), (unsigned)(key & 0xFFFFF));
    g_tune_log += line;
    for (size_t c = 0; c < cands.size(); c++) {
      snprintf(line, sizeof(line), " %s/%u %.3f", schedule_name(cands[c].schedule), cands[c].wgs, times[c]);
      g_tune_log += line;
    }
    snprintf(line, sizeof(line), " ms -> %s/%u\n", schedule_name(cands[best].schedule), cands[best].wgs);
    g_tune_log += line;
  }
  out->schedule = cands[best].schedule;
  out->wgs_per_cu = cands[best].wgs;
  out->candidates = (uint32_t)cands.size();
  out->reserved = 0;
  out->best_ms = times[best];
  out->grid_ms = times[0];
  return BL_AMD_SUCCESS;
}

}  // namespace

extern "C" {

BLAmdResult bl_amd_context_collect_tile_costs(BLAmdContext* ctx, float* features, float* targets, size_t cap, size_t* n_out) {
  if (!ctx) return kErrInvalid;
  if (uint32_t r = check_target(ctx)) return r;
  return collect_tiles_of(ctx->runner, ctx->stats, features, targets, cap, n_out);
}

BLAmdResult bl_amd_batch_collect_tile_costs(BLAmdBatch* b, float* features, float* targets, size_t cap, size_t* n_out) {
  if (!b) return kErrInvalid;
  return collect_tiles_of(b->runner, b->stats, features, targets, cap, n_out);
}

BLAmdResult bl_amd_context_autotune(BLAmdContext* ctx, uint32_t reps, BLAmdTune* out) {
  if (!ctx) return kErrInvalid;
  if (uint32_t r = check_target(ctx)) return r;
  return autotune_runner(ctx->runner, ctx->stats, reps, out);
}

BLAmdResult bl_amd_batch_autotune(BLAmdBatch* b, uint32_t reps, BLAmdTune* out) {
  if (!b) return kErrInvalid;
  return autotune_runner(b->runner, b->stats, reps, out);
}

size_t bl_amd_autotune_log(char* buf, size_t cap) {
  std::lock_guard<std::mutex> lock(device_queue().mu);
  if (buf && cap) {
    const size_t k = std::min(cap - 1, g_tune_log.size());
    memcpy(buf, g_tune_log.data(), k);
    buf[k] = 0;
  }
  return g_tune_log.size();
}

BLAmdResult bl_amd_autotune_reset(void) {
  std::lock_guard<std::mutex> lock(device_queue().mu);
  g_tuned.clear();
  g_tune_log.clear();
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_context_collect_costs(BLAmdContext* ctx, float* features, float* targets, size_t cap, size_t* n_out) {
  if (!ctx || !n_out) return kErrInvalid;
  if (uint32_t r = check_target(ctx)) return r;
  ctx->runner.collect = true;
  uint32_t r = ctx->runner.replay(ctx->stats);
  ctx->runner.collect = false;
  if (r) return r;
  const size_t rows = ctx->runner.collected_t.size() / cm::kOut;
  const size_t k = std::min(rows, cap);
  if (k && (!features || !targets)) return kErrInvalid;
  if (k) {
    memcpy(features, ctx->runner.collected_x.data(), k * cm::kIn * sizeof(float));
    memcpy(targets, ctx->runner.collected_t.data(), k * cm::kOut * sizeof(float));
  }
  *n_out = rows;
  return BL_AMD_SUCCESS;
}

}  // extern "C"

namespace {

// Device state for the MLP entry points (training and tests).
struct Trainer {
  hipStream_t stream = nullptr;
  DevBuf x, t, y, params, partial, grad, m1, m2;
  float* host = nullptr;
  size_t host_cap = 0;
  bool init() {
    if (stream) return true;
    return hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess;
  }
  ~Trainer() {
    if (host) (void)hipHostFree(host);
    if (stream) (void)hipStreamDestroy(stream);
  }
  bool host_reserve(size_t n) {
    if (n <= host_cap) return true;
    if (host) (void)hipHostFree(host);
    host = nullptr;
    host_cap = 0;
    if (hipHostMalloc((void**)&host, n * sizeof(float), hipHostMallocDefault) != hipSuccess) return false;
    host_cap = n;
    return true;
  }
};

Trainer& trainer() {
  static Trainer t;
  return t;
}

}  // namespace

extern "C" {

BLAmdResult bl_amd_cost_model_forward(const float* X, size_t n, const float* params, float* Y) {
  if (!params || (n && (!X || !Y)) || n > 0x7FFFFFFFu) return kErrInvalid;
  if (!n) return BL_AMD_SUCCESS;
  Trainer& tr = trainer();
  if (!tr.init()) return kErrDevice;
  if (!tr.x.reserve(n * cm::kIn * 4) || !tr.y.reserve(n * cm::kOut * 4) || !tr.params.reserve(cm::kParams * 4)) return kErrOOM;
  HIP_OK(hipMemcpyAsync(tr.x.ptr, X, n * cm::kIn * 4, hipMemcpyHostToDevice, tr.stream));
  HIP_OK(hipMemcpyAsync(tr.params.ptr, params, cm::kParams * 4, hipMemcpyHostToDevice, tr.stream));
  cm::launch_mlp_forward(tr.x.as<float>(), (uint32_t)n, tr.params.as<float>(), tr.y.as<float>(), tr.stream);
  HIP_OK(hipMemcpyAsync(Y, tr.y.ptr, n * cm::kOut * 4, hipMemcpyDeviceToHost, tr.stream));
  HIP_OK(hipStreamSynchronize(tr.stream));
  return hipGetLastError() == hipSuccess ? BL_AMD_SUCCESS : kErrDevice;
}

BLAmdResult bl_amd_cost_model_grad(const float* X, const float* T, size_t n, const float* params, float* grad, float* loss) {
  if (!params || !grad || !n || !X || !T || n > 0x7FFFFFFFu) return kErrInvalid;
  Trainer& tr = trainer();
  if (!tr.init()) return kErrDevice;
  const uint32_t groups = cm::mlp_groups((uint32_t)n);
  if (!tr.x.reserve(n * cm::kIn * 4) || !tr.t.reserve(n * cm::kOut * 4) || !tr.params.reserve(cm::kParams * 4) ||
      !tr.partial.reserve((size_t)groups * (cm::kParams + 1) * 4) || !tr.grad.reserve((cm::kParams + 1) * 4))
    return kErrOOM;
  HIP_OK(hipMemcpyAsync(tr.x.ptr, X, n * cm::kIn * 4, hipMemcpyHostToDevice, tr.stream));
  HIP_OK(hipMemcpyAsync(tr.t.ptr, T, n * cm::kOut * 4, hipMemcpyHostToDevice, tr.stream));
  HIP_OK(hipMemcpyAsync(tr.params.ptr, params, cm::kParams * 4, hipMemcpyHostToDevice, tr.stream));
  cm::launch_mlp_grad(tr.x.as<float>(), tr.t.as<float>(), (uint32_t)n, tr.params.as<float>(), tr.partial.as<float>(), tr.stream);
  cm::launch_mlp_reduce(tr.partial.as<float>(), groups, (uint32_t)n, tr.grad.as<float>(), tr.stream);
  std::vector<float> g(cm::kParams + 1);
  HIP_OK(hipMemcpyAsync(g.data(), tr.grad.ptr, g.size() * 4, hipMemcpyDeviceToHost, tr.stream));
  HIP_OK(hipStreamSynchronize(tr.stream));
  if (hipError_t e_ = hipGetLastError(); e_ != hipSuccess) { report_hip("kernel launch", e_, __LINE__); return kErrDevice; }
  memcpy(grad, g.data(), cm::kParams * 4);
  if (loss) *loss = g[cm::kParams];
  return BL_AMD_SUCCESS;
}

BLAmdResult bl_amd_cost_model_train(const float* X, const float* T, size_t n, float* params, uint32_t epochs, uint32_t batch,
                                    float lr, float* epoch_loss) {
  if (!X || !T || !params || !n || !batch || n > 0x7FFFFFFFu || batch > (1u << 20)) return kErrInvalid;
  Trainer& tr = trainer();
  if (!tr.init()) return kErrDevice;
  const uint32_t groups = cm::mlp_groups(batch);
  const size_t n_batches = (n + batch - 1) / batch;
  if (!tr.x.reserve(n * cm::kIn * 4) || !tr.t.reserve(n * cm::kOut * 4) || !tr.params.reserve(cm::kParams * 4) ||
      !tr.partial.reserve((size_t)groups * (cm::kParams + 1) * 4) || !tr.grad.reserve((cm::kParams + 1) * 4) ||
      !tr.m1.reserve(cm::kParams * 4) || !tr.m2.reserve(cm::kParams * 4) || !tr.host_reserve(n_batches))
    return kErrOOM;
  HIP_OK(hipMemcpyAsync(tr.x.ptr, X, n * cm::kIn * 4, hipMemcpyHostToDevice, tr.stream));
  HIP_OK(hipMemcpyAsync(tr.t.ptr, T, n * cm::kOut * 4, hipMemcpyHostToDevice, tr.stream));
  HIP_OK(hipMemcpyAsync(tr.params.ptr, params, cm::kParams * 4, hipMemcpyHostToDevice, tr.stream));
  launch_fill_u32(tr.m1.as<uint32_t>(), 0u, cm::kParams, tr.stream);   // 0.0f
  launch_fill_u32(tr.m2.as<uint32_t>(), 0u, cm::kParams, tr.stream);
  uint32_t step = 0;
  for (uint32_t e = 0; e < epochs; e++) {
    for (size_t b = 0; b < n_batches; b++) {
      const size_t start = b * batch;
      const uint32_t rows = (uint32_t)std::min<size_t>(batch, n - start);
      const float* xb = tr.x.as<float>() + start * cm::kIn;
      const float* tb = tr.t.as<float>() + start * cm::kOut;
      cm::launch_mlp_grad(xb, tb, rows, tr.params.as<float>(), tr.partial.as<float>(), tr.stream);
      cm::launch_mlp_reduce(tr.partial.as<float>(), cm::mlp_groups(rows), rows, tr.grad.as<float>(), tr.stream);
      cm::launch_adam(tr.params.as<float>(), tr.grad.as<float>(), tr.m1.as<float>(), tr.m2.as<float>(), ++step, lr, tr.stream);
      HIP_OK(hipMemcpyAsync(tr.host + b, tr.grad.as<float>() + cm::kParams, 4, hipMemcpyDeviceToHost, tr.stream));
    }
    HIP_OK(hipStreamSynchronize(tr.stream));
    if (epoch_loss) {
      double acc = 0.0;
      for (size_t b = 0; b < n_batches; b++) {
        const size_t rows = std::min<size_t>(batch, n - b * batch);
        acc += (double)tr.host[b] * (double)rows;
      }
      epoch_loss[e] = (float)(acc / (double)n);
    }
  }
  HIP_OK(hipMemcpyAsync(params, tr.params.ptr, cm::kParams * 4, hipMemcpyDeviceToHost, tr.stream));
  HIP_OK(hipStreamSynchronize(tr.stream));
  return hipGetLastError() == hipSuccess ? BL_AMD_SUCCESS : kErrDevice;
}

BLAmdResult bl_amd_batch_get_stats(const BLAmdBatch* b, BLAmdStats* out) {
  if (!b || !out) return kErrInvalid;
  copy_stats(b->stats, out);
  return BL_AMD_SUCCESS;
}

}  // extern "C"
