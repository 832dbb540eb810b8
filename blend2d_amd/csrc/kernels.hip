// Fill-path kernels for gfx950.
//
//   edges    one thread per source segment (line / quad / cubic): transform,
//            clip, split, flatten -> 24.8 edge segments (count pass, then an
//            emit pass at prefix-scanned offsets).
//   bin      per segment: count, scatter into (band, command) lists; per band
//            a compacted, submission-ordered list of the commands touching it.
//   render   one 256-thread workgroup per 64x16 tile: the destination tile is
//            loaded once into registers (128-bit per lane), every command of
//            the band is rasterized into LDS cells with LDS atomics, turned
//            into coverage by a per-scanline DPP scan, fetched, composited
//            in submission order, and the tile is stored once.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_types.h"
#include "edge_device.h"
#include "kernels.h"
#include "pipe_device.h"
#include "raster_device.h"

#include <mutex>

namespace b2amd {

// ---------------------------------------------------------------------------
// Scan (exclusive, uint32) -- three-level: block sums, scan of sums, apply.
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total in `total`.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  uint32_t incl = wave_incl_scan(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; w++) {
    uint32_t s = wsum[w];
    if (w < wave) wprefix += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wprefix + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t* __restrict__ in, size_t n, uint32_t* __restrict__ sums) {
  size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanItems;
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) if (base + k < n) v += in[base + k];
  uint32_t total;
  block_excl_scan(v, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n,
                                                          const uint32_t* __restrict__ prefix) {
  size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanItems;
  uint32_t x[kScanItems];
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) { x[k] = (base + k < n) ? in[base + k] : 0u; v += x[k]; }
  uint32_t total;
  uint32_t excl = block_excl_scan(v, total);
  uint32_t p = (prefix ? prefix[blockIdx.x] : 0u) + excl;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = p;
    p += x[k];
  }
  // The element one past the end receives the grand total, written by the
  // thread that owns the last element (thread 0 of block 0 when n == 0).
  const bool owner = n ? (base <= n - 1 && n - 1 < base + kScanItems) : base == 0;
  if (owner) {
    uint32_t tail = (prefix ? prefix[blockIdx.x] : 0u) + excl;
    for (int k = 0; k < kScanItems && base + k < n; k++) tail += x[k];
    out[n] = tail;
  }
}

size_t scan_tmp_words(size_t n) {
  size_t words = 0;
  while (n > (size_t)kScanBlock) {
    size_t nb = (n + kScanBlock - 1) / kScanBlock;
    words += 2 * (nb + 1);
    n = nb;
  }
  return words + 64;
}

void scan_exclusive(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp, hipStream_t s) {
  if (n == 0) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanThreads), 0, s, in, out, (size_t)0, (const uint32_t*)nullptr);
    return;
  }
  size_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanThreads), 0, s, in, out, n, (const uint32_t*)nullptr);
    return;
  }
  uint32_t* sums = tmp;
  uint32_t* sums_scan = tmp + (nb + 1);
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, sums);
  scan_exclusive(sums, sums_scan, nb, tmp + 2 * (nb + 1), s);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n, (const uint32_t*)sums_scan);
}

// ---------------------------------------------------------------------------
// Utilities
// ---------------------------------------------------------------------------

// Scratch of the execution that allocates nothing
// (bl_amd_context_replay_noalloc): every intermediate buffer of a frame is a
// slice of this region.  Each device holds its own copy for the life of the
// process; every call on a device shares it, so calls that overlap on one
// device race (the library's flushes and replays are serialized by its
// device queue).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int dev, void** ptr, size_t* bytes) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    // the array's host-side name is not a device address: look the address
    // up on this device (current) and use only that
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = kNoAllocScratchBytes;
  return hipSuccess;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

__global__ void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

void launch_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
}

// One gradient table per workgroup, every entry computed on its own
// (lut_device.h).  The stops of a job were validated and sorted on the
// host (offsets in [0, 1], n >= 1), so every position lies in [0, size).
__global__ void __launch_bounds__(256) k_gradient_luts(const lut::Job* __restrict__ jobs,
                                                       const lut::Stop* __restrict__ stops,
                                                       uint32_t* __restrict__ luts) {
  const lut::Job j = jobs[blockIdx.x];
  const lut::Stop* st = stops + j.stop0;
  if (j.wide) {
    for (uint32_t e = threadIdx.x; e < j.size; e += 256u) {
      const uint64_t v = lut::entry64(st, j.n_stops, j.size, e);
      luts[(size_t)j.lut + 2u * e] = (uint32_t)v;
      luts[(size_t)j.lut + 2u * e + 1u] = (uint32_t)(v >> 32);
    }
    return;
  }
  for (uint32_t e = threadIdx.x; e < j.size; e += 256u) luts[(size_t)j.lut + e] = lut::entry(st, j.n_stops, j.size, e);
}

void launch_gradient_luts(const lut::Job* jobs, const lut::Stop* stops, uint32_t n_jobs, uint32_t* luts, hipStream_t st) {
  if (!n_jobs) return;
  hipLaunchKernelGGL(k_gradient_luts, dim3(n_jobs), dim3(256), 0, st, jobs, stops, luts);
}

__global__ void k_init_bounds(CmdBounds* b, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { b[i].min_x = INT32_MAX; b[i].max_x = INT32_MIN; b[i].min_y = INT32_MAX; b[i].max_y = INT32_MIN; }
}

void launch_init_bounds(CmdBounds* b, uint32_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_init_bounds, dim3((n + 255) / 256), dim3(256), 0, st, b, n);
}

// ---------------------------------------------------------------------------
// Stage 1: edges
// ---------------------------------------------------------------------------

struct CountEmit {
  uint32_t n = 0;
  int32_t min_x = INT32_MAX, max_x = INT32_MIN, min_y = INT32_MAX, max_y = INT32_MIN;
  __device__ __forceinline__ void push(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t) {
    n++;
    int32_t lo = (x0 < x1 ? x0 : x1) >> 8, hi = (x0 < x1 ? x1 : x0) >> 8;
    min_x = lo < min_x ? lo : min_x;
    max_x = hi > max_x ? hi : max_x;
    min_y = y0 < min_y ? y0 : min_y;
    max_y = y1 > max_y ? y1 : max_y;
  }
};

struct WriteEmit {
  DevSeg* segs;
  uint32_t* seg_cmd;
  uint32_t pos;
  uint32_t cmd;
  __device__ __forceinline__ void push(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign) {
    segs[pos] = DevSeg{x0, y0, (int32_t)((uint32_t)x1 | (sign << 31)), y1};
    seg_cmd[pos] = cmd;
    pos++;
  }
};

__device__ __forceinline__ edge::P xform_pt(const CmdXform& m, const double* vtx, uint32_t i) {
  double x = vtx[2 * i], y = vtx[2 * i + 1];
  if (m.scale_only) return edge::P{x * m.m00 + m.m20, y * m.m11 + m.m21};
  return edge::P{x * m.m00 + y * m.m10 + m.m20, x * m.m01 + y * m.m11 + m.m21};
}

// Edge work items: a line, one monotonic piece of a curve, or the border line
// of a curve lying wholly left or right of the clip box.
// Splitting curves into their pieces up front gives the flattening kernels
// several independent, shorter work items per curve.
enum ItemKind : uint32_t { kItemLine = 0, kItemQuad = 1, kItemCubic = 2, kItemBorderLeft = 3, kItemBorderRight = 4 };

__device__ __forceinline__ edge::Clip canvas_clip(const DevCanvas& cv) {
  edge::Clip clip;
  clip.x0 = 0.0; clip.y0 = 0.0;
  clip.x1 = (double)cv.w * 256.0;
  clip.y1 = (double)cv.h * 256.0;
  clip.ix0 = (int32_t)clip.x0;
  clip.ix1 = (int32_t)clip.x1;
  return clip;
}

// Work items of source segment `sg`; written to `out` when it is not null.
// `bad` (optional) receives kFailInvalidGeometry for a curve that cannot
// be flattened.
__device__ __forceinline__ uint32_t source_items(const FrameDev& f, const SrcSeg& sg, EdgeItem* out, uint32_t* bad = nullptr) {
  const CmdXform& m = f.xforms[sg.cmd];
  if (sg.kind == kSrcLine) {
    if (out) {
      edge::P* op = (edge::P*)out->p;
      op[0] = xform_pt(m, f.vtx, sg.i0);
      op[1] = xform_pt(m, f.vtx, sg.i1);
      out->cmd = sg.cmd;
      out->kind = kItemLine;
      out->any = 0u;
    }
    return 1u;
  }
  const int npts = sg.kind == kSrcQuad ? 3 : 4;
  edge::P cv[4];
  cv[0] = xform_pt(m, f.vtx, sg.i0);
  cv[1] = xform_pt(m, f.vtx, sg.i1);
  cv[2] = xform_pt(m, f.vtx, sg.i1 + 1);
  cv[3] = npts == 4 ? xform_pt(m, f.vtx, sg.i1 + 2) : cv[2];
  // A curve with a NaN or infinite control point has no defined flattening
  // (the reference subdivides it to its recursion limit or trips its own
  // assertions): the fill fails with BL_ERROR_INVALID_GEOMETRY.
  bool finite = true;
  for (int i = 0; i < npts; i++) finite = finite && isfinite(cv[i].x) && isfinite(cv[i].y);
  if (!finite) {
    if (bad) *bad = kFailInvalidGeometry;
    return 0u;
  }
  const edge::Clip clip = canvas_clip(f.canvases[f.cmds[sg.cmd].canvas]);
  uint32_t common = 0xFu, any = 0u;
  for (int i = 0; i < npts; i++) { const uint32_t fl = edge::fxy(clip, cv[i]); common &= fl; any |= fl; }
  if (common) {
    if (common & (edge::CY0 | edge::CY1)) return 0u;   // above or below the canvas: nothing
    // wholly left or right of the canvas: the border line over its y range
    // (edgebuilder_p.h:1618-1890, Builder::curve_split)
    if (out) {
      edge::P* op = (edge::P*)out->p;
      op[0] = edge::P{0.0, edge::dclamp(cv[0].y, clip.y0, clip.y1)};
      op[1] = edge::P{0.0, edge::dclamp(cv[npts - 1].y, clip.y0, clip.y1)};
      out->cmd = sg.cmd;
      out->kind = (common & edge::CX0) ? kItemBorderLeft : kItemBorderRight;
      out->any = any;
    }
    return 1u;
  }
  CountEmit dummy;
  edge::Builder<CountEmit> b(dummy, clip, f.tol_sq);
  edge::P spline[8 * 3 + 1];
  const int pieces = npts == 3 ? b.split_quad(cv, spline) : b.split_cubic(cv, spline);
  if (out) {
    for (int k = 0; k < pieces; k++) {
      const edge::P* sp = spline + k * (npts - 1);
      edge::P* op = (edge::P*)out[k].p;
      for (int i = 0; i < npts; i++) op[i] = sp[i];
      out[k].cmd = sg.cmd;
      out[k].kind = npts == 3 ? kItemQuad : kItemCubic;
      out[k].any = any;
    }
  }
  return (uint32_t)pieces;
}

// A failed work item fails its whole command (rare; the plain stores race
// benignly, every writer sets the same bits).
__device__ __forceinline__ void fail_cmd(const FrameScratch& s, uint32_t cmd, uint32_t bits) {
  atomicOr(&s.cmd_err[cmd], bits);
  *s.err_any = 1u;
}

__global__ void k_items_count(FrameDev f, FrameScratch s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n_src) return;
  uint32_t bad = 0;
  const SrcSeg sg = f.src[i];
  s.item_cnt[i] = source_items(f, sg, nullptr, &bad);
  if (bad) fail_cmd(s, sg.cmd, bad);
}

__global__ void k_items_write(FrameDev f, FrameScratch s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n_src) return;
  (void)source_items(f, f.src[i], s.items + s.item_off[i]);
}

// Flattens one work item into `e`; returns its kFailInvalidGeometry /
// kTooManySegments bits.
template<typename Emit>
__device__ __forceinline__ uint32_t build_item(const FrameDev& f, const EdgeItem& it, Emit& e) {
  edge::Builder<Emit> b(e, canvas_clip(f.canvases[f.cmds[it.cmd].canvas]), f.tol_sq);
  edge::P stack[edge::Builder<Emit>::kLimit * 4];
  const edge::P* ip = (const edge::P*)it.p;
  switch (it.kind) {
    case kItemLine: b.line(ip[0], ip[1]); break;
    case kItemQuad: b.curve_piece(ip, 3, it.any, stack); break;
    case kItemCubic: b.curve_piece(ip, 4, it.any, stack); break;
    case kItemBorderLeft: b.acc_left(ip[0].y, ip[1].y); break;
    default: b.acc_right(ip[0].y, ip[1].y); break;
  }
  return b.invalid;
}

// Per-command bounds of a wave's work items with one set of atomics per run
// of equal command indices: each lane folds in the lanes d = 1, 2, 4, ...,
// 32 above it that carry its command (min / max are idempotent, so the
// overlapping windows are harmless), and the first lane of each run issues
// the atomics.  Every lane of the wave must call it; lanes without segments
// pass the neutral bounds, lanes past the item total pass cmd 0xFFFFFFFF.
// Items of one command are contiguous, which keeps the runs long.
__device__ __forceinline__ void wave_cmd_bounds(CmdBounds* __restrict__ bounds, uint32_t cmd, int32_t mnx, int32_t mxx,
                                                int32_t mny, int32_t mxy) {
  const int lane = (int)(threadIdx.x & 63u);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t oc = __shfl_down(cmd, d, 64);
    const int32_t a = __shfl_down(mnx, d, 64), b = __shfl_down(mxx, d, 64);
    const int32_t c = __shfl_down(mny, d, 64), e = __shfl_down(mxy, d, 64);
    if (lane + d < 64 && oc == cmd) {
      mnx = a < mnx ? a : mnx;
      mxx = b > mxx ? b : mxx;
      mny = c < mny ? c : mny;
      mxy = e > mxy ? e : mxy;
    }
  }
  const uint32_t prev = __shfl_up(cmd, 1, 64);
  if ((lane == 0 || prev != cmd) && cmd != 0xFFFFFFFFu && mnx <= mxx) {
    CmdBounds* bd = &bounds[cmd];
    atomicMin(&bd->min_x, mnx);
    atomicMax(&bd->max_x, mxx);
    atomicMin(&bd->min_y, mny);
    atomicMax(&bd->max_y, mxy);
  }
}

// Counting pass over the work items (slots past the item total get 0, so
// the scan may run over the host's upper bound `cap`).
__global__ __launch_bounds__(64) void k_edges_count(FrameDev f, FrameScratch s, uint32_t cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool slot = i < cap;
  const bool active = slot && i < s.item_off[f.n_src];
  CountEmit e;
  uint32_t cmd = 0xFFFFFFFFu;
  if (active) {
    const EdgeItem it = s.items[i];
    cmd = it.cmd;
    const uint32_t bad = build_item(f, it, e);
    if (bad) fail_cmd(s, cmd, bad);
  }
  if (slot) s.src_count[i] = e.n;
  wave_cmd_bounds(s.bounds, cmd, e.min_x, e.max_x, e.min_y, e.max_y);
}

__global__ __launch_bounds__(64) void k_edges_emit(FrameDev f, FrameScratch s, uint32_t cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap || i >= s.item_off[f.n_src]) return;
  const EdgeItem it = s.items[i];
  WriteEmit e{s.segs, s.seg_cmd, s.src_off[i], it.cmd};
  build_item(f, it, e);
}

// Pool mode: segments staged per thread in LDS before the block's slot
// range is known (work items average little more than one segment).
constexpr uint32_t kPoolStage = 4;
constexpr uint32_t kPoolThreads = 256;
constexpr uint32_t kPoolBlocks = 4096;

struct StageEmit {
  DevSeg* st;   // this thread's kPoolStage slots
  uint32_t n = 0;
  int32_t min_x = INT32_MAX, max_x = INT32_MIN, min_y = INT32_MAX, max_y = INT32_MIN;
  __device__ __forceinline__ void push(int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t sign) {
    if (n < kPoolStage) st[n] = DevSeg{x0, y0, (int32_t)((uint32_t)x1 | (sign << 31)), y1};
    n++;
    int32_t lo = (x0 < x1 ? x0 : x1) >> 8, hi = (x0 < x1 ? x1 : x0) >> 8;
    min_x = lo < min_x ? lo : min_x;
    max_x = hi > max_x ? hi : max_x;
    min_y = y0 < min_y ? y0 : min_y;
    max_y = y1 > max_y ? y1 : max_y;
  }
};

// Single pass into a pool sized ahead of time.  Each work item is flattened
// once: its first kPoolStage segments go to the thread's slots in LDS while
// all of them are counted (with their bounds); the block then takes one slot
// range from the cursor -- one atomic per block and round, since the cursor
// is one address shared by all XCDs -- and copies the staged segments out
// with coalesced stores, each slot's owner found among the threads' prefix
// sums.  An item with more segments than it staged is flattened again
// straight into its slots.  When the block's range does not fit below the
// capacity the overflow flag is raised, the items that still fit are
// written and the slots of the others below the capacity are marked unused
// (binning is skipped after an overflow and the frame is redone with exact
// counting).  Blocks stride over the work items the device counted (at most
// `item_cap`, the size of the item buffer).
__global__ __launch_bounds__(kPoolThreads) void k_edges_emit_pool(FrameDev f, FrameScratch s, uint32_t cap,
                                                                  uint32_t item_cap) {
  __shared__ DevSeg stage[kPoolThreads * kPoolStage];
  __shared__ uint32_t pre[kPoolThreads];
  __shared__ uint32_t cmd_of[kPoolThreads];
  __shared__ uint32_t wave_sum[kPoolThreads / 64];
  __shared__ uint32_t block_base;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t n_items = min(s.item_off[f.n_src], item_cap);
  for (uint32_t i0 = blockIdx.x * kPoolThreads; i0 < n_items; i0 += gridDim.x * kPoolThreads) {
    const uint32_t i = i0 + tid;
    const bool active = i < n_items;
    EdgeItem it;
    StageEmit e{stage + tid * kPoolStage};
    if (active) {
      it = s.items[i];
      const uint32_t bad = build_item(f, it, e);
      if (bad) fail_cmd(s, it.cmd, bad);
    }
    const uint32_t need = e.n;
    uint32_t incl = need;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if ((int)lane >= d) incl += t;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, total = 0;
    for (uint32_t w = 0; w < kPoolThreads / 64; w++) {
      wave_off += w < wave ? wave_sum[w] : 0u;
      total += wave_sum[w];
    }
    if (tid == 0) block_base = total ? atomicAdd(s.seg_cursor, total) : 0u;
    const uint32_t excl = wave_off + incl - need;
    pre[tid] = excl;
    cmd_of[tid] = active ? it.cmd : 0u;
    __syncthreads();
    const uint32_t base = block_base;
    if (base <= cap && cap - base >= total) {
      for (uint32_t j = tid; j < total; j += kPoolThreads) {
        // owner: the last thread whose range starts at or before j
        uint32_t lo = 0, hi = kPoolThreads - 1u;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1u) >> 1;
          if (pre[mid] <= j) lo = mid;
          else hi = mid - 1u;
        }
        const uint32_t k = j - pre[lo];
        if (k < kPoolStage) {
          s.segs[base + j] = stage[lo * kPoolStage + k];
          s.seg_cmd[base + j] = cmd_of[lo];
        }
      }
      if (need > kPoolStage) {
        WriteEmit w{s.segs, s.seg_cmd, base + excl, it.cmd};
        build_item(f, it, w);
      }
    }
    else if (need) {
      const uint32_t start = base + excl;
      if (start <= cap && cap - start >= need) {
        if (need <= kPoolStage) {
          for (uint32_t k = 0; k < need; k++) {
            s.segs[start + k] = stage[tid * kPoolStage + k];
            s.seg_cmd[start + k] = it.cmd;
          }
        }
        else {
          WriteEmit w{s.segs, s.seg_cmd, start, it.cmd};
          build_item(f, it, w);
        }
      }
      else {
        for (uint32_t k = start; k < cap; k++) {
          s.segs[k] = DevSeg{0, 0, 0, 0};
          s.seg_cmd[k] = 0xFFFFFFFFu;
        }
        atomicOr(s.flags, 1u);
      }
    }
    wave_cmd_bounds(s.bounds, active ? it.cmd : 0xFFFFFFFFu, e.min_x, e.max_x, e.min_y, e.max_y);
    __syncthreads();   // the next round reuses the staging slots and block_base
  }
}

void launch_edges_pool(const FrameDev& f, const FrameScratch& s, uint32_t cap, uint32_t item_cap, hipStream_t st) {
  if (!f.n_src || !item_cap) return;
  const uint32_t blocks = std::min<uint32_t>((item_cap + kPoolThreads - 1) / kPoolThreads, kPoolBlocks);
  hipLaunchKernelGGL(k_edges_emit_pool, dim3(blocks), dim3(kPoolThreads), 0, st, f, s, cap, item_cap);
}

void launch_edge_items(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (!f.n_src) return;
  hipLaunchKernelGGL(k_items_count, dim3((f.n_src + 255) / 256), dim3(256), 0, st, f, s);
  scan_exclusive(s.item_cnt, s.item_off, f.n_src, s.scan_tmp, st);
  hipLaunchKernelGGL(k_items_write, dim3((f.n_src + 255) / 256), dim3(256), 0, st, f, s);
}

void launch_edges(const FrameDev& f, const FrameScratch& s, bool count, uint32_t cap, hipStream_t st) {
  if (!f.n_src || !cap) return;
  dim3 grid((cap + 63) / 64), block(64);
  if (count) hipLaunchKernelGGL(k_edges_count, grid, block, 0, st, f, s, cap);
  else hipLaunchKernelGGL(k_edges_emit, grid, block, 0, st, f, s, cap);
}

// ---------------------------------------------------------------------------
// Stage 2: binning
// ---------------------------------------------------------------------------

__device__ __forceinline__ void seg_bands(const DevSeg& g, int bands, int& b0, int& b1) {
  if (g.y1 <= g.y0) { b0 = 1; b1 = 0; return; }
  int ey0 = g.y0 >> 8;
  int ey1 = (g.y1 - 1) >> 8;
  b0 = ey0 / kTileH;
  b1 = ey1 / kTileH;
  b0 = b0 < 0 ? 0 : b0;
  b1 = b1 > bands - 1 ? bands - 1 : b1;
}

// Number of valid segments: `n`, or in pool mode min(*seg_cursor, n).
__device__ __forceinline__ uint32_t seg_limit(const FrameScratch& s, uint32_t n) {
  if (!s.seg_cursor) return n;
  uint32_t c = *s.seg_cursor;
  return c < n ? c : n;
}

// Pool mode: a segment-pool overflow (flag bit 0) leaves the frame to be
// redone with exact counting, so binning is skipped.
__device__ __forceinline__ bool pool_overflowed(const FrameScratch& s) { return s.flags && (*s.flags & 1u); }

// Grid of the binning kernels: one thread per segment slot, at most
// kBinBlocks workgroups striding over the slots (in pool mode `n` is the pool
// capacity, an upper bound of the slots used).
constexpr uint32_t kBinBlocks = 8192;
inline dim3 bin_grid(uint32_t n) { return dim3(std::min<uint32_t>((n + 255) / 256, kBinBlocks)); }

// Runs of equal keys among the lanes of a wave (key ~0 = no key).  Each lane
// gets the lane index of its run's first lane and the run's length there.
struct WaveRun {
  uint32_t head;   // lane index of the first lane of this lane's run
  uint32_t len;    // run length (meaningful in the first lane)
  bool first;
};
__device__ __forceinline__ WaveRun wave_run(uint64_t key) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t plo = __shfl_up((uint32_t)key, 1, 64), phi = __shfl_up((uint32_t)(key >> 32), 1, 64);
  const bool first = lane == 0 || (((uint64_t)phi << 32) | plo) != key;
  const unsigned long long heads = __ballot(first);
  WaveRun r;
  r.first = first;
  r.head = 63u - (uint32_t)__builtin_clzll(heads & ((2ull << lane) - 1ull));
  const unsigned long long above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
  r.len = (above ? (uint32_t)__builtin_ctzll(above) : 64u) - lane;
  return r;
}

// Segments that stay within one band (the common case) are counted per run
// of equal (band, command) cells among the lanes of a wave, one atomic per
// run; segments crossing bands add to each of their cells.
__global__ void k_bin_count(FrameDev f, FrameScratch s, uint32_t n) {
  const uint32_t lim = seg_limit(s, n);
  if (pool_overflowed(s)) return;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < lim; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    uint64_t key = ~0ull;
    if (i < lim) {
      const DevSeg g = s.segs[i];
      const uint32_t cmd = s.seg_cmd[i];
      if (cmd != 0xFFFFFFFFu && !s.cmd_err[cmd]) {
        const DevCanvas& cv = f.canvases[f.cmds[cmd].canvas];
        int b0, b1;
        seg_bands(g, (int)cv.bands, b0, b1);
        const uint64_t row = cv.mat_base + (uint64_t)(cmd - cv.cmd_base);
        if (b0 == b1) key = row + (uint64_t)b0 * cv.n_cmds;
        else
          for (int b = b0; b <= b1; b++) atomicAdd(&s.mat_cnt[row + (uint64_t)b * cv.n_cmds], 1u);
      }
    }
    const WaveRun r = wave_run(key);
    if (r.first && key != ~0ull) atomicAdd(&s.mat_cnt[key], r.len);
  }
}

void launch_bin_count(const FrameDev& f, const FrameScratch& s, uint32_t n_segs, hipStream_t st) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_bin_count, bin_grid(n_segs), dim3(256), 0, st, f, s, n_segs);
}

__global__ void k_presence(FrameScratch s, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s.mat_pres[i] = s.mat_cnt[i] ? 1u : 0u;
}

// Box fills are present in every band their box touches; a box fill with a
// non-solid style also marks its bands (they take the general compositor).
// 16 threads per command, each taking every 16th band of the box.
__global__ void k_box_presence(FrameDev f, FrameScratch s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = i >> 4;
  if (c >= f.n_cmds) return;
  const DevCmd& cmd = f.cmds[c];
  if (cmd.kind == kCmdAnalytic) return;
  const DevCanvas& cv = f.canvases[cmd.canvas];
  const int b0 = cmd.box[1] / kTileH, b1 = (cmd.box[3] - 1) / kTileH;
  const uint64_t row = cv.mat_base + (uint64_t)(c - cv.cmd_base);
  const bool nonsolid = f.styles[cmd.style].kind != kFetchSolid;
  // k_render_box_fast handles aligned boxes of a solid colour, SRC_OVER or
  // SRC_COPY, on 32-bit canvases
  const bool slow = nonsolid || cmd.kind != kCmdBoxA || (cmd.op != kOpSrcOver && cmd.op != kOpSrcCopy) ||
                    cv.format == kFmtA8;
  for (int b = b0 + (int)(i & 15u); b <= b1; b += 16) {
    s.mat_pres[row + (uint64_t)b * cv.n_cmds] = 1u;
    if (nonsolid) s.band_nonsolid[cv.band_base + b] = 1u;
    if (slow) s.band_slow[cv.band_base + b] = 1u;
  }
}

void launch_presence(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (f.n_mat) hipLaunchKernelGGL(k_presence, dim3((unsigned)((f.n_mat + 255) / 256)), dim3(256), 0, st, s, f.n_mat);
  if (f.n_bands) launch_fill_u32(s.band_nonsolid, 0u, f.n_bands, st);
  if (f.n_bands) launch_fill_u32(s.band_slow, 0u, f.n_bands, st);
  if (f.n_cmds) hipLaunchKernelGGL(k_box_presence, dim3((unsigned)(((uint64_t)f.n_cmds * 16 + 255) / 256)), dim3(256), 0, st, f, s);
}

// As k_bin_count: a run of lanes with the same single-band cell takes its
// slots with one cursor atomic; lane k of the run writes slot base + k.
__global__ void k_bin_scatter(FrameDev f, FrameScratch s, uint32_t n, uint32_t* cursor) {
  const uint32_t lim = seg_limit(s, n);
  if (pool_overflowed(s)) return;
  // pool mode: the binned pool is sized ahead of time; a larger total is an
  // overflow (the frame falls back to exact counting)
  if (s.binned_cap && s.mat_off[f.n_mat] > s.binned_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(s.flags, 2u);
    return;
  }
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < lim; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    uint64_t key = ~0ull;
    DevSeg g{0, 0, 0, 0};
    if (i < lim) {
      g = s.segs[i];
      const uint32_t cmd = s.seg_cmd[i];
      if (cmd != 0xFFFFFFFFu && !s.cmd_err[cmd]) {
        const DevCanvas& cv = f.canvases[f.cmds[cmd].canvas];
        int b0, b1;
        seg_bands(g, (int)cv.bands, b0, b1);
        const uint64_t row = cv.mat_base + (uint64_t)(cmd - cv.cmd_base);
        if (b0 == b1) key = row + (uint64_t)b0 * cv.n_cmds;
        else
          for (int b = b0; b <= b1; b++) {
            const uint32_t pos = atomicAdd(&cursor[row + (uint64_t)b * cv.n_cmds], 1u);
            s.binned[pos] = g;
          }
      }
    }
    const WaveRun r = wave_run(key);
    uint32_t base = 0;
    if (r.first && key != ~0ull) base = atomicAdd(&cursor[key], r.len);
    base = __shfl(base, (int)r.head, 64);
    if (key != ~0ull) s.binned[base + (lane - r.head)] = g;
  }
}

void launch_bin_scatter(const FrameDev& f, const FrameScratch& s, uint32_t n_segs, uint32_t* cursor, hipStream_t st) {
  if (!n_segs) return;
  hipLaunchKernelGGL(k_bin_scatter, bin_grid(n_segs), dim3(256), 0, st, f, s, n_segs, cursor);
}

// Command lists of the bands: every present (band, command) cell becomes one
// entry, in command order within its band, carrying what the render filter
// and the record loads need (so they are one dependent load away).
__global__ void k_list_build(FrameDev f, FrameScratch s, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !s.mat_pres[i]) return;
  // canvas of matrix cell i: the last one whose matrix base is <= i
  uint32_t lo = 0, hi = f.n_canvases;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (f.canvases[mid].mat_base <= i) lo = mid; else hi = mid;
  }
  const DevCanvas& cv = f.canvases[lo];
  const uint32_t cmd = cv.cmd_base + (uint32_t)((i - cv.mat_base) % cv.n_cmds);
  const uint32_t li = s.list_off[i];
  s.list_ent[li] = make_uint4(cmd, s.cmd_tiles[cmd], s.mat_off[i], s.mat_cnt[i]);
  s.list_style[li] = f.cmds[cmd].style;
}

void launch_list_build(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (f.n_mat) hipLaunchKernelGGL(k_list_build, dim3((unsigned)((f.n_mat + 255) / 256)), dim3(256), 0, st, f, s, f.n_mat);
}

// Per command, the range of tile columns it can touch, packed as
// first | last << 16 (first > last when it touches none).  Analytic fills
// touch cells min_x .. max_x + 1 of their edges; boxes their pixel box.
__global__ void k_cmd_tiles(FrameDev f, FrameScratch s) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.n_cmds) return;
  const DevCmd& cmd = f.cmds[c];
  const DevCanvas& cv = f.canvases[cmd.canvas];
  int x0, x1;
  if (cmd.kind == kCmdAnalytic) {
    const CmdBounds bd = s.bounds[c];
    x0 = bd.min_x;
    x1 = bd.max_x + 1;
  }
  else {
    x0 = cmd.box[0];
    x1 = cmd.box[2] - 1;
  }
  const int last = (int)cv.tiles_x - 1;
  int t0 = x0 < 0 ? 0 : x0 / kTileW;
  int t1 = x1 < 0 ? -1 : x1 / kTileW;
  t1 = t1 > last ? last : t1;
  s.cmd_tiles[c] = (x0 > x1 || t0 > t1) ? 0xFFFFu : ((uint32_t)t0 | ((uint32_t)t1 << 16));
}

void launch_cmd_tiles(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (f.n_cmds) hipLaunchKernelGGL(k_cmd_tiles, dim3((f.n_cmds + 255) / 256), dim3(256), 0, st, f, s);
}

// ---------------------------------------------------------------------------
// Stage 3: render
// ---------------------------------------------------------------------------

// LDS cell buffer of one command in a tile: 16 rows x 64 cells, then one
// backdrop word per row (the summed contributions of cells left of the tile).
// Within a row the 16 groups of four cells are stored XOR-swizzled by the
// row index, so the same cell column of different rows (the rows of one
// steep edge, rasterized by neighbouring lanes) falls into different LDS
// banks; a group stays four consecutive words, so the compositor still
// reads a thread's four cells with one 128-bit load.
constexpr int kCellWords = kTileH * kTileW + kTileH;
constexpr int kBackdrop = kTileH * kTileW;

__device__ __forceinline__ int cell_index(int row, int x) {
  return row * kTileW + ((((x >> 2) ^ row) & 15) << 2) + (x & 3);
}

__device__ __forceinline__ uint32_t calc_mask(uint32_t cov, uint32_t rule_mask, uint32_t alpha) {
  uint32_t m = ((uint32_t)((int32_t)cov >> 9) & rule_mask) - 256u;
  int32_t mi = (int32_t)m;
  uint32_t a = (uint32_t)(mi < 0 ? -mi : mi);
  a = a > 256u ? 256u : a;
  return (a * alpha) >> 8;
}

__device__ __forceinline__ uint32_t row_excl_scan16(uint32_t v) {
  const int q = threadIdx.x & 15;
  uint32_t incl = v;
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    uint32_t t = __shfl_up(incl, d, 16);
    if (q >= d) incl += t;
  }
  return incl - v;
}

// Per-command record staged in LDS for the render loop.
struct RenderRecData {
  DevCmd cmd;
  DevStyle style;
  int32_t min_y;       // fill-row origin of analytic fills (24.8)
  int32_t pad;
};
union RenderRec {
  RenderRecData r;
  uint32_t w[sizeof(RenderRecData) / 4];
};
constexpr int kRecWords = (int)(sizeof(RenderRecData) / 4);
static_assert(sizeof(DevCmd) % 8 == 0, "DevStyle must stay 8-byte aligned in the record");

// Commands rasterized together.  The segments of a batch's commands are
// spread over all lanes of the workgroup, each command into its own LDS cell
// buffer; after one barrier every thread composites its pixels through the
// batch's commands in submission order.  Rasterization of different commands
// is independent, so only the composite has to follow the command order.
#ifndef B2AMD_RENDER_BATCH
#define B2AMD_RENDER_BATCH 7
#endif
constexpr int kBatch = B2AMD_RENDER_BATCH;
static_assert(kBatch >= 1 && kBatch * kRecWords <= 2 * kTileThreads, "two record words per thread");

// Word `w` of the record of command `cmd_i` (style index `style_i`).
__device__ __forceinline__ uint32_t rec_word(const FrameDev& f, const FrameScratch& s, uint32_t cmd_i, uint32_t style_i, int w) {
  constexpr int kCmdW = (int)(sizeof(DevCmd) / 4), kStyleW = (int)(sizeof(DevStyle) / 4);
  if (w < kCmdW) return ((const uint32_t*)(f.cmds + cmd_i))[w];
  if (w < kCmdW + kStyleW) return ((const uint32_t*)(f.styles + style_i))[w - kCmdW];
  if (w == kCmdW + kStyleW) return (uint32_t)s.bounds[cmd_i].min_y;
  return 0u;
}

// Word `gw` of the records of the batch that starts at selected command `kb`.
__device__ __forceinline__ uint32_t batch_rec_word(const FrameDev& f, const FrameScratch& s, const uint32_t* sel_cmd,
                                                   const uint32_t* sel_style, uint32_t kb, int gw) {
  const int b = gw / kRecWords, w = gw - b * kRecWords;
  return rec_word(f, s, sel_cmd[kb + b], sel_style[kb + b], w);
}

// Inclusive prefix of the segment counts of the batch's `nb` commands.
__device__ __forceinline__ void batch_prefix(const uint32_t* sel_cnt, uint32_t kb, uint32_t nb, uint32_t end[kBatch]) {
  uint32_t acc = 0;
#pragma unroll
  for (int b = 0; b < kBatch; b++) {
    acc += (uint32_t)b < nb ? sel_cnt[kb + b] : 0u;
    end[b] = acc;
  }
}

// Batch slot and in-command index of the batch's segment `j` (j < end[kBatch - 1]).
__device__ __forceinline__ uint32_t batch_find(const uint32_t end[kBatch], uint32_t j, uint32_t* local) {
  uint32_t b = 0, base = 0;
#pragma unroll
  for (int i = 0; i < kBatch - 1; i++) {
    if (j >= end[i]) { b = (uint32_t)i + 1u; base = end[i]; }
  }
  *local = j - base;
  return b;
}

// Tile rows one raster task covers (a task pays the rasterizer's set-up and
// advance once for its rows).
#ifndef B2AMD_ROWS_PER_TASK
#define B2AMD_ROWS_PER_TASK 1
#endif
constexpr int kRowsPerTask = B2AMD_ROWS_PER_TASK;

// First pass over one binned segment of an analytic fill: a segment right of
// the tile is dropped; one left of it adds its signed vertical extent per
// covered row to the backdrop; for one that reaches the tile the number of
// raster tasks for the tile rows it covers is returned (raster_rows).
__device__ __forceinline__ uint32_t seg_tile_rows(const DevSeg& g, uint32_t* cells, int X0, int Y0) {
  const int32_t x1 = g.x1 & 0x7FFFFFFF;
  const int32_t cmin = (g.x0 < x1 ? g.x0 : x1) >> 8;
  const int32_t cmax = (g.x0 < x1 ? x1 : g.x0) >> 8;
  int r0 = g.y0 >> 8, r1 = (g.y1 - 1) >> 8;
  if (r0 < Y0) r0 = Y0;
  if (r1 > Y0 + kTileH - 1) r1 = Y0 + kTileH - 1;
  if (cmin >= X0 + kTileW || r1 < r0) return 0u;
  if (cmax + 1 < X0) {
    // every cell of every row lies left of the tile: each row adds its y
    // extent, signed by the edge direction (bit 31 of x1)
    const int32_t sgn = g.x1 < 0 ? -1 : 1;
    for (int r = r0; r <= r1; r++) {
      const int lo_y = g.y0 > r * 256 ? g.y0 : r * 256;
      const int hi_y = g.y1 < (r + 1) * 256 ? g.y1 : (r + 1) * 256;
      atomicAdd(&cells[kBackdrop + r - Y0], (uint32_t)(sgn * (hi_y - lo_y)) << 9);
    }
    return 0u;
  }
  return (uint32_t)((r1 - r0 + kRowsPerTask) / kRowsPerTask);
}

// Rows `row` .. `row` + kRowsPerTask - 1 (within the tile) of a segment that
// reaches the tile: the closed-form row emitter adds the cells inside the
// tile and one backdrop value per row for the cells left of it.
__device__ __forceinline__ void raster_rows(const DevSeg& g, uint32_t* cells, int X0, int Y0, int row) {
  ras::Edge e;
  if (!ras::edge_init(e, g.x0, g.y0, g.x1 & 0x7FFFFFFF, g.y1)) return;
  const int32_t sgn = g.x1 < 0 ? -1 : 1;
  const int last = row + kRowsPerTask - 1 < e.row1 ? row + kRowsPerTask - 1 : e.row1;
  for (int r = row; r <= last && r < Y0 + kTileH; r++) {
    const int lr = r - Y0;
    const uint32_t bd = ras::emit_row(e, r, X0, kTileW, sgn, [&](int32_t k, uint32_t v) {
      if (v) atomicAdd(&cells[cell_index(lr, k)], v);
    });
    if (bd) atomicAdd(&cells[kBackdrop + lr], bd);
  }
}

// Destination pixels px .. px + 3 of a row (zero where outside the canvas).
__device__ __forceinline__ void load4(const uint8_t* rowp, int px, int w, bool a8, uint32_t d[4]) {
  const bool full4 = px + 3 < w;
  if (a8) {
    if (full4 && (((uintptr_t)(rowp + px)) & 3) == 0) {
      const uint32_t v = *(const uint32_t*)(rowp + px);
#pragma unroll
      for (int k = 0; k < 4; k++) d[k] = (v >> (8 * k)) & 0xFFu;
    }
    else {
#pragma unroll
      for (int k = 0; k < 4; k++) d[k] = px + k < w ? rowp[px + k] : 0u;
    }
  }
  else {
    const uint32_t* r32 = (const uint32_t*)rowp;
    if (full4 && (((uintptr_t)(r32 + px)) & 15) == 0) {
      const uint4 v = *(const uint4*)(r32 + px);
      d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
    else {
#pragma unroll
      for (int k = 0; k < 4; k++) d[k] = px + k < w ? r32[px + k] : 0u;
    }
  }
}

__device__ __forceinline__ void store4(uint8_t* rowp, int px, int w, bool a8, const uint32_t d[4]) {
  const bool full4 = px + 3 < w;
  if (a8) {
    if (full4 && (((uintptr_t)(rowp + px)) & 3) == 0) {
      *(uint32_t*)(rowp + px) = d[0] | (d[1] << 8) | (d[2] << 16) | (d[3] << 24);
    }
    else {
#pragma unroll
      for (int k = 0; k < 4; k++) if (px + k < w) rowp[px + k] = (uint8_t)d[k];
    }
  }
  else {
    uint32_t* r32 = (uint32_t*)rowp;
    if (full4 && (((uintptr_t)(r32 + px)) & 15) == 0) {
      *(uint4*)(r32 + px) = make_uint4(d[0], d[1], d[2], d[3]);
    }
    else {
#pragma unroll
      for (int k = 0; k < 4; k++) if (px + k < w) r32[px + k] = d[k];
    }
  }
}

// Masks of pixels px .. px + 3 of row y for a box fill.
__device__ __forceinline__ void box_masks4(const FrameDev& f, const DevCmd& c, int y, int px, uint32_t m[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) m[k] = 0u;
  if (y < c.box[1] || y >= c.box[3]) return;
  if (c.kind == kCmdBoxA) {
#pragma unroll
    for (int k = 0; k < 4; k++) m[k] = (px + k >= c.box[0] && px + k < c.box[2]) ? c.alpha : 0u;
  }
  else if (c.kind == kCmdBoxM) {
    // FillMask with a VMaskA8 command (rendercommandprocsync_p.h:65-84,
    // compopgeneric_p.h:165-183): byte x - x0 of the mask row, times the
    // alpha below full alpha.  The view is exactly the box (host-checked), so
    // every byte read lies in row y - y0 of the view.
    const DevImage& mi = f.images[c.box_mask[0]];
    const uint8_t* row = (const uint8_t*)mi.ptr + (int64_t)(y - c.box[1]) * mi.stride;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = px + k;
      if (x >= c.box[0] && x < c.box[2]) {
        const uint32_t v = row[x - c.box[0]];
        m[k] = c.alpha >= 255u ? v : pipe::div255(v * c.alpha);
      }
    }
  }
  else {
    const int kind = (y == c.box_row_first) ? 0 : (y == c.box_row_last) ? 2 : 1;
    const uint32_t rm = c.box_mask[kind];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = px + k;
      if (x >= c.box[0] && x < c.box[2]) {
        const int col = (x == c.box[0]) ? 0 : (x == c.box[2] - 1) ? 2 : 1;
        m[k] = (rm >> (8 * col)) & 0xFFu;
      }
    }
  }
}

// Fetch and composite of pixels px .. px + 3 under masks m; returns how many
// were composited.
template<bool kSolid = false, bool kDither = false>
__device__ __forceinline__ uint32_t composite4(const FrameDev& f, const DevCmd& c, const DevStyle& st, const pipe::Row& rs,
                                               int px, int w, bool a8, const uint32_t m[4], uint32_t d[4]) {
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!m[k] || px + k >= w) continue;
    const uint32_t src = kSolid ? st.solid : pipe::fetch<kDither>(st, rs, px + k, f.luts, f.images, c.src_fmt);
    d[k] = a8 ? pipe::composite8(c.op, d[k], src >> 24, m[k]) : pipe::composite32(c.op, d[k], src, m[k]);
    n++;
  }
  return n;
}

// The same with all four fetches issued before the first is used, so their
// gradient-LUT / texel loads are in flight together (more registers).
template<bool kDither>
__device__ __forceinline__ uint32_t composite4_overlapped(const FrameDev& f, const DevCmd& c, const DevStyle& st,
                                                          const pipe::Row& rs, int px, int w, bool a8,
                                                          const uint32_t m[4], uint32_t d[4]) {
  constexpr bool kSolid = false;
  bool on[4];
  uint32_t src[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    on[k] = m[k] && px + k < w;
    src[k] = 0u;
    if (on[k]) src[k] = kSolid ? st.solid : pipe::fetch<kDither>(st, rs, px + k, f.luts, f.images, c.src_fmt);
  }
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!on[k]) continue;
    d[k] = a8 ? pipe::composite8(c.op, d[k], src[k] >> 24, m[k]) : pipe::composite32(c.op, d[k], src[k], m[k]);
    n++;
  }
  return n;
}

// Workgroup -> tile.  Workgroups are dealt to the 8 XCDs round-robin; the
// mapping gives each XCD a contiguous range of tiles, so the tiles of a band
// (which read the same command lists, records and segments) share one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t wg, uint32_t n_tiles) {
  const uint32_t per = (n_tiles + kXcds - 1) / kXcds;
  return (wg % kXcds) * per + wg / kXcds;
}

// One analytic tile (header `slot`): the whole workgroup rasterizes the
// band's commands that reach the tile and composites them in order.
template<bool kDither>
__device__ __forceinline__ void render_tile(const FrameDev& f, const FrameScratch& s, uint32_t slot) {
  __shared__ uint32_t cells[kBatch][kCellWords];
  const int t = threadIdx.x;
  const uint4 th = s.tile_hdr[slot];
  const DevCanvas cv = f.canvases[th.x];
  const uint32_t lb = th.y, le = th.z;
  const uint32_t tcol = th.w & 0xFFFFu;
  const uint32_t band = th.w >> 16;
  const uint32_t tile = cv.tile_base + band * cv.tiles_x + tcol;
  // A pool overflow leaves the frame incomplete: nothing is composited and
  // the host redoes the frame with exact counting.
  if (s.flags && *s.flags) {
    if (t == 0) s.tile_px[tile] = 0;
    return;
  }
  const int X0 = (int)tcol * kTileW;
  const int Y0 = (int)band * kTileH;

  const int row = t >> 4;
  const int q = t & 15;
  const int y = Y0 + row;
  const int px = X0 + q * 4;
  const bool row_ok = y < (int)cv.h;
  const bool a8 = cv.format == kFmtA8;

  // zero LDS (each thread its own swizzled group of four cells)
  const int grp = cell_index(row, q * 4);
#pragma unroll
  for (int b = 0; b < kBatch; b++) *(uint4*)&cells[b][grp] = make_uint4(0, 0, 0, 0);
  for (int i = t; i < kBatch * kTileH; i += kTileThreads) cells[i / kTileH][kBackdrop + i % kTileH] = 0;

  // load the destination tile
  uint32_t d[4] = {0, 0, 0, 0};
  uint8_t* rowp = (uint8_t*)cv.ptr + (int64_t)y * cv.stride;
  if (row_ok) load4(rowp, px, (int)cv.w, a8, d);

  const int lane = t & 63, wave = t >> 6;
  // Batch segment index of this thread in the first 256 of a batch: the
  // batch's segments are dealt to the four waves in turn, so every SIMD of
  // the CU takes part in rasterization.
  const uint32_t jt = ((uint32_t)lane << 2) | (uint32_t)wave;
  bool touched = false;
  uint32_t npix = 0;

  // The band's command list is filtered 256 entries at a time down to the
  // commands whose tile-column range holds this tile (ballot compaction,
  // submission order kept); the filter also gathers, in parallel, each
  // selected command's (band, command) cell extent and style index.
  __shared__ uint32_t sel_cmd[kTileThreads], sel_off[kTileThreads], sel_cnt[kTileThreads], sel_style[kTileThreads];
  __shared__ uint32_t sel_wave[kTileThreads / 64];
  // Row tasks of the segments that reach the tile: the segment of each
  // thread, and its first task index (low 12 bits) and batch slot (bits 12+).
  __shared__ DevSeg task_seg[kTileThreads];
  __shared__ uint16_t task_start[kTileThreads];
  __shared__ uint32_t task_wave[kTileThreads / 64];
  // Records (command, style, fill-row origin) of the batch being applied
  // and of the next one (double buffer): the next batch's records and first
  // 256 segments are loaded while the current batch is rasterized and
  // composited, so the chain of dependent global loads is off the critical
  // path.
  __shared__ RenderRec rec[2][kBatch];

  for (uint32_t chunk = lb; chunk < le; chunk += kTileThreads) {
    const uint32_t li = chunk + (uint32_t)t;
    uint4 ent = make_uint4(0, 0, 0, 0);
    uint32_t style_i = 0;
    bool keep = false;
    if (li < le) {
      ent = s.list_ent[li];
      style_i = s.list_style[li];
      keep = tcol >= (ent.y & 0xFFFFu) && tcol <= (ent.y >> 16);
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) sel_wave[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = 0, n_sel = 0;
#pragma unroll
    for (int w = 0; w < kTileThreads / 64; w++) {
      const uint32_t cw = sel_wave[w];
      off += w < wave ? cw : 0u;
      n_sel += cw;
    }
    if (keep) {
      const uint32_t slot = off + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
      sel_cmd[slot] = ent.x;
      sel_off[slot] = ent.z;
      sel_cnt[slot] = ent.w;
      sel_style[slot] = style_i;
    }
    __syncthreads();
    if (n_sel == 0) continue;   // uniform; sel_wave is rewritten only after the next chunk's ballot

    // first batch of the chunk: records and first segments
    DevSeg pre_seg{0, 0, 0, 0};
    uint32_t pre_b = 0;
    {
      const uint32_t nb = n_sel < (uint32_t)kBatch ? n_sel : (uint32_t)kBatch;
      uint32_t end[kBatch];
      batch_prefix(sel_cnt, 0, nb, end);
      if (t < (int)nb * kRecWords) rec[0][t / kRecWords].w[t % kRecWords] = batch_rec_word(f, s, sel_cmd, sel_style, 0, t);
      if (t + kTileThreads < (int)nb * kRecWords) {
        const int gw = t + kTileThreads;
        rec[0][gw / kRecWords].w[gw % kRecWords] = batch_rec_word(f, s, sel_cmd, sel_style, 0, gw);
      }
      if (jt < end[kBatch - 1]) {
        uint32_t loc;
        pre_b = batch_find(end, jt, &loc);
        pre_seg = s.binned[sel_off[pre_b] + loc];
      }
    }
    __syncthreads();

    for (uint32_t kb = 0, bi = 0; kb < n_sel; kb += kBatch, bi++) {
      const int cur = (int)(bi & 1u);
      const uint32_t nb = n_sel - kb < (uint32_t)kBatch ? n_sel - kb : (uint32_t)kBatch;
      uint32_t end[kBatch];
      batch_prefix(sel_cnt, kb, nb, end);
      const uint32_t total = end[kBatch - 1];
      const DevSeg seg = pre_seg;
      const uint32_t seg_b = pre_b;

      // prefetch the next batch
      const uint32_t kn = kb + kBatch;
      const bool has_next = kn < n_sel;
      uint32_t pw0 = 0, pw1 = 0;
      int nwn = 0;
      if (has_next) {
        const uint32_t nbn = n_sel - kn < (uint32_t)kBatch ? n_sel - kn : (uint32_t)kBatch;
        nwn = (int)nbn * kRecWords;
        if (t < nwn) pw0 = batch_rec_word(f, s, sel_cmd, sel_style, kn, t);
        if (t + kTileThreads < nwn) pw1 = batch_rec_word(f, s, sel_cmd, sel_style, kn, t + kTileThreads);
        uint32_t endn[kBatch];
        batch_prefix(sel_cnt, kn, nbn, endn);
        pre_seg = DevSeg{0, 0, 0, 0};
        pre_b = 0;
        if (jt < endn[kBatch - 1]) {
          uint32_t loc;
          pre_b = batch_find(endn, jt, &loc);
          pre_seg = s.binned[sel_off[kn + pre_b] + loc];
        }
      }

      // Rasterize the batch.  Segments left of the tile add their row
      // covers to the backdrop; every segment that reaches the tile becomes
      // one task per tile row it covers, so a lane's serial cell walk is one
      // row long and the rows are spread over all lanes.
      for (uint32_t rb = 0; rb < total; rb += kTileThreads) {
        const uint32_t j = rb + jt;
        DevSeg g = seg;
        uint32_t gb = seg_b;
        if (rb) {
          g = DevSeg{0, 0, 0, 0};
          gb = 0;
          if (j < total) {
            uint32_t loc;
            gb = batch_find(end, j, &loc);
            g = s.binned[sel_off[kb + gb] + loc];
          }
        }
        const uint32_t rows = j < total ? seg_tile_rows(g, cells[gb], X0, Y0) : 0u;
        uint32_t incl = rows;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
          const uint32_t v = __shfl_up(incl, dd, 64);
          if (lane >= dd) incl += v;
        }
        if (lane == 63) task_wave[wave] = incl;
        if (rows) task_seg[t] = g;
        __syncthreads();
        uint32_t before = 0, n_tasks = 0;
#pragma unroll
        for (int w = 0; w < kTileThreads / 64; w++) {
          const uint32_t cw = task_wave[w];
          before += w < wave ? cw : 0u;
          n_tasks += cw;
        }
        task_start[t] = (uint16_t)((before + incl - rows) | (gb << 12));
        __syncthreads();
        for (uint32_t k = (uint32_t)t; k < n_tasks; k += kTileThreads) {
          // owner: the last thread whose first task index is <= k
          int o = 0;
#pragma unroll
          for (int step = kTileThreads / 2; step > 0; step >>= 1)
            if ((uint32_t)(task_start[o + step] & 0xFFFu) <= k) o += step;
          const uint32_t ts = task_start[o];
          const DevSeg og = task_seg[o];
          const int r0 = (og.y0 >> 8) > Y0 ? (og.y0 >> 8) : Y0;
          raster_rows(og, cells[ts >> 12], X0, Y0, r0 + kRowsPerTask * (int)(k - (ts & 0xFFFu)));
        }
        __syncthreads();   // cells complete (and task arrays free for the next round)
      }

      // composite the batch's commands in order
      for (uint32_t b = 0; b < nb; b++) {
        const DevCmd& c = rec[cur][b].r.cmd;
        uint32_t m[4] = {0, 0, 0, 0};

        if (c.kind == kCmdAnalytic) {
          uint32_t* cb = cells[b];
          uint4 cv4 = *(uint4*)&cb[grp];
          uint32_t bdv = cb[kBackdrop + row];
          *(uint4*)&cb[grp] = make_uint4(0, 0, 0, 0);
          uint32_t sum = cv4.x + cv4.y + cv4.z + cv4.w;
          uint32_t base = (256u << 9) + bdv + row_excl_scan16(sum);
          // every lane of the row has read the backdrop before the scan
          if (q == 0) cb[kBackdrop + row] = 0;
          uint32_t c0 = base + cv4.x, c1 = c0 + cv4.y, c2 = c1 + cv4.z, c3 = c2 + cv4.w;
          m[0] = calc_mask(c0, c.rule_mask, c.alpha);
          m[1] = calc_mask(c1, c.rule_mask, c.alpha);
          m[2] = calc_mask(c2, c.rule_mask, c.alpha);
          m[3] = calc_mask(c3, c.rule_mask, c.alpha);
        }
        else {
          box_masks4(f, c, y, px, m);
        }

        if (row_ok && (m[0] | m[1] | m[2] | m[3]) != 0) {
          touched = true;
          const DevStyle& st = rec[cur][b].r.style;
          const int fill_row0 = (c.kind == kCmdAnalytic) ? (rec[cur][b].r.min_y >> 8) : (int)c.fetch_row0;
          const pipe::Row rs = pipe::row_setup(st, y, fill_row0, c.kind == kCmdAnalytic ? cv.ref_band_h : 0u);
          npix += composite4_overlapped<kDither>(f, c, st, rs, px, (int)cv.w, a8, m, d);
        }
      }
      if (has_next) {
        if (t < nwn) rec[cur ^ 1][t / kRecWords].w[t % kRecWords] = pw0;
        if (t + kTileThreads < nwn) {
          const int gw = t + kTileThreads;
          rec[cur ^ 1][gw / kRecWords].w[gw % kRecWords] = pw1;
        }
      }
      __syncthreads();   // next records visible; cells cleared; these records no longer read
    }
  }

  // Composited-pixel statistics: one value per tile, summed by k_sum_u64.
  __shared__ uint32_t wave_px[kTileThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) npix += __shfl_xor(npix, o, 64);
  if (lane == 0) wave_px[wave] = npix;
  __syncthreads();
  if (t == 0) {
    unsigned long long tp = 0;
#pragma unroll
    for (int w = 0; w < kTileThreads / 64; w++) tp += wave_px[w];
    s.tile_px[tile] = tp;
  }

  if (!touched || !row_ok) return;
  store4(rowp, px, (int)cv.w, a8, d);
}

// Four workgroups per CU: LDS (7 cell buffers) and the register budget are
// sized for it.  One workgroup per analytic tile, XCD-contiguous.
template<bool kDither>
__global__ __launch_bounds__(kTileThreads, 4) void k_render(FrameDev f, FrameScratch s, uint32_t n) {
  const uint32_t slot = xcd_tile(blockIdx.x, n);
  // uniform: padding workgroups of the XCD mapping, and (when n is an upper
  // bound) slots past the analytic tiles
  if (slot >= n || slot >= s.tile_off[f.n_bands]) return;
  render_tile<kDither>(f, s, slot);
}

// The same tiles drained from a work queue by a persistent grid: every
// workgroup takes the next queue position with one atomic until the queue is
// exhausted (every workgroup reaches that exit).  The queue holds the
// analytic tile slots in the order k_tile_scatter wrote (longest predicted
// first), or is the identity (null: band order).  With `s.tile_cycles` set
// (2 x the tile count), the wall-clock ticks (100 MHz) of every tile and
// then its composited pixels are logged (cost-model samples).
template<bool kLog, int kOcc, bool kDither>
__global__ __launch_bounds__(kTileThreads, kOcc) void k_render_queue(FrameDev f, FrameScratch s) {
  __shared__ uint32_t next;
  const uint32_t total = s.tile_off[f.n_bands];
  for (;;) {
    __syncthreads();   // the previous tile's LDS reads are done and `next` is free
    if (threadIdx.x == 0) next = atomicAdd(s.queue_head, 1u);
    __syncthreads();
    const uint32_t i = next;
    if (i >= total) break;   // uniform
    const uint32_t slot = s.queue ? s.queue[i] : i;
    if constexpr (!kLog) {
      render_tile<kDither>(f, s, slot);
    }
    else {
      const long long c0 = wall_clock64();
      render_tile<kDither>(f, s, slot);
      __syncthreads();
      if (threadIdx.x == 0) {
        s.tile_cycles[slot] = (uint32_t)min((long long)0xFFFFFFFF, wall_clock64() - c0);
        // and the tile's composited pixels (written by this thread)
        const uint4 th = s.tile_hdr[slot];
        const DevCanvas& cv = f.canvases[th.x];
        const unsigned long long px = s.tile_px[cv.tile_base + (th.w >> 16) * cv.tiles_x + (th.w & 0xFFFFu)];
        s.tile_cycles[total + slot] = (uint32_t)(px < 0xFFFFFFFFull ? px : 0xFFFFFFFFull);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Work queue of the analytic tiles
// ---------------------------------------------------------------------------

// Per analytic tile the 16 features of the tile cost model, from one walk of
// the band's command list by one wave (the list entries whose tile-column
// range holds the tile are the commands the tile composites):
//   0 log2(1 + commands)          1 log2(1 + binned segments of them)
//   2 log2(1 + segments of commands starting left of the tile)
//   3 log2(1 + list entries of the band)
//   4..8 log2(1 + commands by style: linear, radial, conic, pattern, solid)
//   9 log2(1 + box commands)      10 log2(1 + commands not SRC_OVER)
//   11 A8 canvas                  12 valid rows / 16   13 valid columns / 64
//   14 log2(1 + command batches)  15 log2(1 + largest segment count of one command)
__global__ __launch_bounds__(256) void k_tile_features(FrameDev f, FrameScratch s, uint32_t n) {
  const uint32_t slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (slot >= n || slot >= s.tile_off[f.n_bands]) return;   // uniform per wave
  const uint4 th = s.tile_hdr[slot];
  const DevCanvas& cv = f.canvases[th.x];
  const uint32_t tcol = th.w & 0xFFFFu, band = th.w >> 16;
  uint32_t c[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // n, segs, segs_left, lin, rad, con, pat, solid, box, op, -, max
  for (uint32_t li = th.y + lane; li < th.z; li += 64) {
    const uint4 ent = s.list_ent[li];
    const uint32_t lo = ent.y & 0xFFFFu, hi = ent.y >> 16;
    if (tcol < lo || tcol > hi) continue;
    const DevCmd& cmd = f.cmds[ent.x];
    const uint32_t kind = f.styles[s.list_style[li]].kind;
    c[0] += 1u;
    c[1] += ent.w;
    c[2] += lo < tcol ? ent.w : 0u;
    c[3] += kind == kFetchLinear;
    c[4] += kind == kFetchRadial;
    c[5] += kind == kFetchConic;
    c[6] += kind >= kFetchPatternNN;
    c[7] += kind == kFetchSolid;
    c[8] += cmd.kind != kCmdAnalytic;
    c[9] += cmd.op != kOpSrcOver;
    c[11] = ent.w > c[11] ? ent.w : c[11];
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t v = __shfl_xor(c[k], o, 64);
      c[k] = k == 11 ? (v > c[k] ? v : c[k]) : c[k] + v;
    }
  }
  if (lane != 0) return;
  const uint32_t x0 = tcol * kTileW, y0 = band * kTileH;
  const uint32_t rows = cv.h - y0 < (uint32_t)kTileH ? cv.h - y0 : (uint32_t)kTileH;
  const uint32_t cols = cv.w - x0 < (uint32_t)kTileW ? cv.w - x0 : (uint32_t)kTileW;
  float* o = s.tile_feat + (size_t)slot * 16;
  o[0] = log2f(1.0f + (float)c[0]);
  o[1] = log2f(1.0f + (float)c[1]);
  o[2] = log2f(1.0f + (float)c[2]);
  o[3] = log2f(1.0f + (float)(th.z - th.y));
#pragma unroll
  for (int k = 0; k < 5; k++) o[4 + k] = log2f(1.0f + (float)c[3 + k]);
  o[9] = log2f(1.0f + (float)c[8]);
  o[10] = log2f(1.0f + (float)c[9]);
  o[11] = cv.format == kFmtA8 ? 1.0f : 0.0f;
  o[12] = (float)rows / (float)kTileH;
  o[13] = (float)cols / (float)kTileW;
  o[14] = log2f(1.0f + (float)((c[0] + kBatch - 1) / kBatch));
  o[15] = log2f(1.0f + (float)c[11]);
}

// Queue key of each analytic tile: its predicted cost (log2 ticks from the
// model's first output when `pred` is set, else log2 of binned segments +
// 16 x commands from the features) in 1/8 steps, as a bucket of a
// descending counting sort; bucket sizes into `hist` (256 entries, zeroed),
// counted per workgroup in LDS first (most tiles of a frame fall into a few
// buckets, so per-tile global atomics would queue on them).
constexpr uint32_t kQueueBuckets = 256;
__global__ __launch_bounds__(kQueueBuckets) void k_tile_keys(FrameDev f, FrameScratch s, uint32_t n, const float* __restrict__ pred) {
  __shared__ uint32_t lhist[kQueueBuckets];
  lhist[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot < n && slot < s.tile_off[f.n_bands]) {
    float cost;
    if (pred) {
      cost = pred[(size_t)slot * 2];
    }
    else {
      const float* x = s.tile_feat + (size_t)slot * 16;
      cost = log2f(1.0f + (exp2f(x[1]) - 1.0f) + 16.0f * (exp2f(x[0]) - 1.0f));
    }
    float q = cost * 8.0f;
    q = q != q ? 0.0f : q;   // a NaN prediction sorts last
    const uint32_t key = q <= 0.0f ? 0u : q >= (float)(kQueueBuckets - 1) ? kQueueBuckets - 1 : (uint32_t)q;
    const uint32_t bucket = kQueueBuckets - 1 - key;
    s.tile_key[slot] = bucket;
    atomicAdd(&lhist[bucket], 1u);
  }
  __syncthreads();
  if (lhist[threadIdx.x]) atomicAdd(&s.queue_hist[threadIdx.x], lhist[threadIdx.x]);
}

// Exclusive scan of the bucket sizes in place (one workgroup), and the queue
// head reset.
__global__ __launch_bounds__(kQueueBuckets) void k_queue_scan(FrameScratch s) {
  __shared__ uint32_t v[kQueueBuckets];
  const uint32_t t = threadIdx.x;
  v[t] = s.queue_hist[t];
  __syncthreads();
  for (uint32_t d = 1; d < kQueueBuckets; d <<= 1) {
    const uint32_t a = t >= d ? v[t - d] : 0u;
    __syncthreads();
    v[t] += a;
    __syncthreads();
  }
  s.queue_hist[t] = v[t] - s.queue_hist[t];
  if (t == 0) *s.queue_head = 0u;
}

// Queue positions: each workgroup takes one range per bucket from the
// bucket cursors and numbers its tiles within it in LDS (within a bucket
// the order is any: tiles are independent).
__global__ __launch_bounds__(kQueueBuckets) void k_tile_scatter(FrameDev f, FrameScratch s, uint32_t n) {
  __shared__ uint32_t lcount[kQueueBuckets], lbase[kQueueBuckets];
  lcount[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = slot < n && slot < s.tile_off[f.n_bands];
  uint32_t bucket = 0, rank = 0;
  if (on) {
    bucket = s.tile_key[slot];
    rank = atomicAdd(&lcount[bucket], 1u);
  }
  __syncthreads();
  if (lcount[threadIdx.x]) lbase[threadIdx.x] = atomicAdd(&s.queue_hist[threadIdx.x], lcount[threadIdx.x]);
  __syncthreads();
  if (on) s.queue[lbase[bucket] + rank] = slot;
}

// Canvas of a global band: the last one whose first band is <= b.
__device__ __forceinline__ uint32_t band_canvas(const FrameDev& f, uint32_t b) {
  uint32_t lo = 0, hi = f.n_canvases;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (f.canvases[mid].band_base <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// Per band: analytic fills present (a non-empty binned range) -> its tiles
// go to k_render; otherwise its tile groups go to k_render_box_fast (solid
// SRC_OVER / SRC_COPY aligned boxes only), or to k_render_box, the solid or
// the general instance.
__global__ void k_band_class(FrameDev f, FrameScratch s) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.n_bands) return;
  const DevCanvas& cv = f.canvases[band_canvas(f, b)];
  const uint64_t row_base = cv.mat_base + (uint64_t)(b - cv.band_base) * cv.n_cmds;
  const bool analytic = s.mat_off[row_base + cv.n_cmds] != s.mat_off[row_base];
  const bool solid = !s.band_nonsolid[b];
  const bool fast = solid && !s.band_slow[b];
  s.band_tiles[b] = analytic ? cv.tiles_x : 0u;
  s.band_fgroups[b] = !analytic && fast ? cv.groups_x : 0u;
  s.band_sgroups[b] = !analytic && solid && !fast ? cv.groups_x : 0u;
  s.band_ggroups[b] = !analytic && !solid ? cv.groups_x : 0u;
}

// Compacted headers: {canvas, the band's command list range, column | band
// << 16} for every tile of an analytic band and every group of a box band.
__global__ void k_tile_setup(FrameDev f, FrameScratch s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n_tiles) return;
  uint32_t lo = 0, hi = f.n_canvases;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (f.canvases[mid].tile_base <= i) lo = mid; else hi = mid;
  }
  const DevCanvas& cv = f.canvases[lo];
  const uint32_t local = i - cv.tile_base;
  const uint32_t band = local / cv.tiles_x, col = local % cv.tiles_x;
  const uint32_t gb = cv.band_base + band;
  if (!s.band_tiles[gb]) return;
  const uint64_t row_base = cv.mat_base + (uint64_t)band * cv.n_cmds;
  s.tile_hdr[s.tile_off[gb] + col] = make_uint4(lo, s.list_off[row_base], s.list_off[row_base + cv.n_cmds], col | (band << 16));
}

// Per box tile group two header words: {first tile index, the band's command
// list range, group column | band << 16} and the destination address of the
// group's first pixel {address low, address high, stride, width left of the
// group's first pixel | valid rows << 16 | A8 << 21 | tiles << 24}, so the
// compositor needs no canvas record.
__global__ void k_group_setup(FrameDev f, FrameScratch s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n_groups) return;
  uint32_t lo = 0, hi = f.n_canvases;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (f.canvases[mid].group_base <= i) lo = mid; else hi = mid;
  }
  const DevCanvas& cv = f.canvases[lo];
  const uint32_t local = i - cv.group_base;
  const uint32_t band = local / cv.groups_x, col = local % cv.groups_x;
  const uint32_t gb = cv.band_base + band;
  uint32_t slot;
  if (s.band_fgroups[gb]) slot = s.fgroup_off[gb] + col;
  else if (s.band_sgroups[gb]) slot = s.fgroup_off[f.n_bands] + s.sgroup_off[gb] + col;
  else if (s.band_ggroups[gb]) slot = s.fgroup_off[f.n_bands] + s.sgroup_off[f.n_bands] + s.ggroup_off[gb] + col;
  else return;
  const uint64_t row_base = cv.mat_base + (uint64_t)band * cv.n_cmds;
  const uint32_t tc0 = col * kGroupTiles;
  const uint32_t ntiles = cv.tiles_x - tc0 < (uint32_t)kGroupTiles ? cv.tiles_x - tc0 : (uint32_t)kGroupTiles;
  const uint32_t x0 = tc0 * kTileW, y0 = band * kTileH;
  const uint32_t rows = cv.h - y0 < (uint32_t)kTileH ? cv.h - y0 : (uint32_t)kTileH;
  const uint32_t a8 = cv.format == kFmtA8 ? 1u : 0u;
  const uint64_t ptr = cv.ptr + (uint64_t)((int64_t)y0 * cv.stride) + (uint64_t)x0 * (a8 ? 1u : 4u);
  uint4* h = s.group_hdr + 2 * (size_t)slot;
  h[0] = make_uint4(cv.tile_base + band * cv.tiles_x + tc0, s.list_off[row_base], s.list_off[row_base + cv.n_cmds],
                    col | (band << 16));
  h[1] = make_uint4((uint32_t)ptr, (uint32_t)(ptr >> 32), (uint32_t)cv.stride,
                    (cv.w - x0) | (rows << 16) | (a8 << 21) | (ntiles << 24));
}

// Bands with box fills only: one workgroup per group of up to 4 tiles of a
// band (256 x 16 pixels, 16 pixels per thread).  No coverage cells are
// needed, so the workgroup is light, and one walk of the band's command list
// and records serves four tiles.
//
// A thread's 16 pixels are four quads.  Solid styles (the pure compositing
// load): wave w takes rows 4w .. 4w + 3 of the group and lane l the pixels
// 4l .. 4l + 3 of each, so every load and store instruction of a wave moves
// one contiguous 1 KB row segment (lane-contiguous accesses measured 25 %
// faster than 64 contiguous bytes per thread).  Other styles keep one row
// per thread, four consecutive quads along it, so their per-row fetch setup
// is done once per command.
template<bool kSolid, bool kDither = false>
__device__ __forceinline__ void render_box(const FrameDev& f, const FrameScratch& s, uint32_t n) {
  const uint32_t slot = xcd_tile(blockIdx.x, n);
  if (slot >= n) return;
  // Headers: the fast bands' groups first, then the solid ones, then the
  // general ones.  The check is uniform, as in k_render.
  const uint32_t n_groups = (kSolid ? s.sgroup_off : s.ggroup_off)[f.n_bands];
  if (slot >= n_groups) return;
  const size_t hi = 2 * ((size_t)slot + s.fgroup_off[f.n_bands] + (kSolid ? 0u : s.sgroup_off[f.n_bands]));
  const uint4 gh = s.group_hdr[hi];
  const uint4 gd = s.group_hdr[hi + 1];
  const uint32_t pool_flags = s.flags ? *s.flags : 0u;
  const int t = threadIdx.x;
  const uint32_t lb = gh.y, le = gh.z;
  const uint32_t band = gh.w >> 16;
  const uint32_t tc0 = (gh.w & 0xFFFFu) * kGroupTiles;
  const uint32_t ntiles = gd.w >> 24;
  const uint32_t tc1 = tc0 + ntiles - 1;
  const uint32_t tile0 = gh.x;
  if (pool_flags) {
    if ((uint32_t)t < ntiles) s.tile_px[tile0 + t] = 0;
    return;
  }

  const int lane = t & 63, wave = t >> 6;
  const int X0 = (int)tc0 * kTileW;
  const int w = X0 + (int)(gd.w & 0xFFFFu);        // canvas width
  const int rows = (int)((gd.w >> 16) & 0x1Fu);    // rows of the band inside the canvas
  const bool a8 = (gd.w >> 21) & 1u;
  // quad g: group row qrow(g), first pixel qx(g)
  const int row0 = kSolid ? wave * 4 : t >> 4;
  auto qrow = [&](int g) { return kSolid ? row0 + g : row0; };
  auto qx = [&](int g) { return kSolid ? X0 + lane * 4 : X0 + (t & 15) * 16 + 4 * g; };
  const int tile_of_thread = kSolid ? lane >> 4 : (t & 15) >> 2;
  // the thread's pixel span, for skipping commands that miss it
  const int span_x0 = qx(0), span_x1 = kSolid ? qx(0) + 4 : qx(0) + 16;
  const int span_y0 = (int)band * kTileH + row0, span_y1 = span_y0 + (kSolid ? 4 : 1);
  const bool any_row = row0 < rows;

  // destination pointer of the group's row 0, offset so that pixel x is at
  // base + x * bpp
  uint8_t* const base = (uint8_t*)(((uint64_t)gd.y << 32) | gd.x) - (int64_t)X0 * (a8 ? 1 : 4);
  const int64_t stride = (int32_t)gd.z;
  uint32_t d[4][4];
#pragma unroll
  for (int g = 0; g < 4; g++) {
#pragma unroll
    for (int k = 0; k < 4; k++) d[g][k] = 0u;
    if (qrow(g) < rows && qx(g) < w) load4(base + qrow(g) * stride, qx(g), w, a8, d[g]);
  }

  bool touched = false;
  uint32_t npix = 0;
  __shared__ uint32_t sel_cmd[kTileThreads], sel_style[kTileThreads];
  __shared__ uint32_t sel_wave[kTileThreads / 64];
  __shared__ RenderRec rec[2][kBatch];

  for (uint32_t chunk = lb; chunk < le; chunk += kTileThreads) {
    const uint32_t li = chunk + (uint32_t)t;
    uint4 ent = make_uint4(0, 0, 0, 0);
    uint32_t style_i = 0;
    bool keep = false;
    if (li < le) {
      ent = s.list_ent[li];
      style_i = s.list_style[li];
      keep = (ent.y & 0xFFFFu) <= tc1 && (ent.y >> 16) >= tc0;
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) sel_wave[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = 0, n_sel = 0;
#pragma unroll
    for (int wv = 0; wv < kTileThreads / 64; wv++) {
      const uint32_t cw = sel_wave[wv];
      off += wv < wave ? cw : 0u;
      n_sel += cw;
    }
    if (keep) {
      const uint32_t slot = off + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
      sel_cmd[slot] = ent.x;
      sel_style[slot] = style_i;
    }
    __syncthreads();
    if (n_sel == 0) continue;

    {
      const uint32_t nb = n_sel < (uint32_t)kBatch ? n_sel : (uint32_t)kBatch;
      if (t < (int)nb * kRecWords) rec[0][t / kRecWords].w[t % kRecWords] = batch_rec_word(f, s, sel_cmd, sel_style, 0, t);
      if (t + kTileThreads < (int)nb * kRecWords) {
        const int gw = t + kTileThreads;
        rec[0][gw / kRecWords].w[gw % kRecWords] = batch_rec_word(f, s, sel_cmd, sel_style, 0, gw);
      }
    }
    __syncthreads();

    for (uint32_t kb = 0, bi = 0; kb < n_sel; kb += kBatch, bi++) {
      const int cur = (int)(bi & 1u);
      const uint32_t nb = n_sel - kb < (uint32_t)kBatch ? n_sel - kb : (uint32_t)kBatch;
      const uint32_t kn = kb + kBatch;
      const bool has_next = kn < n_sel;
      uint32_t pw0 = 0, pw1 = 0;
      int nwn = 0;
      if (has_next) {
        const uint32_t nbn = n_sel - kn < (uint32_t)kBatch ? n_sel - kn : (uint32_t)kBatch;
        nwn = (int)nbn * kRecWords;
        if (t < nwn) pw0 = batch_rec_word(f, s, sel_cmd, sel_style, kn, t);
        if (t + kTileThreads < nwn) pw1 = batch_rec_word(f, s, sel_cmd, sel_style, kn, t + kTileThreads);
      }
      for (uint32_t b = 0; b < nb; b++) {
        const DevCmd& c = rec[cur][b].r.cmd;
        if (c.kind == kCmdAnalytic || !any_row) continue;
        if (span_y1 <= c.box[1] || span_y0 >= c.box[3] || span_x1 <= c.box[0] || span_x0 >= c.box[2]) continue;
        touched = true;
        const DevStyle& st = rec[cur][b].r.style;
        const pipe::Row rs = kSolid ? pipe::Row{} : pipe::row_setup(st, span_y0, (int)c.fetch_row0, 0u);
        if (kSolid && !a8 && c.kind == kCmdBoxA && (c.op == kOpSrcOver || c.op == kOpSrcCopy) && c.alpha) {
          // The pure compositing load: a solid colour under one constant
          // mask (an aligned box's alpha) over PRGB32.  Everything that
          // depends on the source alone is formed once per command; per
          // pixel only the destination terms of composite32 remain (same
          // arithmetic, so the same bits), selected where the box and the
          // canvas cover the pixel.
          const uint32_t m = c.alpha, src = st.solid;
          const bool over = c.op == kOpSrcOver;
          const uint32_t s_rb = (src & 0x00FF00FFu) * m, s_ag = ((src >> 8) & 0x00FF00FFu) * m;
          const uint32_t s1 = pipe::div255x2(s_rb) | (pipe::div255x2(s_ag) << 8);
          const uint32_t k = over ? (s1 >> 24) ^ 0xFFu : m ^ 0xFFu;   // destination factor
          const uint32_t add_rb = over ? 0u : s_rb, add_ag = over ? 0u : s_ag;
          const uint32_t add_px = over ? s1 : 0u;
          const int bx0 = c.box[0], bx1 = c.box[2] < w ? c.box[2] : w;
#pragma unroll
          for (int g = 0; g < 4; g++) {
            const int yg = (int)band * kTileH + qrow(g);
            const bool row_in = qrow(g) < rows && yg >= c.box[1] && yg < c.box[3];
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const bool in = row_in && span_x0 + j >= bx0 && span_x0 + j < bx1;
              const uint32_t dv = d[g][j];
              const uint32_t rb = pipe::div255x2((dv & 0x00FF00FFu) * k + add_rb);
              const uint32_t ag = pipe::div255x2(((dv >> 8) & 0x00FF00FFu) * k + add_ag);
              d[g][j] = in ? (rb | (ag << 8)) + add_px : dv;
              npix += in ? 1u : 0u;
            }
          }
          continue;
        }
        // One composite body for the four pixel quads: quad g is always in
        // d[0], the quads rotate through it.
#pragma unroll 1
        for (int g = 0; g < 4; g++) {
          uint32_t m[4];
          const int px = qx(g);
          box_masks4(f, c, (int)band * kTileH + qrow(g), px, m);
          if (qrow(g) < rows) npix += composite4<kSolid, kDither>(f, c, st, rs, px, w, a8, m, d[0]);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const uint32_t r0 = d[0][k];
            d[0][k] = d[1][k];
            d[1][k] = d[2][k];
            d[2][k] = d[3][k];
            d[3][k] = r0;
          }
        }
      }
      if (has_next) {
        if (t < nwn) rec[cur ^ 1][t / kRecWords].w[t % kRecWords] = pw0;
        if (t + kTileThreads < nwn) {
          const int gw = t + kTileThreads;
          rec[cur ^ 1][gw / kRecWords].w[gw % kRecWords] = pw1;
        }
      }
      __syncthreads();
    }
  }

  // Composited pixels per tile of the group (all of a thread's pixels lie in
  // one tile).
  __shared__ uint32_t tile_cnt[kGroupTiles];
  if (t < kGroupTiles) tile_cnt[t] = 0;
  __syncthreads();
  if (npix) atomicAdd(&tile_cnt[tile_of_thread], npix);
  __syncthreads();
  if ((uint32_t)t < ntiles) s.tile_px[tile0 + t] = tile_cnt[t];

  if (!touched) return;
#pragma unroll
  for (int g = 0; g < 4; g++)
    if (qrow(g) < rows && qx(g) < w) store4(base + qrow(g) * stride, qx(g), w, a8, d[g]);
}

// Bands whose box fills are all aligned boxes of a solid colour with
// SRC_OVER or SRC_COPY on a 32-bit canvas -- the pure compositing load.  One
// workgroup per group of up to 4 tiles (256 x 16 pixels); wave w takes rows
// 4w .. 4w + 3 and lane l the pixels 4l .. 4l + 3 of each, so every load and
// store instruction of a wave moves one contiguous 1 KB row segment.  The
// band's commands are read 256 at a time into a compact record each (box,
// alpha, op, colour), with no compaction or selection pass; a thread skips
// the commands whose box misses its pixels.  Per command everything that
// depends on the source alone is formed once, per pixel only the
// destination terms of pipe::composite32 remain (same arithmetic, so the
// same bits).  Few registers and records of 24 bytes let eight workgroups
// share a CU, which is what keeps enough loads in flight.
struct FastBoxRec {
  int32_t x0, y0, x1, y1;   // the box (pixels, x1 / y1 exclusive)
  uint32_t alpha_op;        // alpha | op << 8
  uint32_t solid;           // premultiplied colour
};

__global__ __launch_bounds__(kTileThreads, 8) void k_render_box_fast(FrameDev f, FrameScratch s, uint32_t n) {
  __shared__ FastBoxRec rec[kTileThreads];
  __shared__ uint32_t tile_cnt[kGroupTiles];
  const uint32_t slot = xcd_tile(blockIdx.x, n);
  if (slot >= n) return;
  // the fast groups' headers come first: 2 * slot lies inside the header
  // buffer (two words per group of the frame) for every slot below the
  // launch count, so the count, the header and the pool flag load together
  const uint32_t n_groups = s.fgroup_off[f.n_bands];
  const uint4 gh = s.group_hdr[2 * (size_t)slot];
  const uint4 gd = s.group_hdr[2 * (size_t)slot + 1];
  const uint32_t pool_flags = s.flags ? *s.flags : 0u;
  if (slot >= n_groups) return;   // uniform
  const int t = threadIdx.x;
  const uint32_t lb = gh.y, le = gh.z;
  const uint32_t band = gh.w >> 16;
  const uint32_t tc0 = (gh.w & 0xFFFFu) * kGroupTiles;
  const uint32_t ntiles = gd.w >> 24;
  const uint32_t tile0 = gh.x;
  if (pool_flags) {
    if ((uint32_t)t < ntiles) s.tile_px[tile0 + t] = 0;
    return;
  }
  const int lane = t & 63, row0 = (t >> 6) * 4;
  const int X0 = (int)tc0 * kTileW;
  const int w = X0 + (int)(gd.w & 0xFFFFu);        // canvas width
  const int rows = (int)((gd.w >> 16) & 0x1Fu);    // rows of the band inside the canvas
  const int px = X0 + lane * 4;
  const int y0 = (int)band * kTileH + row0;
  uint8_t* const base = (uint8_t*)(((uint64_t)gd.y << 32) | gd.x) - (int64_t)X0 * 4;
  const int64_t stride = (int32_t)gd.z;
  uint32_t d[4][4];
#pragma unroll
  for (int g = 0; g < 4; g++) {
#pragma unroll
    for (int k = 0; k < 4; k++) d[g][k] = 0u;
    if (row0 + g < rows && px < w) load4(base + (row0 + g) * stride, px, w, false, d[g]);
  }

  uint32_t npix = 0;
  bool touched = false;
  for (uint32_t chunk = lb; chunk < le; chunk += kTileThreads) {
    const uint32_t nc = le - chunk < (uint32_t)kTileThreads ? le - chunk : (uint32_t)kTileThreads;
    __syncthreads();   // the previous chunk's records are consumed
    if ((uint32_t)t < nc) {
      const DevCmd& c = f.cmds[s.list_ent[chunk + t].x];
      rec[t] = FastBoxRec{c.box[0], c.box[1], c.box[2], c.box[3], c.alpha | (c.op << 8),
                          f.styles[s.list_style[chunk + t]].solid};
    }
    __syncthreads();
    for (uint32_t k = 0; k < nc; k++) {
      const FastBoxRec r = rec[k];
      if (y0 + 4 <= r.y0 || y0 >= r.y1 || px + 4 <= r.x0 || px >= r.x1 || row0 >= rows) continue;
      touched = true;
      const uint32_t m = r.alpha_op & 0xFFu;
      const bool over = (r.alpha_op >> 8) == kOpSrcOver;
      const uint32_t s_rb = (r.solid & 0x00FF00FFu) * m, s_ag = ((r.solid >> 8) & 0x00FF00FFu) * m;
      const uint32_t s1 = pipe::div255x2(s_rb) | (pipe::div255x2(s_ag) << 8);
      const uint32_t kd = over ? (s1 >> 24) ^ 0xFFu : m ^ 0xFFu;   // destination factor
      const uint32_t add_rb = over ? 0u : s_rb, add_ag = over ? 0u : s_ag;
      const uint32_t add_px = over ? s1 : 0u;
      const int bx1 = r.x1 < w ? r.x1 : w;
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const bool row_in = row0 + g < rows && y0 + g >= r.y0 && y0 + g < r.y1;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const bool in = row_in && px + j >= r.x0 && px + j < bx1;
          const uint32_t dv = d[g][j];
          const uint32_t rb = pipe::div255x2((dv & 0x00FF00FFu) * kd + add_rb);
          const uint32_t ag = pipe::div255x2(((dv >> 8) & 0x00FF00FFu) * kd + add_ag);
          d[g][j] = in ? (rb | (ag << 8)) + add_px : dv;
          npix += in ? 1u : 0u;
        }
      }
    }
  }

  // composited pixels per tile of the group (a thread's pixels lie in tile
  // lane / 16)
  if (t < kGroupTiles) tile_cnt[t] = 0;
  __syncthreads();
  if (npix) atomicAdd(&tile_cnt[lane >> 4], npix);
  __syncthreads();
  if ((uint32_t)t < ntiles) s.tile_px[tile0 + t] = tile_cnt[t];

  if (!touched) return;
#pragma unroll
  for (int g = 0; g < 4; g++)
    if (row0 + g < rows && px < w) store4(base + (row0 + g) * stride, px, w, false, d[g]);
}

// *out += sum of v[0, n): per-workgroup partial sums, one atomic each
// (integer addition, so the result does not depend on the order).
__global__ __launch_bounds__(256) void k_sum_u64(const unsigned long long* v, uint32_t n, unsigned long long* out) {
  __shared__ unsigned long long part[256];
  unsigned long long acc = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) acc += v[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(out, part[0]);
}

// Solid styles need few registers: seven workgroups per CU (72 VGPRs, no
// spills), where the general instance fits four.  Capping it at 64 for eight
// spills and runs slower.
__global__ __launch_bounds__(kTileThreads, 6) void k_render_box_solid(FrameDev f, FrameScratch s, uint32_t n) {
  render_box<true>(f, s, n);
}

template<bool kDither>
__global__ __launch_bounds__(kTileThreads) void k_render_box(FrameDev f, FrameScratch s, uint32_t n) {
  render_box<false, kDither>(f, s, n);
}

void launch_band_class(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (!f.n_tiles) return;
  hipLaunchKernelGGL(k_band_class, dim3((f.n_bands + 255) / 256), dim3(256), 0, st, f, s);
  scan_exclusive(s.band_tiles, s.tile_off, f.n_bands, s.scan_tmp, st);
  scan_exclusive(s.band_fgroups, s.fgroup_off, f.n_bands, s.scan_tmp, st);
  scan_exclusive(s.band_sgroups, s.sgroup_off, f.n_bands, s.scan_tmp, st);
  scan_exclusive(s.band_ggroups, s.ggroup_off, f.n_bands, s.scan_tmp, st);
}

void launch_render_setup(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (!f.n_tiles) return;
  hipLaunchKernelGGL(k_tile_setup, dim3((f.n_tiles + 255) / 256), dim3(256), 0, st, f, s);
  hipLaunchKernelGGL(k_group_setup, dim3((f.n_groups + 255) / 256), dim3(256), 0, st, f, s);
}

void launch_render(const FrameDev& f, const FrameScratch& s, uint32_t n_a, uint32_t n_f, uint32_t n_s, uint32_t n_g,
                   hipStream_t st) {
  if (!f.n_tiles) return;
  // frames with dithered gradients take the instances that carry the dither
  // path; the others leave it out
  if (n_a && f.dither) hipLaunchKernelGGL(k_render<true>, dim3((n_a + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_a);
  else if (n_a) hipLaunchKernelGGL(k_render<false>, dim3((n_a + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_a);
  if (n_f) hipLaunchKernelGGL(k_render_box_fast, dim3((n_f + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_f);
  if (n_s) hipLaunchKernelGGL(k_render_box_solid, dim3((n_s + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_s);
  if (n_g && f.dither) hipLaunchKernelGGL(k_render_box<true>, dim3((n_g + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_g);
  else if (n_g) hipLaunchKernelGGL(k_render_box<false>, dim3((n_g + kXcds - 1) / kXcds * kXcds), dim3(kTileThreads), 0, st, f, s, n_g);
}

void launch_tile_features(const FrameDev& f, const FrameScratch& s, uint32_t n_a, hipStream_t st) {
  if (!f.n_tiles || !n_a) return;
  hipLaunchKernelGGL(k_tile_features, dim3((n_a + 3) / 4), dim3(256), 0, st, f, s, n_a);
}

void launch_tile_queue(const FrameDev& f, const FrameScratch& s, uint32_t n_a, const float* pred, hipStream_t st) {
  if (!f.n_tiles || !n_a) return;
  launch_fill_u32(s.queue_hist, 0u, kQueueBuckets, st);
  hipLaunchKernelGGL(k_tile_keys, dim3((n_a + kQueueBuckets - 1) / kQueueBuckets), dim3(kQueueBuckets), 0, st, f, s, n_a, pred);
  hipLaunchKernelGGL(k_queue_scan, dim3(1), dim3(kQueueBuckets), 0, st, s);
  hipLaunchKernelGGL(k_tile_scatter, dim3((n_a + kQueueBuckets - 1) / kQueueBuckets), dim3(kQueueBuckets), 0, st, f, s, n_a);
}

void launch_render_queue(const FrameDev& f, const FrameScratch& s, uint32_t grid, uint32_t wgs_per_cu, hipStream_t st) {
  if (!f.n_tiles || !grid) return;
  if (!s.queue) launch_fill_u32(s.queue_head, 0u, 1, st);
  // Four workgroups per CU need the four-occupancy instance (128 VGPRs,
  // some spilled); fewer take the three-occupancy one (170 VGPRs).
  if (f.dither) {
    if (s.tile_cycles) hipLaunchKernelGGL((k_render_queue<true, 4, true>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
    else if (wgs_per_cu >= 4) hipLaunchKernelGGL((k_render_queue<false, 4, true>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
    else hipLaunchKernelGGL((k_render_queue<false, 3, true>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
  }
  else if (s.tile_cycles) hipLaunchKernelGGL((k_render_queue<true, 4, false>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
  else if (wgs_per_cu >= 4) hipLaunchKernelGGL((k_render_queue<false, 4, false>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
  else hipLaunchKernelGGL((k_render_queue<false, 3, false>), dim3(grid), dim3(kTileThreads), 0, st, f, s);
}

void launch_render_stats(const FrameDev& f, const FrameScratch& s, hipStream_t st) {
  if (!f.n_tiles) return;
  uint32_t g = (f.n_tiles + 4095) / 4096;
  if (g > 256) g = 256;
  hipLaunchKernelGGL(k_sum_u64, dim3(g), dim3(256), 0, st, (const unsigned long long*)s.tile_px, f.n_tiles, s.pixels);
}

}  // namespace b2amd
