// Learned sizing kernels: per-command features, MLP inference, and the
// forward / backward / Adam kernels that train the MLP on the GPU.
#include <hip/hip_runtime.h>
#include <math.h>

#include "cost_model.h"
#include "device_types.h"

namespace b2amd {
namespace cm {

// ---------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------

struct Pt { double x, y; };

__device__ __forceinline__ Pt fx_pt(const CmdXform& m, const double* vtx, uint32_t i) {
  double x = vtx[2 * i], y = vtx[2 * i + 1];
  if (m.scale_only) return Pt{x * m.m00 + m.m20, y * m.m11 + m.m21};
  return Pt{x * m.m00 + y * m.m10 + m.m20, x * m.m01 + y * m.m11 + m.m21};
}

__device__ __forceinline__ double dist(Pt a, Pt b) { return sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y)); }

__device__ __forceinline__ int32_t clamp_fx(double v) {
  if (!(v > -1073741824.0)) return -1073741824;
  if (!(v < 1073741824.0)) return 1073741824;
  return (int32_t)v;
}

// Sums in 1/16 px, as integers, so the features do not depend on the order
// of the atomic additions.
__device__ __forceinline__ uint32_t px16(double fixed_len) {
  double v = fixed_len * (16.0 / 256.0);
  return v > 4.0e9 ? 4000000000u : (uint32_t)v;
}

__device__ __forceinline__ unsigned long long shfl_down_u64(unsigned long long v, int d) {
  const uint32_t lo = __shfl_down((uint32_t)v, d, 64), hi = __shfl_down((uint32_t)(v >> 32), d, 64);
  return ((unsigned long long)hi << 32) | lo;
}

// Per-source feature terms, summed per command.  The lanes of a wave first
// reduce the runs of equal command indices among themselves (sources of a
// command are contiguous), then the first lane of each run adds the run's
// totals with one set of atomics: integer sums and min / max, so the result
// does not depend on the grouping.
__global__ void k_features_acc(FrameDev f, FrameModel m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(threadIdx.x & 63u);
  uint32_t key = 0xFFFFFFFFu;
  uint32_t cnt[6] = {0, 0, 0, 0, 0, 0};          // line, quad, cubic, vis, left, right
  unsigned long long sum[4] = {0, 0, 0, 0};      // chord, poly, flat, flat_vis
  int32_t bx0 = 1073741824, by0 = 1073741824, bx1 = -1073741824, by1 = -1073741824;
  if (i < f.n_src) {
    const SrcSeg sg = f.src[i];
    const DevCmd& cmd = f.cmds[sg.cmd];
    if (cmd.kind == kCmdAnalytic) {
      const DevCanvas& cv = f.canvases[cmd.canvas];
      const CmdXform& x = f.xforms[sg.cmd];
      Pt p[4];
      const int np = sg.kind == kSrcLine ? 2 : sg.kind == kSrcQuad ? 3 : 4;
      p[0] = fx_pt(x, f.vtx, sg.i0);
      p[1] = fx_pt(x, f.vtx, sg.i1);
      if (np > 2) p[2] = fx_pt(x, f.vtx, sg.i1 + 1);
      if (np > 3) p[3] = fx_pt(x, f.vtx, sg.i1 + 2);
      double x0 = p[0].x, x1 = p[0].x, y0 = p[0].y, y1 = p[0].y, poly = 0.0;
      bool finite = true;
      for (int k = 0; k < np; k++) {
        finite = finite && isfinite(p[k].x) && isfinite(p[k].y);
        x0 = fmin(x0, p[k].x); x1 = fmax(x1, p[k].x);
        y0 = fmin(y0, p[k].y); y1 = fmax(y1, p[k].y);
        if (k) poly += dist(p[k - 1], p[k]);
      }
      if (finite) {
        double dd = 0.0;
        if (np == 3) dd = dist(Pt{p[0].x - 2.0 * p[1].x + p[2].x, p[0].y - 2.0 * p[1].y + p[2].y}, Pt{0.0, 0.0});
        if (np == 4) {
          double a = dist(Pt{p[0].x - 2.0 * p[1].x + p[2].x, p[0].y - 2.0 * p[1].y + p[2].y}, Pt{0.0, 0.0});
          double b = dist(Pt{p[1].x - 2.0 * p[2].x + p[3].x, p[1].y - 2.0 * p[2].y + p[3].y}, Pt{0.0, 0.0});
          dd = fmax(a, b);
        }
        const double W = (double)cv.w * 256.0, H = (double)cv.h * 256.0;
        const bool vis = y1 > 0.0 && y0 < H && y1 > y0;
        const uint32_t flat = px16(sqrt(dd * 256.0));   // sqrt of |2nd difference| in px, x16
        key = sg.cmd;
        cnt[np - 2] = 1u;
        if (vis) {
          cnt[3] = 1u;
          cnt[4] = x0 < 0.0 ? 1u : 0u;
          cnt[5] = x1 > W ? 1u : 0u;
          sum[3] = flat;
        }
        sum[0] = px16(dist(p[0], p[np - 1]));
        sum[1] = px16(poly);
        sum[2] = flat;
        bx0 = clamp_fx(floor(x0));
        by0 = clamp_fx(floor(y0));
        bx1 = clamp_fx(ceil(x1));
        by1 = clamp_fx(ceil(y1));
      }
    }
  }
  // runs of equal keys: run index = number of run starts up to this lane
  const uint32_t prev = __shfl_up(key, 1, 64);   // every lane takes part in the shuffle
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const uint32_t run = (uint32_t)__popcll(heads & ((2ull << lane) - 1ull));
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t orun = __shfl_down(run, d, 64);
    const bool take = lane + d < 64 && orun == run;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const uint32_t o = __shfl_down(cnt[k], d, 64);
      if (take) cnt[k] += o;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const unsigned long long o = shfl_down_u64(sum[k], d);
      if (take) sum[k] += o;
    }
    const int32_t a = __shfl_down(bx0, d, 64), b = __shfl_down(by0, d, 64);
    const int32_t c = __shfl_down(bx1, d, 64), e = __shfl_down(by1, d, 64);
    if (take) {
      bx0 = a < bx0 ? a : bx0;
      by0 = b < by0 ? b : by0;
      bx1 = c > bx1 ? c : bx1;
      by1 = e > by1 ? e : by1;
    }
  }
  if (!head || key == 0xFFFFFFFFu) return;
  CmdAcc* acc = &m.acc[key];
  if (cnt[0]) atomicAdd(&acc->n_line, cnt[0]);
  if (cnt[1]) atomicAdd(&acc->n_quad, cnt[1]);
  if (cnt[2]) atomicAdd(&acc->n_cubic, cnt[2]);
  if (cnt[3]) {
    atomicAdd(&acc->n_vis, cnt[3]);
    if (cnt[4]) atomicAdd(&acc->n_left, cnt[4]);
    if (cnt[5]) atomicAdd(&acc->n_right, cnt[5]);
    atomicAdd(&acc->flat_vis, sum[3]);
  }
  atomicAdd(&acc->chord, sum[0]);
  atomicAdd(&acc->poly, sum[1]);
  atomicAdd(&acc->flat, sum[2]);
  atomicMin(&acc->bx0, bx0);
  atomicMin(&acc->by0, by0);
  atomicMax(&acc->bx1, bx1);
  atomicMax(&acc->by1, by1);
}

__global__ void k_features_init(FrameDev f, FrameModel m) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.n_cmds) return;
  CmdAcc a{};
  a.bx0 = a.by0 = 1073741824;
  a.bx1 = a.by1 = -1073741824;
  m.acc[c] = a;
  if (m.actual_segs) m.actual_segs[c] = 0;
  if (m.actual_bins) m.actual_bins[c] = 0;
  if (c == 0) {
    m.totals[0] = 0ull;
    m.totals[1] = 0ull;
  }
}

__device__ __forceinline__ float lg(double v) { return (float)log2(1.0 + (v > 0.0 ? v : 0.0)); }

__global__ void k_features_finalize(FrameDev f, FrameModel m) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.n_cmds) return;
  const CmdAcc a = m.acc[c];
  const DevCanvas& cv = f.canvases[f.cmds[c].canvas];
  const double W = (double)cv.w * 256.0, H = (double)cv.h * 256.0;
  const uint32_t n = a.n_line + a.n_quad + a.n_cubic;
  float* o = m.features + (size_t)c * kIn;
  if (!n) {
    for (int k = 0; k < kIn; k++) o[k] = 0.0f;
    return;
  }
  const double bw = (double)a.bx1 - (double)a.bx0, bh = (double)a.by1 - (double)a.by0;
  const double vy0 = fmax((double)a.by0, 0.0), vy1 = fmin((double)a.by1, H);
  const double vx0 = fmax((double)a.bx0, 0.0), vx1 = fmin((double)a.bx1, W);
  const double vh = vy1 > vy0 ? (vy1 - vy0) / 256.0 : 0.0;
  const double vw = vx1 > vx0 ? (vx1 - vx0) / 256.0 : 0.0;
  o[0] = lg(a.n_line);
  o[1] = lg(a.n_quad);
  o[2] = lg(a.n_cubic);
  o[3] = lg(a.chord / 16.0);
  o[4] = lg(a.poly / 16.0);
  o[5] = lg(a.flat / 16.0);
  o[6] = lg(bw / 256.0);
  o[7] = lg(bh / 256.0);
  o[8] = lg(vh);
  o[9] = lg(vw);
  o[10] = lg(a.n_vis);
  o[11] = lg(a.n_left);
  o[12] = lg(a.n_right);
  o[13] = lg(a.flat_vis / 16.0);
  o[14] = lg(vh / (double)kTileH);
  o[15] = (float)a.n_vis / (float)n;
}

void launch_features(const FrameDev& f, const FrameModel& m, hipStream_t st) {
  if (!f.n_cmds) return;
  hipLaunchKernelGGL(k_features_init, dim3((f.n_cmds + 255) / 256), dim3(256), 0, st, f, m);
  if (f.n_src) hipLaunchKernelGGL(k_features_acc, dim3((f.n_src + 255) / 256), dim3(256), 0, st, f, m);
  hipLaunchKernelGGL(k_features_finalize, dim3((f.n_cmds + 255) / 256), dim3(256), 0, st, f, m);
}

// ---------------------------------------------------------------------------
// MLP forward (one thread per row; weights are read with wave-uniform
// indices, so they come through the scalar cache)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void mlp_row(const float* __restrict__ P, const float* x, float* y) {
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float acc = P[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; i++) acc += x[i] * P[kW1 + i * kH + j];
    h1[j] = acc > 0.0f ? acc : 0.0f;
  }
  float y0 = P[kB3 + 0], y1 = P[kB3 + 1];
  for (int j = 0; j < kH; j++) {
    float acc = P[kB2 + j];
#pragma unroll
    for (int k = 0; k < kH; k++) acc += h1[k] * P[kW2 + k * kH + j];
    float h2 = acc > 0.0f ? acc : 0.0f;
    y0 += h2 * P[kW3 + j * kOut + 0];
    y1 += h2 * P[kW3 + j * kOut + 1];
  }
  y[0] = y0;
  y[1] = y1;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint32_t n, const float* __restrict__ P,
                                                     float* __restrict__ Y) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) x[i] = X[(size_t)r * kIn + i];
  float y[kOut];
  mlp_row(P, x, y);
  Y[(size_t)r * kOut + 0] = y[0];
  Y[(size_t)r * kOut + 1] = y[1];
}

void launch_mlp_forward(const float* X, uint32_t n, const float* params, float* Y, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, st, X, n, params, Y);
}

// Per-command inference for a frame; predicted counts are summed into the
// frame totals (rounded up to whole segments per command).
__global__ __launch_bounds__(256) void k_infer_frame(FrameDev f, FrameModel m) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.n_cmds) return;
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) x[i] = m.features[(size_t)c * kIn + i];
  float y[kOut] = {0.0f, 0.0f};
  if (f.cmds[c].kind == kCmdAnalytic && (x[0] + x[1] + x[2]) > 0.0f) mlp_row(m.params, x, y);
  m.pred[(size_t)c * kOut + 0] = y[0];
  m.pred[(size_t)c * kOut + 1] = y[1];
  float s = ceilf(exp2f(fminf(y[0], 30.0f)) - 1.0f), b = ceilf(exp2f(fminf(y[1], 30.0f)) - 1.0f);
  atomicAdd(&m.totals[0], (unsigned long long)(s > 0.0f ? s : 0.0f));
  atomicAdd(&m.totals[1], (unsigned long long)(b > 0.0f ? b : 0.0f));
}

void launch_infer_frame(const FrameDev& f, const FrameModel& m, hipStream_t st) {
  if (f.n_cmds) hipLaunchKernelGGL(k_infer_frame, dim3((f.n_cmds + 255) / 256), dim3(256), 0, st, f, m);
}

// Exact per-command counts of a frame that ran the counting pass (the
// training targets).
__global__ void k_actual_segs(FrameDev f, FrameScratch s, FrameModel m, uint32_t n_segs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_segs) return;
  atomicAdd(&m.actual_segs[s.seg_cmd[i]], 1u);
}

__global__ void k_actual_bins(FrameDev f, FrameScratch s, FrameModel m, uint32_t n_segs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_segs) return;
  const DevSeg g = s.segs[i];
  const uint32_t cmd = s.seg_cmd[i];
  const DevCanvas& cv = f.canvases[f.cmds[cmd].canvas];
  int b0 = (g.y0 >> 8) / kTileH, b1 = ((g.y1 - 1) >> 8) / kTileH;
  b0 = b0 < 0 ? 0 : b0;
  b1 = b1 > (int)cv.bands - 1 ? (int)cv.bands - 1 : b1;
  if (b1 >= b0) atomicAdd(&m.actual_bins[cmd], (uint32_t)(b1 - b0 + 1));
}

void launch_actuals(const FrameDev& f, const FrameScratch& s, const FrameModel& m, uint32_t n_segs, hipStream_t st) {
  if (n_segs) hipLaunchKernelGGL(k_actual_segs, dim3((n_segs + 255) / 256), dim3(256), 0, st, f, s, m, n_segs);
  if (n_segs) hipLaunchKernelGGL(k_actual_bins, dim3((n_segs + 255) / 256), dim3(256), 0, st, f, s, m, n_segs);
}

// ---------------------------------------------------------------------------
// Training: gradient of the mean squared error, 64 rows per workgroup
// ---------------------------------------------------------------------------

constexpr int kRows = 64;
constexpr int kPad = kH + 1;   // LDS row pitch (breaks the 64-way bank conflict)

__global__ __launch_bounds__(kRows) void k_mlp_grad(const float* __restrict__ X, const float* __restrict__ T, uint32_t n,
                                                    const float* __restrict__ P, float* __restrict__ partial) {
  __shared__ float sx[kRows][kIn + 1];
  __shared__ float sh1[kRows][kPad];
  __shared__ float sd1[kRows][kPad];
  __shared__ float sd2[kRows][kPad];
  __shared__ float sdy[kRows][kOut];
  const int t = threadIdx.x;
  const uint32_t r = blockIdx.x * kRows + t;
  const bool valid = r < n;

  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) {
    x[i] = valid ? X[(size_t)r * kIn + i] : 0.0f;
    sx[t][i] = x[i];
  }
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float acc = P[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; i++) acc += x[i] * P[kW1 + i * kH + j];
    h1[j] = acc > 0.0f ? acc : 0.0f;
    sh1[t][j] = h1[j];
  }
  float y0 = P[kB3 + 0], y1 = P[kB3 + 1];
  for (int j = 0; j < kH; j++) {
    float acc = P[kB2 + j];
#pragma unroll
    for (int k = 0; k < kH; k++) acc += h1[k] * P[kW2 + k * kH + j];
    float h2 = acc > 0.0f ? acc : 0.0f;
    sd2[t][j] = h2;
    y0 += h2 * P[kW3 + j * kOut + 0];
    y1 += h2 * P[kW3 + j * kOut + 1];
  }
  const float e0 = valid ? y0 - T[(size_t)r * kOut + 0] : 0.0f;
  const float e1 = valid ? y1 - T[(size_t)r * kOut + 1] : 0.0f;
  const float dy0 = 2.0f * e0, dy1 = 2.0f * e1;
  sdy[t][0] = dy0;
  sdy[t][1] = dy1;
  float loss = e0 * e0 + e1 * e1;
  __syncthreads();

  float* out = partial + (size_t)blockIdx.x * (kParams + 1);
  // dW3 [j][o] = sum_s h2[s][j] dy[s][o]; db3
  for (int p = t; p < kH * kOut; p += kRows) {
    const int j = p / kOut, o = p % kOut;
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sd2[s][j] * sdy[s][o];
    out[kW3 + p] = acc;
  }
  if (t < kOut) {
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sdy[s][t];
    out[kB3 + t] = acc;
  }
  __syncthreads();
  // dh2 = (h2 > 0) (dy . W3[j]) over this row
  for (int j = 0; j < kH; j++) {
    const float h2 = sd2[t][j];
    sd2[t][j] = h2 > 0.0f ? dy0 * P[kW3 + j * kOut + 0] + dy1 * P[kW3 + j * kOut + 1] : 0.0f;
  }
  // dh1 = (h1 > 0) (W2 dh2)
#pragma unroll
  for (int k = 0; k < kH; k++) {
    float acc = 0.0f;
    for (int j = 0; j < kH; j++) acc += sd2[t][j] * P[kW2 + k * kH + j];
    sd1[t][k] = h1[k] > 0.0f ? acc : 0.0f;
  }
  __syncthreads();
  // dW1 [i][j] = sum_s x[s][i] dh1[s][j]
  for (int p = t; p < kIn * kH; p += kRows) {
    const int i = p / kH, j = p % kH;
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sx[s][i] * sd1[s][j];
    out[kW1 + p] = acc;
  }
  {
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sd1[s][t];
    out[kB1 + t] = acc;
  }
  // dW2 [k][j] = sum_s h1[s][k] dh2[s][j]
  for (int p = t; p < kH * kH; p += kRows) {
    const int k = p / kH, j = p % kH;
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sh1[s][k] * sd2[s][j];
    out[kW2 + p] = acc;
  }
  {
    float acc = 0.0f;
    for (int s = 0; s < kRows; s++) acc += sd2[s][t];
    out[kB2 + t] = acc;
  }
  // loss: wave sum (one wave per workgroup)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
  if (t == 0) out[kParams] = loss;
}

uint32_t mlp_groups(uint32_t n) { return (n + kRows - 1) / kRows; }

void launch_mlp_grad(const float* X, const float* T, uint32_t n, const float* params, float* partial, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_mlp_grad, dim3(mlp_groups(n)), dim3(kRows), 0, st, X, T, n, params, partial);
}

__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t groups, float inv_n, float* __restrict__ grad) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > (uint32_t)kParams) return;
  float acc = 0.0f;
  for (uint32_t g = 0; g < groups; g++) acc += partial[(size_t)g * (kParams + 1) + p];
  grad[p] = acc * inv_n;
}

void launch_mlp_reduce(const float* partial, uint32_t groups, uint32_t n, float* grad, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + 255) / 256), dim3(256), 0, st, partial, groups, 1.0f / (float)n, grad);
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       float lr, float c1, float c2) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gi = g[p];
  float a = b1 * m1[p] + (1.0f - b1) * gi;
  float b = b2 * m2[p] + (1.0f - b2) * gi * gi;
  m1[p] = a;
  m2[p] = b;
  P[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

void launch_adam(float* params, const float* grad, float* m1, float* m2, uint32_t step, float lr, hipStream_t st) {
  const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, st, params, grad, m1, m2, lr, c1, c2);
}

}  // namespace cm
}  // namespace b2amd
