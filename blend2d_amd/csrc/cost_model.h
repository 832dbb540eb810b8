// Learned sizing: a small MLP that predicts, per fill command, how many 24.8
// edge segments the edge builder will emit and how many (segment, band)
// copies binning will make.  The predictions size the segment and bin pools
// so a frame can skip the counting pass of the edge builder; an overflow of
// either pool makes the frame fall back to exact counting, so pixels never
// depend on the model.
//
// Network: 16 features -> 64 (ReLU) -> 64 (ReLU) -> 2, fp32.  Targets and
// outputs are log2(1 + count).  Parameters are one flat array:
//   W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]   (row = input)
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kernels.h"

namespace b2amd {
namespace cm {

constexpr int kIn = 16, kH = 64, kOut = 2;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kIn * kH;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH * kOut;
constexpr int kParams = kB3 + kOut;          // 5378

// Per-command feature accumulators filled by one pass over the source
// segments (atomics), turned into the 16 features by the finalize kernel.
struct CmdAcc {
  uint32_t n_line, n_quad, n_cubic, n_vis;       // counts; n_vis: sources reaching the canvas rows
  uint32_t n_left, n_right, pad0, pad1;          // sources reaching left of x = 0 / right of x = w
  unsigned long long chord, poly, flat, flat_vis; // sums in 1/16 px: chord, control polygon, sqrt(max |2nd difference|)
  int32_t bx0, by0, bx1, by1;                    // control-point bbox, 24.8 (clamped to +-2^30)
};

// Device views for one frame.
struct FrameModel {
  const float* params;     // kParams
  CmdAcc* acc;             // n_cmds
  float* features;         // n_cmds x kIn
  float* pred;             // n_cmds x kOut (log2(1 + count))
  unsigned long long* totals;  // [0] predicted segments, [1] predicted bins (sums over commands)
  uint32_t* actual_segs;   // n_cmds, exact segments per command (logging)
  uint32_t* actual_bins;   // n_cmds, exact (segment, band) copies per command (logging)
};

void launch_features(const FrameDev& f, const FrameModel& m, hipStream_t st);
void launch_infer_frame(const FrameDev& f, const FrameModel& m, hipStream_t st);
void launch_actuals(const FrameDev& f, const FrameScratch& s, const FrameModel& m, uint32_t n_segs, hipStream_t st);

// Generic MLP entry points (training / tests).
void launch_mlp_forward(const float* X, uint32_t n, const float* params, float* Y, hipStream_t st);
// Gradient of the mean squared error over rows [0, n): per-workgroup partial
// sums into `partial` (n_groups(n) x (kParams + 1), last column = loss sum).
uint32_t mlp_groups(uint32_t n);
void launch_mlp_grad(const float* X, const float* T, uint32_t n, const float* params, float* partial, hipStream_t st);
// Reduces the partials (fixed order) into grad[kParams + 1] scaled by 1/n.
void launch_mlp_reduce(const float* partial, uint32_t groups, uint32_t n, float* grad, hipStream_t st);
void launch_adam(float* params, const float* grad, float* m1, float* m2, uint32_t step, float lr, hipStream_t st);

}  // namespace cm
}  // namespace b2amd
