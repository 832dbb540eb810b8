// Host-side style setup of the fill path: matrices, premultiplication,
// gradient LUTs, fetch constants and comp-op simplification.  All double
// precision work of the reference frontend stays on the host, in the same
// operation order, so the constants the kernels receive are bit-identical.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "device_types.h"
#include "lut_device.h"

namespace b2amd {

struct Mat { double m00, m01, m10, m11, m20, m21; };

enum TransformType { kTTIdentity = 0, kTTTranslate = 1, kTTScale = 2, kTTSwap = 3, kTTAffine = 4, kTTInvalid = 5 };

// FormatExt values used by comp-op simplification (core/format_p.h).
enum FmtExt : uint32_t { kFxNone = 0, kFxPRGB32 = 1, kFxXRGB32 = 2, kFxA8 = 3, kFxFRGB32 = 4, kFxZERO32 = 5 };

// BLCompOp values (core/context.h) and the internal ops simplification yields.
enum CompOp : uint32_t {
  kCompSrcOver = 0, kCompSrcCopy = 1, kCompSrcIn = 2, kCompDstOver = 5, kCompDstCopy = 6, kCompDstIn = 7,
  kCompClear = 11, kCompPlus = 12, kCompModulate = 14, kCompMultiply = 15, kCompScreen = 16,
  kCompNop = 255
};

enum SolidId : uint32_t { kSolidNone = 0, kSolidTransparent = 1, kSolidOpaqueBlack = 2, kSolidOpaqueWhite = 3 };

struct Simplified { uint32_t op, dst, src, solid; };

int transform_type(const Mat& m);
bool invert(Mat& dst, const Mat& s);
Mat multiply(const Mat& a, const Mat& b);

int nearby_to_int(double x);
int round_to_int(double x);
int64_t floor_to_int64(double x);

uint32_t premultiply(uint32_t argb32);

using Stop = lut::Stop;
// add_stop semantics of BLGradient (sorted; a third stop at one offset replaces the second).
void add_stop(std::vector<Stop>& stops, double offset, uint32_t rgba32);
uint32_t gradient_info(const std::vector<Stop>& stops, bool& solid, uint32_t& format);

bool init_linear(DevStyle& s, uint32_t lut_size, const double v[4], uint32_t extend, const Mat& t);
bool init_radial(DevStyle& s, uint32_t lut_size, const double v[6], uint32_t extend, const Mat& t);
bool init_conic(DevStyle& s, uint32_t lut_size, const double v[4], const Mat& t);
// Returns 1 on success, 0 when the transform is not invertible or the image
// is empty (the fill is a no-op).
int init_pattern(DevStyle& s, uint32_t img_w, uint32_t img_h, uint32_t extend, uint32_t quality, const Mat& t);

Simplified simplify(uint32_t op, uint32_t dst_fmt, uint32_t src_fmt);

}  // namespace b2amd
