// PLUS reference vectors: the reference's own PLUS operator
// (CompOp_Plus_Op, pipeline/reference/compopgeneric_p.h:66-81, compiled
// here from the reference headers; its portable pipeline does not wire it
// in, fixedpiperuntime.cpp:254) applied to a grid that covers every
// (destination, source, mask) channel triple once:
//
//   for m in 0..255, pixel i in 0..16383, channel c in 0..3:
//     p = 4 i + c;  d_c = p >> 8;  s_c = p & 255
//
// (channels are independent in PLUS).  Writes the destination grid, the
// source grid and the results for every m, little-endian uint32, to the file
// named by argv[1] as m-major rows of 16384 results.
#include <blend2d/blend2d.h>
#include <blend2d/pipeline/reference/compopgeneric_p.h>

#include <cstdint>
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  using namespace bl::Pipeline::Reference;
  using Px = Pixel::P32_A8R8G8B8;
  constexpr uint32_t kPixels = 16384;
  std::vector<uint32_t> d(kPixels), s(kPixels), out((size_t)kPixels * 256);
  for (uint32_t i = 0; i < kPixels; i++) {
    uint32_t dv = 0, sv = 0;
    for (uint32_t c = 0; c < 4; c++) {
      const uint32_t p = 4 * i + c;
      dv |= (p >> 8) << (8 * c);
      sv |= (p & 255u) << (8 * c);
    }
    d[i] = dv;
    s[i] = sv;
  }
  for (uint32_t m = 0; m < 256; m++)
    for (uint32_t i = 0; i < kPixels; i++)
      out[(size_t)m * kPixels + i] = CompOp_Plus_Op<Px>::op_prgb32_prgb32(Px::from_value(d[i]), Px::from_value(s[i]), m).value();
  FILE* f = fopen(argv[1], "wb");
  if (!f) return 1;
  fwrite(out.data(), 4, out.size(), f);
  fclose(f);
  return 0;
}
