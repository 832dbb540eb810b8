// Checks the per-entry gradient tables of blend2d_amd/csrc/lut_device.h
// (compiled here for the host), 32-bit and 64-bit (dithered gradients),
// against the oracle's span-by-span tables (orc_gradient_lut /
// orc_gradient_lut64, oracle/orc_scene.c, pinned by the reference-rendered
// gradient goldens) for random stop lists: sorted offsets drawn from a mix of
// exact positions (0, 0.002, 0.5, 0.998, 1), clusters and uniform values,
// pairs of equal offsets, 1-12 stops and up to 64 (the oracle's capacity).
//
//   lut_check <cases> <seed>   prints "cases N entries M mismatches K"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <vector>

#include "../../blend2d_amd/csrc/lut_device.h"

extern "C" void orc_gradient_lut(const double* offsets, const uint32_t* rgba32, size_t n, uint32_t* out,
                                 uint32_t* lut_size_out);
extern "C" void orc_gradient_lut64(const double* offsets, const uint32_t* rgba32, size_t n, uint64_t* out,
                                   uint32_t* lut_size_out);

int main(int argc, char** argv) {
  const int cases = argc > 1 ? atoi(argv[1]) : 20000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 1);
  auto uni = [&]() { return (double)(rng() >> 11) * (1.0 / 9007199254740992.0); };
  const double exact[] = {0.0, 0.002, 0.5, 0.998, 1.0, 0.25, 0.75};
  long entries = 0, bad = 0;
  std::vector<uint32_t> out(1024);
  std::vector<uint64_t> out64(1024);
  for (int c = 0; c < cases; c++) {
    size_t n = 1 + rng() % 12;
    if (c % 97 == 0) n = 40 + rng() % 25;   // the oracle keeps up to 64 stops
    std::vector<double> off(n);
    const double lo = uni() * 0.5, span = uni();
    for (size_t i = 0; i < n; i++) {
      const uint64_t kind = rng() % 8;
      if (kind == 0) off[i] = exact[rng() % 7];
      else if (kind == 1) off[i] = std::min(1.0, lo + span * 0.01 * uni());
      else off[i] = uni();
    }
    std::sort(off.begin(), off.end());
    // at most two stops at one offset (add_stop replaces a third)
    for (size_t i = 2; i < n; i++)
      if (off[i] == off[i - 2]) off[i] = std::min(1.0, off[i] + 1e-3);
    std::sort(off.begin(), off.end());
    for (size_t i = 2; i < n; i++)
      if (off[i] == off[i - 2]) { n = i; break; }
    off.resize(n);
    std::vector<uint32_t> col(n);
    for (size_t i = 0; i < n; i++) {
      uint32_t v = (uint32_t)rng();
      const uint64_t a = rng() % 4;
      if (a == 0) v |= 0xFF000000u;
      else if (a == 1) v &= 0x00FFFFFFu;
      col[i] = v;
    }
    uint32_t size = 0;
    orc_gradient_lut(off.data(), col.data(), n, out.data(), &size);
    std::vector<b2amd::lut::Stop> st(n);
    for (size_t i = 0; i < n; i++) {
      uint64_t v = 0;
      for (int k = 0; k < 4; k++) v |= ((uint64_t)((col[i] >> (8 * k)) & 0xFFu) * 0x101u) << (16 * k);
      st[i] = b2amd::lut::Stop{off[i], v};
    }
    for (uint32_t e = 0; e < size; e++) {
      entries++;
      const uint32_t got = b2amd::lut::entry(st.data(), (uint32_t)n, size, e);
      if (got != out[e]) {
        if (bad < 5) printf("case %d n %zu size %u entry %u: got %08x want %08x\n", c, n, size, e, got, out[e]);
        bad++;
      }
    }
    // the 64-bit table of the same stops (dithered gradients)
    uint32_t size64 = 0;
    orc_gradient_lut64(off.data(), col.data(), n, out64.data(), &size64);
    for (uint32_t e = 0; e < size64; e++) {
      entries++;
      const uint64_t got = b2amd::lut::entry64(st.data(), (uint32_t)n, size64, e);
      if (got != out64[e]) {
        if (bad < 5)
          printf("case %d n %zu size64 %u entry %u: got %016llx want %016llx\n", c, n, size64, e,
                 (unsigned long long)got, (unsigned long long)out64[e]);
        bad++;
      }
    }
  }
  printf("cases %d entries %ld mismatches %ld\n", cases, entries, bad);
  return bad ? 1 : 0;
}
