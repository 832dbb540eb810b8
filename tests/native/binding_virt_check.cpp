// Every slot of the GPU backend's BLContextVirt table is set
// (blend2d_amd/binding/amdcontext.cpp): the table is a struct of function
// pointers, so a slot left null would crash the first call that reaches it.
#include <stdio.h>

#include "../../blend2d_amd/binding/amdcontext.cpp"

int main() {
  const BLContextVirt& v = bl_amd_binding::amd_virt;
  const void* const* slot = reinterpret_cast<const void* const*>(&v);
  const size_t n = sizeof(v) / sizeof(void*);
  size_t null_slots = 0;
  for (size_t i = 0; i < n; i++)
    if (!slot[i]) null_slots++;
  printf("slots %zu null %zu\n", n, null_slots);
  return null_slots ? 1 : 0;
}
