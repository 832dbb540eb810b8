# Builds the reference Blend2D library (portable, non-JIT pipeline) from its
# sources under /root/reference into oracle/_ref/, plus the scene driver
# oracle/ref_render.cpp linked against it.  This is test infrastructure: the
# result is only ever used as a checker (tests/) and as the CPU baseline
# (bench.py), never by the product library.
#
#   make -f oracle/ref.mk -j8            (from the repository root)
#
# Flags follow the reference's own release configuration for GCC on x86_64
# (CMakeLists.txt:650-733): SSE2 baseline, per-file SIMD flags for the
# *_sse4_2 / *_avx / *_avx2 sources, -DBL_BUILD_NO_JIT because asmjit is not
# vendored.  No header, library or generated file is substituted.

REF      ?= /root/reference
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
CXX      ?= g++

SRCS := $(shell cd $(REF) && find blend2d -name '*.cpp' ! -name '*_test*.cpp' ! -name '*_asimd*.cpp' | sort)
OBJS := $(patsubst %.cpp,$(OBJ)/%.o,$(SRCS))

BASEFLAGS := -std=c++17 -O2 -fPIC -fvisibility=hidden -I$(REF) \
  -fno-math-errno -fno-threadsafe-statics -fno-exceptions -fno-rtti \
  -fmerge-all-constants -fno-semantic-interposition -fno-trapping-math \
  -fno-finite-math-only -ftree-vectorize -msse2 -DBL_BUILD_OPT_AVX512 \
  -DBL_BUILD_NO_JIT -w

flags_for = $(strip \
  $(if $(findstring _sse3.cpp,$1),-msse3) \
  $(if $(findstring _ssse3.cpp,$1),-mssse3) \
  $(if $(findstring _sse4_1.cpp,$1),-msse4.1) \
  $(if $(findstring _sse4_2.cpp,$1),-mpopcnt -mpclmul -msse4.2) \
  $(if $(findstring _avx.cpp,$1),-mpopcnt -mpclmul -mavx) \
  $(if $(findstring _avx2.cpp,$1),-mpopcnt -mpclmul -mbmi -mbmi2 -mavx2 -DBL_TARGET_OPT_POPCNT -DBL_TARGET_OPT_BMI2) \
  $(if $(findstring _avx512.cpp,$1),-mpopcnt -mpclmul -mbmi -mbmi2 -mavx512f -mavx512bw -mavx512dq -mavx512cd -mavx512vl -DBL_TARGET_OPT_POPCNT -DBL_TARGET_OPT_BMI2))

all: $(OUT)/libblend2d_ref.so $(OUT)/libref_render.so $(OUT)/libamd_ref_render.so

$(OBJ)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(BASEFLAGS) $(call flags_for,$<) -c $< -o $@

$(OUT)/libblend2d_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -lpthread

$(OUT)/libref_render.so: oracle/ref_render.cpp blend2d_amd/csrc/scene_format.h $(OUT)/libblend2d_ref.so
	$(CXX) -std=c++17 -O2 -fPIC -shared -I$(REF) -Iblend2d_amd/csrc -o $@ oracle/ref_render.cpp \
	  -L$(OUT) -lblend2d_ref -Wl,-rpath,'$$ORIGIN'

.PHONY: all

# The same scene driver bound to the MI355X backend through its BLContext
# implementation (blend2d_amd/binding/amdcontext.cpp, compiled against the
# reference headers): the integration test of that binding.
amd: $(OUT)/libamd_ref_render.so

$(OUT)/libamd_ref_render.so: oracle/ref_render.cpp blend2d_amd/binding/amdcontext.cpp blend2d_amd/csrc/scene_format.h \
                             include/blend2d_amd.h $(OUT)/libblend2d_ref.so blend2d_amd/libblend2d_amd.so
	$(CXX) -std=c++17 -O2 -fPIC -shared -DREF_RENDER_AMD -DBL_BUILD_NO_JIT -I$(REF) -Iblend2d_amd/csrc -Iinclude -o $@ \
	  oracle/ref_render.cpp blend2d_amd/binding/amdcontext.cpp \
	  -L$(OUT) -lblend2d_ref -Lblend2d_amd -lblend2d_amd -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../blend2d_amd'

.PHONY: amd
