# Builds the CPU oracle (plain-C restatement of the reference fill path) into
# oracle/_build/liborc.so.  Test infrastructure only.
#
#   make -f oracle/orc.mk            (from the repository root)
#
# -ffp-contract=off keeps every multiply-add unfused, as the reference's
# Math::madd is (support/math_p.h:134-135); no -mfma.

OUT    := oracle/_build
CC     ?= gcc
CFLAGS := -std=c11 -O2 -fPIC -ffp-contract=off -fno-math-errno -Wall -Wno-unused-label -Iblend2d_amd/csrc
SRCS   := oracle/orc_edges.c oracle/orc_raster.c oracle/orc_pipe.c oracle/orc_style.c oracle/orc_scene.c
HDRS   := oracle/orc.h oracle/orc_core.h blend2d_amd/csrc/scene_format.h

all: $(OUT)/liborc.so

$(OUT)/liborc.so: $(SRCS) $(HDRS)
	@mkdir -p $(OUT)
	$(CC) $(CFLAGS) -shared -o $@ $(SRCS) -lm

.PHONY: all
