# Builds the reference's Glucose GPU solver (glucose-syrup/gpu, a caller of
# the check path) from its own sources where they lie under the reference
# tree, against this repository's include/gpuShareLib/GpuClauseSharer.h and
# libgpusharesat.so instead of the reference's CUDA gpuShareLib: the
# integration check that the reference's solver threads link and run
# unchanged.  Test infrastructure; output only into oracle/_ref/.
#
#   make -C oracle -f solver.mk          (after python -m gpusharesat_amd.build)
REF ?= /root/reference
REPO := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(REPO)/oracle/_ref
SRCS := $(wildcard $(REF)/glucose-syrup/gpu/*.cc) $(wildcard $(REF)/glucose-syrup/core/*.cc) \
        $(filter-out %/Main.cc,$(wildcard $(REF)/glucose-syrup/simp/*.cc)) \
        $(wildcard $(REF)/glucose-syrup/utils/*.cc) \
        $(wildcard $(REF)/glucose-syrup/satUtils/*.cc) $(REF)/gpuShareLib/Utils.cc
OBJS := $(patsubst $(REF)/%.cc,$(OUT)/obj/%.o,$(SRCS))
# this repository's include/ first, so gpuShareLib/GpuClauseSharer.h is ours;
# the reference's host-only helpers (Utils.h, Profiler.h) come from its tree
CXXFLAGS := -std=c++14 -O2 -w -I$(REPO)/include -I$(REF) -I$(REF)/glucose-syrup \
            -D__STDC_LIMIT_MACROS -D__STDC_FORMAT_MACROS -DNDEBUG

all: $(OUT)/glucose-gpu

$(OUT)/obj/%.o: $(REF)/%.cc
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/glucose-gpu: $(OBJS) $(REPO)/gpusharesat_amd/libgpusharesat.so
	$(CXX) -o $@ $(OBJS) -L$(REPO)/gpusharesat_amd -lgpusharesat \
	    -Wl,-rpath,'$$ORIGIN/../../gpusharesat_amd' -lz -lpthread

clean:
	rm -rf $(OUT)/obj $(OUT)/glucose-gpu

.PHONY: all clean
