# Builds the reference's two GPU solvers (glucose-syrup/gpu and
# rel-newtech/gpu, the callers of the check path) from their own sources
# where they lie under the reference
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
# the second solver tree (rel-newtech), same interface
RSRCS := $(wildcard $(REF)/rel-newtech/gpu/*.cc) \
         $(filter-out %/Main.cc,$(wildcard $(REF)/rel-newtech/core/*.cc)) \
         $(filter-out %/Main.cc,$(wildcard $(REF)/rel-newtech/simp/*.cc)) \
         $(wildcard $(REF)/rel-newtech/utils/*.cc) $(REF)/gpuShareLib/Utils.cc
ROBJS := $(patsubst $(REF)/%.cc,$(OUT)/robj/%.o,$(RSRCS))
# this repository's include/ first, so gpuShareLib/GpuClauseSharer.h is ours;
# the reference's host-only helpers (Utils.h, Profiler.h) come from its tree
CXXFLAGS := -std=c++14 -O2 -w -I$(REPO)/include -I$(REF) -I$(REF)/glucose-syrup \
            -D__STDC_LIMIT_MACROS -D__STDC_FORMAT_MACROS -DNDEBUG

all: $(OUT)/glucose-gpu $(OUT)/rel-newtech-gpu

$(OUT)/obj/%.o: $(REF)/%.cc
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/robj/%.o: $(REF)/%.cc
	@mkdir -p $(dir $@)
	$(CXX) $(subst glucose-syrup,rel-newtech,$(CXXFLAGS)) -c $< -o $@

$(OUT)/rel-newtech-gpu: $(ROBJS) $(REPO)/gpusharesat_amd/libgpusharesat.so
	$(CXX) -o $@ $(ROBJS) -L$(REPO)/gpusharesat_amd -lgpusharesat \
	    -Wl,-rpath,'$$ORIGIN/../../gpusharesat_amd' -lz -lpthread

$(OUT)/glucose-gpu: $(OBJS) $(REPO)/gpusharesat_amd/libgpusharesat.so
	$(CXX) -o $@ $(OBJS) -L$(REPO)/gpusharesat_amd -lgpusharesat \
	    -Wl,-rpath,'$$ORIGIN/../../gpusharesat_amd' -lz -lpthread

clean:
	rm -rf $(OUT)/obj $(OUT)/robj $(OUT)/glucose-gpu $(OUT)/rel-newtech-gpu

.PHONY: all clean
