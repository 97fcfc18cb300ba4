# Builds the reference MBS-fast CPU path from its own sources, in place, into
# oracle/_ref/.  This is test infrastructure only: the objects produced here are
# used to write golden fixtures and to time the reference CPU baseline.  Nothing
# in the product links against them.
#
# Usage (from the repository root, only where the reference tree is present):
#   make -f oracle/ref.mk -j8            # oracle/_ref/mbs_po + libmbsref.a
#
# The recipe compiles every C++ source of the reference except the CUDA units
# (src/cuda/*.cu and the CUDA runtime query src/cuda/GpuSupport.cpp); the CPU
# stub cpu/GpuSupportStub.cpp takes the place of the latter, exactly as the
# reference's own CPU build does.  No MPI, no CUDA.  AVX2+FMA only, so the
# binary also runs on any x86-64 host that has them.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
ARCH     ?= -mavx2 -mfma
CXXFLAGS := -O3 -funroll-loops $(ARCH) -std=gnu++11 -fopenmp -w \
            -DMBS_GIT_DESCRIBE=\"reference-snapshot\"
INC      := -I$(REF)/src -I$(REF)/src/math -I$(REF)/src/handler \
            -I$(REF)/src/common -I$(REF)/src/geometry \
            -I$(REF)/src/geometry/intrinsic -I$(REF)/src/particle \
            -I$(REF)/src/scattering -I$(REF)/src/tracer \
            -I$(REF)/src/cuda -I$(REF)/src/bigint

SRCS := $(shell find $(REF)/src -path $(REF)/src/cuda -prune -o -name '*.cpp' -print) \
        $(shell find $(REF)/src/bigint -name '*.cc') \
        $(REF)/cpu/GpuSupportStub.cpp
OBJS := $(patsubst $(REF)/%,$(OUT)/obj/%.o,$(SRCS))
MAIN_OBJ := $(OUT)/obj/src/main.cpp.o
LIB_OBJS := $(filter-out $(MAIN_OBJ),$(OBJS))

# The GPU-slot variant: HandlerPO.cpp and main.cpp are compiled with
# -DUSE_CUDA so that HandlerPO's GPU members and the GpuSupport queries are
# left undefined, and a shim translation unit supplies them.  The fixture
# writer tests/golden/ref_dump_shim.cpp is such a shim.
SLOT_SRCS := $(REF)/src/handler/HandlerPO.cpp $(REF)/src/main.cpp
SLOT_OBJS := $(patsubst $(REF)/%,$(OUT)/objslot/%.o,$(SLOT_SRCS))
BASE_OBJS := $(filter-out $(patsubst $(REF)/%,$(OUT)/obj/%.o,$(SLOT_SRCS) $(REF)/cpu/GpuSupportStub.cpp),$(OBJS))

all: $(OUT)/mbs_po $(OUT)/libmbsref.a $(OUT)/mbs_po_dump $(OUT)/ref_fixed_harness \
     $(OUT)/mbs_po_hip $(OUT)/ref_sobol_dump $(OUT)/dropin_harness

# The reference's Sobol2D points, for the orientation-set fixture.
$(OUT)/ref_sobol_dump: tests/golden/ref_sobol_dump.cpp $(OUT)/obj/src/math/Sobol.cpp.o
	$(CXX) $(CXXFLAGS) $(INC) $^ -o $@

# The reference CLI with the MI355X backend: the GPU-slot objects plus the
# product's drop-in translation unit mbs-fast_amd/host/HandlerPOHip.cpp,
# linked against the engine library (found at run time through the rpath).
ENGINE_SO := mbs-fast_amd/libmbs_hip.so
$(OUT)/hip_shim.o: mbs-fast_amd/host/HandlerPOHip.cpp include/mbs_hip.h
	$(CXX) $(CXXFLAGS) $(INC) -Iinclude -c $< -o $@

$(OUT)/mbs_po_hip: $(BASE_OBJS) $(SLOT_OBJS) $(OUT)/hip_shim.o $(ENGINE_SO)
	$(CXX) $(CXXFLAGS) $(BASE_OBJS) $(SLOT_OBJS) $(OUT)/hip_shim.o -o $@ \
	    -L$(dir $(ENGINE_SO)) -lmbs_hip -Wl,-rpath,'$$ORIGIN/../../mbs-fast_amd' -lm -lgomp

# Harness of the drop-in members the CLI does not reach on one GPU (the
# pack-cache token, the theta probe): the reference library, the GPU-slot
# HandlerPO and the product's drop-in over the engine library.
$(OUT)/dropin_harness: tests/dropin_harness.cpp $(filter-out $(OUT)/obj/src/main.cpp.o,$(BASE_OBJS)) $(OUT)/objslot/src/handler/HandlerPO.cpp.o $(OUT)/hip_shim.o $(ENGINE_SO)
	$(CXX) $(CXXFLAGS) $(INC) tests/dropin_harness.cpp $(filter-out $(OUT)/obj/src/main.cpp.o,$(BASE_OBJS)) \
	    $(OUT)/objslot/src/handler/HandlerPO.cpp.o $(OUT)/hip_shim.o -o $@ \
	    -L$(dir $(ENGINE_SO)) -lmbs_hip -Wl,-rpath,'$$ORIGIN/../../mbs-fast_amd' -lm -lgomp

# Fixed-orientation fixture harness: the reference library plus the dump shim.
$(OUT)/ref_fixed_harness: tests/golden/ref_fixed_harness.cpp $(filter-out $(OUT)/obj/src/main.cpp.o,$(BASE_OBJS)) $(OUT)/objslot/src/handler/HandlerPO.cpp.o $(OUT)/dump_shim.o
	$(CXX) $(CXXFLAGS) $(INC) $^ -o $@ -lm -lgomp

$(OUT)/objslot/%.o: $(REF)/%
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DUSE_CUDA $(INC) -c $< -o $@

$(OUT)/dump_shim.o: tests/golden/ref_dump_shim.cpp
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/mbs_po_dump: $(BASE_OBJS) $(SLOT_OBJS) $(OUT)/dump_shim.o
	$(CXX) $(CXXFLAGS) $^ -o $@ -lm -lgomp

$(OUT)/obj/%.o: $(REF)/%
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/libmbsref.a: $(LIB_OBJS)
	rm -f $@
	ar rcs $@ $^

$(OUT)/mbs_po: $(OBJS)
	$(CXX) $(CXXFLAGS) $^ -o $@ -lm -lgomp

.PHONY: all
