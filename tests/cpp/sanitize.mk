# Host-code build under AddressSanitizer + UndefinedBehaviorSanitizer, for
# a machine without a GPU (tests/test_host_sanitizers.py):
#
#   make -f tests/cpp/sanitize.mk -j8
#
# libmdx is compiled for gfx950 as usual, with the sanitizers on its host
# side only (the driver reaches only paths that return before a launch);
# the mirror,
# the Benchmark-mode initial conditions, the Benchmark driver and the oracle
# are compiled with the same clang, so one sanitizer runtime serves them
# all.  Outputs go to tests/cpp/_build/san/.
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
ROOT := $(abspath $(HERE)../..)
OUT := $(ROOT)/tests/cpp/_build/san
ROCM ?= /opt/rocm
HIPCC := $(ROCM)/bin/hipcc
CXX := $(ROCM)/llvm/bin/clang++
CC := $(ROCM)/llvm/bin/clang
SAN := -fsanitize=address -fsanitize=undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer
HIP_SAN := -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
  -Xarch_host -fno-sanitize-recover=undefined -fno-omit-frame-pointer
COMMON := -O1 -g -fPIC -I$(ROOT)/include

all: $(OUT)/sanitize_driver $(OUT)/mdx_benchmark

$(OUT)/%.o: $(ROOT)/mdacp_amd/csrc/%.hip $(lastword $(MAKEFILE_LIST))
	@mkdir -p $(OUT)
	$(HIPCC) --offload-arch=gfx950 -std=c++17 $(COMMON) -ffp-contract=fast-honor-pragmas \
	  -Xclang -target-feature -Xclang -packed-fp32-ops $(HIP_SAN) -c $< -o $@

$(OUT)/confmaker.o: $(ROOT)/mdacp_amd/host/confmaker.cpp
	@mkdir -p $(OUT)
	$(CXX) -std=c++17 $(COMMON) -mavx2 -mfma -ffp-contract=fast $(SAN) -c $< -o $@

$(OUT)/mdacp_gpu.o: $(ROOT)/mdacp_amd/host/mdacp_gpu.cpp
	@mkdir -p $(OUT)
	$(CXX) -std=c++17 $(COMMON) $(SAN) -c $< -o $@

$(OUT)/mdoracle.o: $(ROOT)/oracle/mdoracle.c
	@mkdir -p $(OUT)
	$(CC) -std=c11 $(COMMON) -ffp-contract=off $(SAN) -c $< -o $@

LIB_OBJS := $(OUT)/mdx.o $(OUT)/ensemble.o $(OUT)/mdacp_gpu.o $(OUT)/confmaker.o
LINK := -L$(ROCM)/lib -lamdhip64 -Wl,-rpath,$(ROCM)/lib -lm

$(OUT)/sanitize_driver: $(ROOT)/tests/cpp/sanitize_driver.cc $(LIB_OBJS) $(OUT)/mdoracle.o
	$(CXX) -std=c++17 $(COMMON) $(SAN) -o $@ $^ $(LINK)

$(OUT)/mdx_benchmark: $(ROOT)/mdacp_amd/host/benchmark.cpp $(LIB_OBJS)
	$(CXX) -std=c++17 $(COMMON) $(SAN) -o $@ $^ $(LINK)

.PHONY: all
