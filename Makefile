# MultiBlockMomentsFV on MI355X -- build
#
#   make            multiblockmomentsfv_amd/libmbmfv.so (engine + host C),
#                   build/moments (drop-in for the reference's ./build/moments)
#                   and multiblockmomentsfv_amd/libmbmprobe.so (FP64 issue-rate
#                   probe for bench.py --valu-rate)
#   make oracle     the CPU oracle used by the tests (oracle/liboracle.so)
#   make physcheck  host build of the engine's per-cell functions and their
#                   VJPs for the adjoint tests (tests/native/libphyscheck.so)
#   make clean
#
# All HIP code is compiled for gfx950 only.

HIPCC   ?= /opt/rocm/bin/hipcc
CC      ?= gcc
ARCH    ?= gfx950
PKG     := multiblockmomentsfv_amd
GIT_COMMIT := $(shell git rev-parse --short HEAD 2>/dev/null || echo unknown)

# -ffp-contract=on: a*b+c is fused only within one source expression, so every
# inlined copy of a formula rounds the same way and results do not depend on
# the tiling (the default 'fast' lets the backend fuse per call site).
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -Wall -Iinclude -ffp-contract=on \
            -Xclang -target-feature -Xclang -packed-fp32-ops \
            -Wno-unused-command-line-argument
CFLAGS   := -O2 -fPIC -Wall -Iinclude -DMBM_GIT_COMMIT=\"$(GIT_COMMIT)\"

LIB     := $(PKG)/libmbmfv.so
PROBE   := $(PKG)/libmbmprobe.so
BIN     := build/moments
OBJDIR  := $(PKG)/_obj

HIPSRC  := $(PKG)/csrc/mbm_engine.hip $(PKG)/csrc/mbm_adjoint.hip $(PKG)/csrc/mbm_batch.hip
HIPHDR  := $(wildcard $(PKG)/csrc/*.h) include/mbmfv.h include/mbmfv_host.h

CXX     ?= g++
PHYSCHK := tests/native/libphyscheck.so

.PHONY: all oracle physcheck clean
all: $(LIB) $(BIN) $(PROBE)

$(PROBE): $(PKG)/csrc/probe/fp64_rate.hip $(PKG)/csrc/mbm_phys.h
	$(HIPCC) $(HIPFLAGS) -shared -o $@ $<

physcheck: $(PHYSCHK)

$(PHYSCHK): tests/native/physcheck.cpp $(HIPHDR)
	$(CXX) -std=c++17 -O2 -fPIC -shared -ffp-contract=off -Iinclude -o $@ $<

$(OBJDIR):
	mkdir -p $(OBJDIR)

$(OBJDIR)/%.o: $(PKG)/csrc/%.hip $(HIPHDR) | $(OBJDIR)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(OBJDIR)/mbm_host.o: $(PKG)/host/mbm_host.c $(HIPHDR) | $(OBJDIR)
	$(CC) $(CFLAGS) -c $< -o $@

$(LIB): $(OBJDIR)/mbm_engine.o $(OBJDIR)/mbm_adjoint.o $(OBJDIR)/mbm_batch.o $(OBJDIR)/mbm_host.o
	$(HIPCC) --offload-arch=$(ARCH) -shared -o $@ $^

$(BIN): $(PKG)/host/moments.c $(LIB) include/mbmfv.h include/mbmfv_host.h
	mkdir -p build
	$(CC) $(CFLAGS) -o $@ $< -L$(PKG) -lmbmfv -Wl,-rpath,'$$ORIGIN/../$(PKG)'

oracle:
	$(MAKE) -C oracle

clean:
	rm -rf $(OBJDIR) $(LIB) $(PROBE) build $(PHYSCHK)
