# Builds the reference's own PLINK + randomized-SVD path (PCAone v0.7.2) from
# its source files where they lie, into oracle/_ref/ (git-ignored).  Test
# infrastructure only: golden fixtures and the CPU baseline come from it.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
#
# Compiled: src/{Halko,Data,FilePlink,Utils,Cmd,kfunc}.cpp (the path and the
# helpers it links against) + the vendored zstd library sources (Utils.cpp
# links zstd stream functions) + oracle/ref_driver.cpp (our own main).
# Flags follow the reference Makefile (-O3 -std=c++17 -fopenmp -DNDEBUG,
# Eigen without external BLAS), except -march=x86-64-v3 instead of
# -march=native so the binary also runs on the GPU host's CPU.
REF      ?= /root/reference
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
CXX      := g++
CC       := gcc
ARCHFLAG ?= -march=x86-64-v3
CXXFLAGS := -O3 -std=c++17 -m64 -DNDEBUG $(ARCHFLAG) -fopenmp -DVERSION=\"0.7.2\" -w
INC      := -I$(REF)/src -I$(REF)/external -I$(REF)/external/zstd/lib -I$(REF)/external/pgenlib
CFLAGS   := -O3 -DZSTD_DISABLE_ASM -DZSTD_LEGACY_SUPPORT=0 $(ARCHFLAG) -w -I$(REF)/external/zstd/lib -I$(REF)/external/zstd/lib/common

SRCS     := Halko Data FilePlink Utils Cmd kfunc
CPPOBJ   := $(addprefix $(OBJ)/,$(addsuffix .o,$(SRCS)))
ZSTD_C   := $(wildcard $(REF)/external/zstd/lib/common/*.c) \
            $(wildcard $(REF)/external/zstd/lib/compress/*.c) \
            $(wildcard $(REF)/external/zstd/lib/decompress/*.c)
ZSTDOBJ  := $(patsubst $(REF)/external/zstd/lib/%.c,$(OBJ)/zstd/%.o,$(ZSTD_C))

all: $(OUT)/pcaone_ref

# The same reference objects with libpcaone_amd bound in through
# pcaone_amd/binding/ (the maintainer-side binding, compiled by the
# reference's own compiler and flags): make -f oracle/ref.mk binding
AMDINC   := -Iinclude -Ipcaone_amd/binding
AMDLIB   := -Lpcaone_amd -lpcaone_amd -Wl,-rpath,'$$ORIGIN/../../pcaone_amd' -Wl,-rpath,/opt/rocm/lib
binding: $(OUT)/pcaone_ref_amd

$(OBJ)/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OBJ)/ref_driver.o: oracle/ref_driver.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OBJ)/ref_driver_amd.o: oracle/ref_driver_amd.cpp pcaone_amd/binding/pcaone_reference_binding.hpp include/pcaone_amd.h
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) $(AMDINC) -c $< -o $@

$(OUT)/pcaone_ref_amd: $(CPPOBJ) $(OBJ)/ref_driver_amd.o $(ZSTDOBJ) pcaone_amd/libpcaone_amd.so
	$(CXX) $(CXXFLAGS) -o $@ $(CPPOBJ) $(OBJ)/ref_driver_amd.o $(ZSTDOBJ) $(AMDLIB) -lz -lgomp -lpthread -ldl -lm

$(OBJ)/zstd/%.o: $(REF)/external/zstd/lib/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/pcaone_ref: $(CPPOBJ) $(OBJ)/ref_driver.o $(ZSTDOBJ)
	$(CXX) $(CXXFLAGS) -o $@ $^ -lz -lgomp -lpthread -ldl -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean binding
