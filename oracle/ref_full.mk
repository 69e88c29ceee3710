# Builds the reference's complete program (PCAone v0.7.2: src/Main.cpp and
# every subsystem it links) from its source files where they lie, into
# oracle/_ref/pcaone_full (git-ignored).  Test infrastructure only: golden
# fixtures of the --svd 0 (IRAM) and --svd 3 (exact PCA) branches, which live
# in src/Arnoldi.cpp and src/Main.cpp:180-217.
#
#   make -f oracle/ref_full.mk REF=/root/reference -j8
#
# Compiled: every src/*.cpp, the vendored bgen, pgenlib (reader) and zstd
# sources; flags as oracle/ref.mk (the reference Makefile's, Eigen without
# external BLAS/LAPACK, -march=x86-64-v3).
REF      ?= /root/reference
OUT      := oracle/_ref
OBJ      := $(OUT)/full_obj
CXX      := g++
CC       := gcc
ARCHFLAG ?= -march=x86-64-v3
CXXFLAGS := -O3 -std=c++17 -m64 -DNDEBUG $(ARCHFLAG) -fopenmp -DVERSION=\"0.7.2\" -w
INC      := -I$(REF)/src -I$(REF)/external -I$(REF)/external/zstd/lib -I$(REF)/external/pgenlib \
            -I$(REF)/external/pgenlib/include -I$(REF)/external/bgen
CFLAGS   := -O3 -DZSTD_DISABLE_ASM -DZSTD_LEGACY_SUPPORT=0 $(ARCHFLAG) -w -I$(REF)/external/zstd/lib -I$(REF)/external/zstd/lib/common

SRC_CPP  := $(wildcard $(REF)/src/*.cpp)
BGEN_CPP := $(wildcard $(REF)/external/bgen/*.cpp)
PGEN_CC  := $(wildcard $(REF)/external/pgenlib/include/*.cc) $(wildcard $(REF)/external/pgenlib/*.cc)
PGEN_CPP := $(wildcard $(REF)/external/pgenlib/*.cpp)
ZSTD_C   := $(wildcard $(REF)/external/zstd/lib/common/*.c) \
            $(wildcard $(REF)/external/zstd/lib/compress/*.c) \
            $(wildcard $(REF)/external/zstd/lib/decompress/*.c)
OBJS     := $(patsubst $(REF)/%.cpp,$(OBJ)/%.o,$(SRC_CPP) $(BGEN_CPP) $(PGEN_CPP)) \
            $(patsubst $(REF)/%.cc,$(OBJ)/%.o,$(PGEN_CC)) \
            $(patsubst $(REF)/%.c,$(OBJ)/%.o,$(ZSTD_C))

all: $(OUT)/pcaone_full

$(OBJ)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OBJ)/%.o: $(REF)/%.cc
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OBJ)/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/pcaone_full: $(OBJS)
	$(CXX) $(CXXFLAGS) -o $@ $^ -lz -lgomp -lpthread -ldl -lm

.PHONY: all
