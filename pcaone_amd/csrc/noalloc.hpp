// The packed product with every intermediate in a fixed device scratch array
// (no device allocation); rows, W and Y are device pointers.
#pragma once
#include <cstdint>

#include "gemm.hpp"

namespace pcaone {

class Context;

constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;  // per GPU, held until the process ends

void packed_gemm_noalloc(Context& x, Form f, const uint8_t* rows, uint64_t nrows, uint32_t ncodes,
                         const double* W, uint32_t L, double* Y);

}  // namespace pcaone
