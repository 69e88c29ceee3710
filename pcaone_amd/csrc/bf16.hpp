// BF16 rounding helpers shared by the operand-split kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pcaone {

// Round-to-nearest-even FP32 -> BF16 bits (inputs are finite).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

}  // namespace pcaone
