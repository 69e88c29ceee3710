// BF16 / FP16 rounding helpers shared by the operand-split kernels.
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

// Round-to-nearest-even FP64 -> FP16 bits, and back.
__device__ __forceinline__ uint16_t fp16_bits(double x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ __forceinline__ double fp16_value(uint16_t b) { return (double)__builtin_bit_cast(_Float16, b); }

}  // namespace pcaone
