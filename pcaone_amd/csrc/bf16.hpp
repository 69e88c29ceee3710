// BF16 / FP16 / FP8 rounding helpers shared by the operand-split kernels.
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

// Round-to-nearest-even FP64 -> FP8 E4M3 (the OCP format: exponent bias 7,
// largest finite 448 = 0x7E, subnormal step 2^-9) for |y| <= 448; *val
// receives the value of the returned bits.  Each binade [2^E, 2^(E+1)) holds
// 8 steps of 2^(E-3); below 2^-6 the step stays 2^-9.
__device__ __forceinline__ uint8_t e4m3_bits(double y, double* val) {
  const double a = fabs(y);
  if (!(a > 0.0)) {
    *val = 0.0;
    return 0;
  }
  int e;
  (void)frexp(a, &e);  // a = f 2^e, f in [0.5, 1): leading bit 2^(e-1)
  const int E = e - 1 < -6 ? -6 : e - 1;
  double r = rint(ldexp(a, 3 - E));  // steps of 2^(E-3), ties to even (exact scaling)
  r = ldexp(r, E - 3);
  if (r > 448.0) r = 448.0;
  *val = y < 0.0 ? -r : r;
  if (r == 0.0) return 0;
  int er;
  (void)frexp(r, &er);
  const int Er = er - 1;
  uint8_t bits;
  if (Er < -6)
    bits = (uint8_t)ldexp(r, 9);  // subnormal: r / 2^-9 in 1..7
  else
    bits = (uint8_t)(((Er + 7) << 3) | ((int)ldexp(r, 3 - Er) - 8));
  return (uint8_t)(bits | (y < 0.0 ? 0x80 : 0));
}

}  // namespace pcaone
