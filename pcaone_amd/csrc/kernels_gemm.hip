// Packed-genotype GEMM on CDNA4 matrix cores (see gemm.hpp for the contract).
//
// Workgroup = WAVES (4 or 8) waves; wave w owns WM row blocks of 16 rows and all NC column
// blocks of 16.  Per K-step of KS = 256 codes (128 for FP64) the workgroup
// stages the W tile (pieces x LP x KS, transposed) in LDS; every lane loads
// its KS/4 codes of each of its rows (16 bytes) straight from HBM and decodes
// them in registers.  Lane group q = lane/16 owns codes [q KS/4, (q+1) KS/4)
// of the step, so one load feeds all MFMAs of the step; the B fragment of the same K
// slots is a contiguous run of the transposed W row.  W rows are stored
// unpadded with their 16-byte vectors XOR-swizzled by the row index, so the
// reads are conflict-free and a double-buffered BF16 hi/lo tile at L = 80
// fits the 160 KB of LDS.
//
// Pipelining: the next step's W tile and genotype words are loaded into
// registers while the current step computes; the W tile is double-buffered
// in LDS when two copies fit (one barrier per step), single-buffered
// otherwise (FP64 at large LP; two barriers per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "device.hpp"
#include "gemm.hpp"

namespace pcaone {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

constexpr size_t kLdsMax = 160 * 1024;

// code -> top byte of v in BF16/FP32/FP64: 00 -> +2 (0x40), 01 -> 0, 10 -> 0, 11 -> -2 (0xC0)
constexpr uint32_t kLutTopByte = 0xC0000040u;

// code -> int8 v/2: 00 -> +1, 01 -> 0, 10 -> 0, 11 -> -1
constexpr uint32_t kLutI8 = 0xFF000001u;

__device__ __forceinline__ uint32_t spread(uint32_t w, int t) { return (w >> (2 * t)) & 0x03030303u; }
__device__ __forceinline__ uint32_t top_bytes(uint32_t codes4) { return __builtin_amdgcn_perm(0u, kLutTopByte, codes4); }

// Sixteen int8 values v/2 (codes 0..15 of word w, in order).
__device__ __forceinline__ i32x4_t decode_i8x16(uint32_t w) {
  i32x4_t d;
#pragma unroll
  for (int t = 0; t < 4; ++t) d[t] = (int)__builtin_amdgcn_perm(0u, kLutI8, spread(w, t));
  return d;
}

// Eight BF16 values (codes 8*half .. 8*half+7 of word w).
__device__ __forceinline__ bf16x8_t decode_bf16x8(uint32_t w, int half) {
  const uint32_t h0 = top_bytes(spread(w, 2 * half));
  const uint32_t h1 = top_bytes(spread(w, 2 * half + 1));
  u32x4_t d;
  d[0] = __builtin_amdgcn_perm(h0, h0, 0x050C040Cu);
  d[1] = __builtin_amdgcn_perm(h0, h0, 0x070C060Cu);
  d[2] = __builtin_amdgcn_perm(h1, h1, 0x050C040Cu);
  d[3] = __builtin_amdgcn_perm(h1, h1, 0x070C060Cu);
  return __builtin_bit_cast(bf16x8_t, d);
}

// Top 32 bits of the FP32/FP64 value of byte e of `tb` (other bytes zero).
__device__ __forceinline__ uint32_t top_word(uint32_t tb, int e) {
  return __builtin_amdgcn_perm(tb, tb, ((uint32_t)(4 + e) << 24) | 0x0C0C0Cu);
}

template <Form F>
struct FT;
// P: operand pieces; KS: codes per K-step; NACC: accumulator sets (the I8
// pieces carry different weights, 256 and 1, so each has its own).
template <>
struct FT<Form::BF16> {
  using w_t = uint16_t;
  using acc_t = float;
  static constexpr int P = 2, KS = 256, NACC = 1;
};
template <>
struct FT<Form::BF16X1> {
  using w_t = uint16_t;
  using acc_t = float;
  static constexpr int P = 1, KS = 256, NACC = 1;
};
template <>
struct FT<Form::FP16> {
  using w_t = uint16_t;
  using acc_t = float;
  static constexpr int P = 1, KS = 256, NACC = 1;
};
template <>
struct FT<Form::I8> {
  using w_t = uint8_t;
  using acc_t = int32_t;
  static constexpr int P = 2, KS = 256, NACC = 2;
};
template <>
struct FT<Form::F8> {
  using w_t = uint8_t;
  using acc_t = float;
  // (a 128-code step with two 4-wave workgroups per CU was measured slower:
  // 8.66 against 8.21 ms per launch at 50,000 x 1,000,000, L = 80)
  static constexpr int P = 3, KS = 256, NACC = 1;
};
template <>
struct FT<Form::FP32> {
  using w_t = float;
  using acc_t = float;
  static constexpr int P = 1, KS = 256, NACC = 1;
};
template <>
struct FT<Form::FP64> {
  using w_t = double;
  using acc_t = double;
  static constexpr int P = 1, KS = 128, NACC = 1;  // a 256-code FP64 tile would not fit in LDS at L = 128
};

template <Form F, int NC>
constexpr size_t tile_bytes() {
  return (size_t)FT<F>::P * NC * 16 * FT<F>::KS * sizeof(typename FT<F>::w_t);
}
template <Form F, int NC>
constexpr bool double_buffered() {
  return 2 * tile_bytes<F, NC>() <= kLdsMax;
}
template <Form F, int NC>
constexpr size_t lds_bytes() {
  return (double_buffered<F, NC>() ? 2 : 1) * tile_bytes<F, NC>();
}

template <Form F>
using accv_t = typename std::conditional<F == Form::FP64, f64x4_t,
                                         typename std::conditional<F == Form::I8, i32x4_t, f32x4_t>::type>::type;

// Code words per lane and row block for one K-step (each lane group of 16
// lanes owns a contiguous quarter of the step).
template <Form F>
using aword_t = typename std::conditional<FT<F>::KS == 256, u32x4_t, u32x2_t>::type;

// Position of the 16-byte vector v of LDS W-tile row `row` (an XOR within
// each aligned group of 8 or 16 vectors).  For the 2-byte forms the B reads
// (16 rows at vectors 8q + s, rows 512 bytes apart) are free of bank
// conflicts with the 8-row period and take two passes per read with the
// 16-row one (measured with SQ_LDS_BANK_CONFLICT on MI355X); the other forms
// keep the 16-row period.
template <Form F>
__device__ __forceinline__ int swz(int row, int v) {
  constexpr int VPR = FT<F>::KS * (int)sizeof(typename FT<F>::w_t) / 16;  // vectors per W-tile row
  return v ^ (row & (sizeof(typename FT<F>::w_t) == 2 ? 7 : (VPR < 16 ? VPR - 1 : 15)));
}

// FP8 form, 16x16x128 MFMAs with E4M3 on both sides: substep m takes words 2m
// and 2m + 1 of the lane group's KS / 64 (codes (KS / 4) q + 32 m + [0, 32), in
// order), decoded to E4M3 (+2 = 0x40, 0, -2 = 0xC0: the same bytes as the BF16
// top bytes); the B fragment of a piece is the same 32 K slots of its W row,
// two 16-byte vectors.  The pieces accumulate into one FP32 set.  The decode
// and the MFMAs are separate so that the main loop can issue the next step's
// loads between them.
template <int WM>
__device__ __forceinline__ void f8_decode(i32x8_t (&af)[FT<Form::F8>::KS / 128][WM],
                                          const aword_t<Form::F8> (&aw)[WM]) {
#pragma unroll
  for (int m = 0; m < FT<Form::F8>::KS / 128; ++m)
#pragma unroll
    for (int rb = 0; rb < WM; ++rb)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int t = 0; t < 4; ++t) af[m][rb][4 * i + t] = (int)top_bytes(spread(aw[rb][2 * m + i], t));
}

template <int WM, int NC>
__device__ __forceinline__ void f8_mfma(f32x4_t (&acc)[WM][NC], const i32x8_t (&af)[FT<Form::F8>::KS / 128][WM],
                                        const uint8_t* wl, int q, int r16) {
  using T = FT<Form::F8>;
  constexpr int LP = NC * 16;
  constexpr int LDSP = T::KS;
  constexpr int NSUB = T::KS / 128;
  constexpr int VQ = T::KS / 64;  // 16-byte W vectors per lane group
  auto read_b = [&](int m, int cb, int p) {
    const uint8_t* row = wl + (p * LP + cb * 16 + r16) * LDSP;
    const u32x4_t v0 = *reinterpret_cast<const u32x4_t*>(row + swz<Form::F8>(r16, VQ * q + 2 * m) * 16);
    const u32x4_t v1 = *reinterpret_cast<const u32x4_t*>(row + swz<Form::F8>(r16, VQ * q + 2 * m + 1) * 16);
    i32x8_t b;
#pragma unroll
    for (int e = 0; e < 4; ++e) b[e] = (int)v0[e], b[4 + e] = (int)v1[e];
    return b;
  };
#pragma unroll
  for (int m = 0; m < NSUB; ++m)
#pragma unroll
    for (int cb = 0; cb < NC; ++cb)
#pragma unroll
      for (int p = 0; p < T::P; ++p) {
        const i32x8_t bv = read_b(m, cb, p);
#pragma unroll
        for (int rb = 0; rb < WM; ++rb)  // scale operands 0: the plain (unscaled) product
          acc[rb][cb] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af[m][rb], bv, acc[rb][cb], 0, 0, 0, 0, 0, 0);
      }
}

// All MFMAs of one K-step for one wave.
template <Form F, int WM, int NC>
__device__ __forceinline__ void compute_step(accv_t<F> (&acc)[FT<F>::NACC][WM][NC], const aword_t<F> (&aw)[WM],
                                             const typename FT<F>::w_t* wl, int q, int r16) {
  using T = FT<F>;
  using w_t = typename T::w_t;
  using acc_t = typename T::acc_t;
  constexpr int LP = NC * 16;
  constexpr int LDSP = T::KS;
  if constexpr (F == Form::I8) {
    // 16x16x64 int8 MFMAs: lane group q's word s holds the 16 codes of
    // substep s, and B vector 4q + s of each W row the same 16 K slots.
#pragma unroll
    for (int s = 0; s < T::KS / 64; ++s) {
      u32x4_t bv[NC * T::P];
#pragma unroll
      for (int cb = 0; cb < NC; ++cb)
#pragma unroll
        for (int p = 0; p < T::P; ++p)
          bv[cb * T::P + p] = *reinterpret_cast<const u32x4_t*>(wl + (p * LP + cb * 16 + r16) * LDSP +
                                                                 swz<F>(r16, (T::KS / 64) * q + s) * 16);
      __builtin_amdgcn_sched_barrier(0);
      i32x4_t af[WM];
#pragma unroll
      for (int rb = 0; rb < WM; ++rb) af[rb] = decode_i8x16(aw[rb][s]);
#pragma unroll
      for (int cb = 0; cb < NC; ++cb)
#pragma unroll
        for (int p = 0; p < T::P; ++p)
#pragma unroll
          for (int rb = 0; rb < WM; ++rb)
            acc[p][rb][cb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                af[rb], __builtin_bit_cast(i32x4_t, bv[cb * T::P + p]), acc[p][rb][cb], 0, 0, 0);
    }
  } else if constexpr (F == Form::F8) {
    i32x8_t af[T::KS / 128][WM];
    f8_decode<WM>(af, aw);
    f8_mfma<WM, NC>(acc[0], af, wl, q, r16);
  } else if constexpr (F == Form::BF16 || F == Form::BF16X1 || F == Form::FP16) {
    // Fragment j of substep s + 1 is read from LDS into the registers of
    // fragment j of substep s right after that fragment's last MFMA, and the A
    // fragments of substep s + 1 are decoded behind substep s's MFMAs (two A
    // sets, one B set), so LDS and decode latency stay hidden.  Wave tiles of
    // 8 row blocks keep one A set and two B sets instead (their second A set
    // would spill), and tiles with more than 10 fragments per substep read
    // them just before their MFMAs.
    constexpr int NS = T::KS / 32;
    constexpr bool PF = NC * T::P <= 10;
    if constexpr (PF && WM <= 4) {
      auto read1 = [&](int s, int cb, int p) {
        return *reinterpret_cast<const u32x4_t*>(wl + (p * LP + cb * 16 + r16) * LDSP +
                                                 swz<F>(r16, (T::KS / 32) * q + s) * 8);
      };
      auto decode = [&](int s, bf16x8_t(&dst)[WM]) {
#pragma unroll
        for (int rb = 0; rb < WM; ++rb) dst[rb] = decode_bf16x8(aw[rb][s >> 1], s & 1);
      };
      u32x4_t bv[NC * T::P];
      bf16x8_t af[2][WM];
#pragma unroll
      for (int cb = 0; cb < NC; ++cb)
#pragma unroll
        for (int p = 0; p < T::P; ++p) bv[cb * T::P + p] = read1(0, cb, p);
      decode(0, af[0]);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int cb = 0; cb < NC; ++cb)
#pragma unroll
          for (int p = 0; p < T::P; ++p) {
#pragma unroll
            for (int rb = 0; rb < WM; ++rb) {
              if constexpr (F == Form::FP16)  // +-2 and 0 have the same top byte in FP16 as in BF16
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                    __builtin_bit_cast(f16x8_t, af[s & 1][rb]), __builtin_bit_cast(f16x8_t, bv[cb * T::P + p]),
                    acc[0][rb][cb], 0, 0, 0);
              else
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    af[s & 1][rb], __builtin_bit_cast(bf16x8_t, bv[cb * T::P + p]), acc[0][rb][cb], 0, 0, 0);
            }
            if (s + 1 < NS) bv[cb * T::P + p] = read1(s + 1, cb, p);
          }
        if (s + 1 < NS) decode(s + 1, af[(s + 1) & 1]);
      }
      __builtin_amdgcn_sched_barrier(0);
    } else if constexpr (PF) {
      auto read_b = [&](int s, u32x4_t(&dst)[NC * T::P]) {
#pragma unroll
        for (int cb = 0; cb < NC; ++cb)
#pragma unroll
          for (int p = 0; p < T::P; ++p)
            dst[cb * T::P + p] = *reinterpret_cast<const u32x4_t*>(wl + (p * LP + cb * 16 + r16) * LDSP +
                                                                    swz<F>(r16, (T::KS / 32) * q + s) * 8);
      };
      auto mfma_substep = [&](int s, const u32x4_t(&bv)[NC * T::P]) {
        bf16x8_t af[WM];
#pragma unroll
        for (int rb = 0; rb < WM; ++rb) af[rb] = decode_bf16x8(aw[rb][s >> 1], s & 1);
#pragma unroll
        for (int cb = 0; cb < NC; ++cb) {
#pragma unroll
          for (int p = 0; p < T::P; ++p) {
#pragma unroll
            for (int rb = 0; rb < WM; ++rb) {
              if constexpr (F == Form::FP16)  // +-2 and 0 have the same top byte in FP16 as in BF16
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, af[rb]),
                                                                        __builtin_bit_cast(f16x8_t, bv[cb * T::P + p]),
                                                                        acc[0][rb][cb], 0, 0, 0);
              else
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    af[rb], __builtin_bit_cast(bf16x8_t, bv[cb * T::P + p]), acc[0][rb][cb], 0, 0, 0);
            }
          }
        }
      };
      u32x4_t bvs[2][NC * T::P];
      read_b(0, bvs[0]);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (s + 1 < NS) read_b(s + 1, bvs[(s + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);  // keep the next reads ahead of these MFMAs
        mfma_substep(s, bvs[s & 1]);
      }
    } else {
      // All B fragments of a substep are read before its MFMAs, so the LDS
      // latency is paid once per substep instead of once per fragment.
#pragma unroll
      for (int s = 0; s < T::KS / 32; ++s) {
        u32x4_t bv[NC * T::P];
#pragma unroll
        for (int cb = 0; cb < NC; ++cb)
#pragma unroll
          for (int p = 0; p < T::P; ++p)
            bv[cb * T::P + p] = *reinterpret_cast<const u32x4_t*>(wl + (p * LP + cb * 16 + r16) * LDSP +
                                                                   swz<F>(r16, (T::KS / 32) * q + s) * 8);
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the MFMAs
        bf16x8_t af[WM];
#pragma unroll
        for (int rb = 0; rb < WM; ++rb) af[rb] = decode_bf16x8(aw[rb][s >> 1], s & 1);
#pragma unroll
        for (int cb = 0; cb < NC; ++cb) {
#pragma unroll
          for (int p = 0; p < T::P; ++p) {
#pragma unroll
            for (int rb = 0; rb < WM; ++rb) {
              if constexpr (F == Form::FP16)  // +-2 and 0 have the same top byte in FP16 as in BF16
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, af[rb]),
                                                                        __builtin_bit_cast(f16x8_t, bv[cb * T::P + p]),
                                                                        acc[0][rb][cb], 0, 0, 0);
              else
                acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    af[rb], __builtin_bit_cast(bf16x8_t, bv[cb * T::P + p]), acc[0][rb][cb], 0, 0, 0);
            }
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < T::KS / 16; ++u) {
      acc_t av[WM][4];
#pragma unroll
      for (int rb = 0; rb < WM; ++rb) {
        const uint32_t tb = top_bytes(spread(aw[rb][u >> 2], u & 3));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if constexpr (F == Form::FP32)
            av[rb][e] = __uint_as_float(top_word(tb, e));
          else
            av[rb][e] = __hiloint2double((int)top_word(tb, e), 0);
        }
      }
      // all B values of this u-step are read before its MFMAs (one LDS wait)
      acc_t bw[NC][4];
#pragma unroll
      for (int cb = 0; cb < NC; ++cb) {
        const w_t* row = wl + (cb * 16 + r16) * LDSP;
        if constexpr (F == Form::FP32) {
          const float4 b4 = *reinterpret_cast<const float4*>(row + swz<F>(r16, (T::KS / 16) * q + u) * 4);
          bw[cb][0] = b4.x, bw[cb][1] = b4.y, bw[cb][2] = b4.z, bw[cb][3] = b4.w;
        } else {
          const double2 b0 = *reinterpret_cast<const double2*>(row + swz<F>(r16, (T::KS / 8) * q + 2 * u) * 2);
          const double2 b1 = *reinterpret_cast<const double2*>(row + swz<F>(r16, (T::KS / 8) * q + 2 * u + 1) * 2);
          bw[cb][0] = b0.x, bw[cb][1] = b0.y, bw[cb][2] = b1.x, bw[cb][3] = b1.y;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int cb = 0; cb < NC; ++cb) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int rb = 0; rb < WM; ++rb) {
            if constexpr (F == Form::FP32)
              acc[0][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[rb][e], bw[cb][e], acc[0][rb][cb], 0, 0, 0);
            else
              acc[0][rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb][e], bw[cb][e], acc[0][rb][cb], 0, 0, 0);
          }
      }
    }
  }
}

// One row tile (WAVES x WM x 16 rows) over codes [kb, ke): Y receives the
// tile's partial products (row-major, LP columns) at row `rowtile`.
template <Form F, int WM, int NC, int WAVES>
__device__ __forceinline__ void gemm_tile(const uint32_t* __restrict__ X, uint64_t xpitch, const void* Wt,
                                          uint64_t wpitch, uint64_t wpiece, void* Yv, uint64_t rowtile, uint64_t kb,
                                          uint64_t ke, unsigned char* smem) {
  using T = FT<F>;
  using w_t = typename T::w_t;
  using acc_t = typename T::acc_t;
  constexpr int LP = NC * 16;
  constexpr int LDSP = T::KS;
  constexpr int VE = 16 / sizeof(w_t);
  constexpr int VPR = T::KS / VE;
  constexpr int NV = T::P * LP * VPR;           // 16-byte vectors per W tile
  constexpr int NT = 64 * WAVES;                // threads per workgroup
  constexpr int NVT = (NV + NT - 1) / NT;       // per thread (the last round may be partial)
  constexpr bool EVEN = NV % NT == 0;
  constexpr bool DB = double_buffered<F, NC>();
  constexpr size_t TILE_ELEMS = tile_bytes<F, NC>() / sizeof(w_t);

  w_t* wl0 = reinterpret_cast<w_t*>(smem);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  const uint64_t rowbase = rowtile + (uint64_t)wave * WM * 16;

  accv_t<F> acc[T::NACC][WM][NC];
#pragma unroll
  for (int a = 0; a < T::NACC; ++a)
#pragma unroll
    for (int rb = 0; rb < WM; ++rb)
#pragma unroll
      for (int cb = 0; cb < NC; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][rb][cb][r] = 0;

  const uint32_t* xr[WM];
#pragma unroll
  for (int rb = 0; rb < WM; ++rb) xr[rb] = X + (rowbase + rb * 16 + r16) * xpitch + (T::KS / 64) * q;
  const w_t* wg = static_cast<const w_t*>(Wt);

  // Thread tid moves W-tile vectors v = tid + NT i.
  static_assert(NT % VPR == 0, "a W tile row must split evenly over the workgroup");
  const int vrow0 = tid / VPR, vpart = tid % VPR;  // NT is a multiple of VPR
  constexpr int ROWS_PER_I = NT / VPR;

  // Double-buffered tiles are filled by direct global -> LDS loads: a wave
  // writes 64 consecutive 16-byte slots (64 / VPR whole rows), lane l slot
  // l, and fetches the W vector that the swizzle places there (the XOR
  // swizzle is its own inverse).  No registers hold the tile in flight.
  constexpr bool DMA = DB;
  static_assert(64 % VPR == 0, "a wave's 64 slots must cover whole W-tile rows");
  auto tile_dma = [&](w_t* dst, uint64_t kk) {
#pragma unroll
    for (int i = 0; i < NVT; ++i) {
      const int wrow = wave * (64 / VPR) + ROWS_PER_I * i;  // first row of this wave's slots (uniform)
      if (EVEN || wrow < T::P * LP) {
        const int row = vrow0 + ROWS_PER_I * i, piece = row / LP, col = row - piece * LP;
        const w_t* src = wg + piece * wpiece + col * wpitch + swz<F>(row, vpart) * VE + kk;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                         (__attribute__((address_space(3))) void*)(dst + wrow * LDSP), 16, 0, 0);
      }
    }
  };

  u32x4_t wreg[DMA ? 1 : NVT];
  aword_t<F> anext[WM], acur[WM];

  if (kb < ke) {
#pragma unroll
    for (int rb = 0; rb < WM; ++rb) acur[rb] = *reinterpret_cast<const aword_t<F>*>(xr[rb] + (kb >> 4));
    if constexpr (DMA) {
      tile_dma(wl0, kb);
    } else {
#pragma unroll
      for (int i = 0; i < NVT; ++i) {
        const int row = vrow0 + ROWS_PER_I * i, piece = row / LP, col = row - piece * LP;
        if (EVEN || row < T::P * LP)
          wreg[i] = *reinterpret_cast<const u32x4_t*>(wg + piece * wpiece + col * wpitch + vpart * VE + kb);
      }
#pragma unroll
      for (int i = 0; i < NVT; ++i)
        if (EVEN || vrow0 + ROWS_PER_I * i < T::P * LP)
          *reinterpret_cast<u32x4_t*>(wl0 + (vrow0 + ROWS_PER_I * i) * LDSP +
                                      swz<F>(vrow0 + ROWS_PER_I * i, vpart) * VE) = wreg[i];
    }
    // Nothing may stay pending into the loop: the wait inserted for a load
    // still in flight at the loop head would also cover the loop's own
    // prefetch and expose its latency every step.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int buf = 0;
    for (uint64_t k0 = kb; k0 < ke; k0 += T::KS) {
      const bool more = k0 + T::KS < ke;
      const uint64_t kn = k0 + T::KS;
      if (more) {
        if constexpr (DMA) {
          // the other buffer was last read in the previous step, before its barrier
          tile_dma(wl0 + (buf ^ 1) * TILE_ELEMS, kn);
        } else {
#pragma unroll
          for (int i = 0; i < NVT; ++i) {
            const int row = vrow0 + ROWS_PER_I * i, piece = row / LP, col = row - piece * LP;
            if (EVEN || row < T::P * LP)
              wreg[i] = *reinterpret_cast<const u32x4_t*>(wg + piece * wpiece + col * wpitch + vpart * VE + kn);
          }
        }
#pragma unroll
        for (int rb = 0; rb < WM; ++rb) anext[rb] = *reinterpret_cast<const aword_t<F>*>(xr[rb] + (kn >> 4));
      }
      w_t* cur = wl0 + (DB ? buf * TILE_ELEMS : 0);
      compute_step<F, WM, NC>(acc, acur, cur, q, r16);
      if (more) {
        if constexpr (!DMA) {
          w_t* nxt = wl0 + (DB ? (buf ^ 1) * TILE_ELEMS : 0);
          if constexpr (!DB) __syncthreads();
#pragma unroll
          for (int i = 0; i < NVT; ++i)
            if (EVEN || vrow0 + ROWS_PER_I * i < T::P * LP)
              *reinterpret_cast<u32x4_t*>(nxt + (vrow0 + ROWS_PER_I * i) * LDSP +
                                          swz<F>(vrow0 + ROWS_PER_I * i, vpart) * VE) = wreg[i];
        }
#pragma unroll
        for (int rb = 0; rb < WM; ++rb) acur[rb] = anext[rb];
      }
      // (Letting the next step's code words cross the barrier, a bare barrier
      // after a wait for the tile DMA only, measured the same on one box: FP8
      // 8.65-8.73 ms, BF16 10.89-11.03 ms per launch either way.)
      if constexpr (DMA) __builtin_amdgcn_s_waitcnt(0);  // this wave's tile loads have landed
      __syncthreads();
      buf ^= 1;
    }
  }

  // Accumulator layout: lane group q, register r holds row 4q + r of a 16x16
  // block for the FP32- and int32-accumulating MFMAs, and row q + 4r for the
  // FP64 one.  I8: the chunk's product is 256 (A q_hi) + A q_lo, exact in
  // int32 for chunks of at most kI8ChunkCap codes.
  acc_t* Y = static_cast<acc_t*>(Yv);
#pragma unroll
  for (int rb = 0; rb < WM; ++rb)
#pragma unroll
    for (int cb = 0; cb < NC; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = F == Form::FP64 ? q + 4 * r : 4 * q + r;
        acc_t v;
        if constexpr (F == Form::I8)
          v = (int32_t)((uint32_t)acc[0][rb][cb][r] << 8) + acc[1][rb][cb][r];
        else
          v = acc[0][rb][cb][r];
        Y[(rowbase + rb * 16 + row) * LP + cb * 16 + r16] = v;
      }
}

template <Form F, int WM, int NC, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_packed_gemm(GemmArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using acc_t = typename FT<F>::acc_t;
  const uint64_t kb = a.k_begin + (uint64_t)blockIdx.y * a.kchunk;
  const uint64_t ke = (kb + a.kchunk < a.k_end) ? kb + a.kchunk : a.k_end;
  gemm_tile<F, WM, NC, WAVES>(a.X, a.xpitch, a.Wt, a.wpitch, a.wpiece,
                              static_cast<acc_t*>(a.Y) + (uint64_t)blockIdx.y * a.ychunk,
                              (uint64_t)blockIdx.x * (WAVES * WM * 16), kb, ke, smem);
}

// Persistent grouped form: the workgroups walk a list of (cohort, row tile,
// chunk) items of many independent products; items of inactive cohorts are
// skipped.  Every workgroup reaches the end of the list.
template <Form F, int WM, int NC, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_grouped_gemm(GroupedGemmArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using acc_t = typename FT<F>::acc_t;
  for (uint32_t it = blockIdx.x; it < a.nitems; it += gridDim.x) {
    const GemmItem item = a.items[it];
    if (a.active && !a.active[item.cohort]) continue;
    const GemmCohort c = a.cohorts[item.cohort];
    gemm_tile<F, WM, NC, WAVES>(c.X, c.xpitch, c.Wt, c.wpitch, c.wpiece,
                                static_cast<acc_t*>(c.Y) + (uint64_t)item.chunk * c.ychunk, item.row0, item.k0,
                                item.k1, smem);
  }
}

// Launches the grouped kernel, or (occ != null) reports how many of its
// workgroups fit on one CU.
template <Form F, int WM, int NC, int WAVES>
void launch_grouped_t(const GroupedGemmArgs& a, unsigned grid, hipStream_t s, int* occ) {
  constexpr size_t lds = lds_bytes<F, NC>();
  const void* fn = reinterpret_cast<const void*>(&k_grouped_gemm<F, WM, NC, WAVES>);
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  if (occ) {
    PCAONE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, fn, 64 * WAVES, lds));
    return;
  }
  hipLaunchKernelGGL((k_grouped_gemm<F, WM, NC, WAVES>), dim3(grid), dim3(64 * WAVES), lds, s, a);
  check_launch("grouped_gemm");
}

// Launches the kernel, or (occ != null) reports how many of its workgroups
// fit on one CU.
template <Form F, int WM, int NC, int WAVES>
void launch_t(const GemmArgs& a, hipStream_t s, int* occ) {
  constexpr size_t lds = lds_bytes<F, NC>();
  const void* fn = reinterpret_cast<const void*>(&k_packed_gemm<F, WM, NC, WAVES>);
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  if (occ) {
    PCAONE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, fn, 64 * WAVES, lds));
    return;
  }
  const uint64_t tile = (uint64_t)WAVES * WM * 16;
  const uint64_t nchunks = (a.k_end - a.k_begin + a.kchunk - 1) / a.kchunk;
  dim3 grid((unsigned)(a.rows / tile), (unsigned)nchunks);
  hipLaunchKernelGGL((k_packed_gemm<F, WM, NC, WAVES>), grid, dim3(64 * WAVES), lds, s, a);
  check_launch("packed_gemm");
}

// Which (form, wm, waves) combinations are compiled.
template <Form F>
constexpr bool shape_ok(int wm, int waves) {
  if (F == Form::F8) return (waves == 4 && (wm == 1 || wm == 2 || wm == 4)) || (waves == 8 && (wm == 2 || wm == 4));
  if (waves == 4) return wm == 1 || wm == 2 || (wm == 4 && F != Form::FP64) || (wm == 8 && F != Form::FP64 && F != Form::FP32);
  if (waves == 8) return (wm == 2 || wm == 4) && F != Form::FP64;
  return false;
}

// Calls fn.template operator()<NC>() for lp = 16 NC.
template <Form F, int WAVES, class Fn>
bool with_nc(int lp, Fn&& fn) {
  switch (lp / 16) {
#define PCAONE_NC(n) \
  case n:            \
    fn.template operator()<n>();    \
    return true;
    PCAONE_NC(1) PCAONE_NC(2) PCAONE_NC(3) PCAONE_NC(4) PCAONE_NC(5) PCAONE_NC(6) PCAONE_NC(7) PCAONE_NC(8)
#undef PCAONE_NC
    default:
      break;
  }
  return false;
}

template <Form F, int WM, int WAVES>
void launch_nc(const GemmArgs& a, hipStream_t s, int* occ) {
  if constexpr (shape_ok<F>(WM, WAVES)) {
    if (with_nc<F, WAVES>(a.lp, [&]<int NC>() { launch_t<F, WM, NC, WAVES>(a, s, occ); })) return;
  }
  fail(kErrArgument, "packed_gemm: unsupported shape (wm " + std::to_string(WM) + ", waves " +
                         std::to_string(WAVES) + ", lp " + std::to_string(a.lp) + ")");
}

template <Form F, int WM, int WAVES>
void launch_grouped_nc(int lp, const GroupedGemmArgs& a, unsigned grid, hipStream_t s, int* occ) {
  if constexpr (shape_ok<F>(WM, WAVES)) {
    if (with_nc<F, WAVES>(lp, [&]<int NC>() { launch_grouped_t<F, WM, NC, WAVES>(a, grid, s, occ); })) return;
  }
  fail(kErrArgument, "grouped_gemm: unsupported shape (wm " + std::to_string(WM) + ", waves " +
                         std::to_string(WAVES) + ", lp " + std::to_string(lp) + ")");
}

template <Form F>
void launch_grouped_shape(const GemmShape& g, int lp, const GroupedGemmArgs& a, unsigned grid, hipStream_t s,
                          int* occ) {
  if (g.waves == 8) {
    if (g.wm == 2) return launch_grouped_nc<F, 2, 8>(lp, a, grid, s, occ);
    if (g.wm == 4) return launch_grouped_nc<F, 4, 8>(lp, a, grid, s, occ);
  } else if (g.waves == 4) {
    if (g.wm == 1) return launch_grouped_nc<F, 1, 4>(lp, a, grid, s, occ);
    if (g.wm == 2) return launch_grouped_nc<F, 2, 4>(lp, a, grid, s, occ);
    if (g.wm == 4) return launch_grouped_nc<F, 4, 4>(lp, a, grid, s, occ);
    if (g.wm == 8) return launch_grouped_nc<F, 8, 4>(lp, a, grid, s, occ);
  }
  fail(kErrArgument, "grouped_gemm: unsupported wave tile " + std::to_string(g.wm) + " x " + std::to_string(g.waves));
}

template <Form F>
void launch_shape(const GemmShape& g, const GemmArgs& a, hipStream_t s, int* occ = nullptr) {
  if (g.waves == 8) {
    if (g.wm == 2) return launch_nc<F, 2, 8>(a, s, occ);
    if (g.wm == 4) return launch_nc<F, 4, 8>(a, s, occ);
  } else if (g.waves == 4) {
    if (g.wm == 1) return launch_nc<F, 1, 4>(a, s, occ);
    if (g.wm == 2) return launch_nc<F, 2, 4>(a, s, occ);
    if (g.wm == 4) return launch_nc<F, 4, 4>(a, s, occ);
    if (g.wm == 8) return launch_nc<F, 8, 4>(a, s, occ);
  }
  fail(kErrArgument, "packed_gemm: unsupported wave tile " + std::to_string(g.wm) + " x " + std::to_string(g.waves));
}

}  // namespace

GemmShape gemm_default_shape(Form f, int lp) {
  // Tuned at 50,000 x 1,000,000, L = 80 on MI355X: 8 waves x 4 row blocks
  // (two waves per SIMD sharing one LDS tile) beats 4 waves x 8 for BF16.
  if (f == Form::FP64) return {1, 4};
  if (f == Form::FP32) return lp > 96 ? GemmShape{2, 4} : GemmShape{4, 8};
  if (f == Form::BF16) return lp > 80 ? GemmShape{4, 4} : GemmShape{4, 8};
  if (f == Form::I8) return lp > 64 ? GemmShape{2, 8} : GemmShape{4, 8};  // two accumulator sets
  if (f == Form::F8) return lp > 80 ? GemmShape{2, 8} : GemmShape{4, 8};  // 16 A registers per row block
  return lp > 96 ? GemmShape{4, 4} : GemmShape{4, 8};
}

bool gemm_shape_supported(Form f, const GemmShape& g, int lp) {
  if (lp < 16 || lp > 128 || lp % 16) return false;
  switch (f) {
    case Form::FP64: return shape_ok<Form::FP64>(g.wm, g.waves);
    case Form::FP32: return shape_ok<Form::FP32>(g.wm, g.waves);
    case Form::BF16: return shape_ok<Form::BF16>(g.wm, g.waves);
    case Form::BF16X1: return shape_ok<Form::BF16X1>(g.wm, g.waves);
    case Form::FP16: return shape_ok<Form::FP16>(g.wm, g.waves);
    case Form::I8: return shape_ok<Form::I8>(g.wm, g.waves);
    case Form::F8: return shape_ok<Form::F8>(g.wm, g.waves);
  }
  return false;
}

void launch_grouped_gemm(Form f, const GemmShape& g, int lp, const GroupedGemmArgs& a, unsigned grid,
                         hipStream_t s) {
  require(lp >= 16 && lp <= 128 && lp % 16 == 0, kErrArgument, "grouped_gemm: bad lp");
  require(gemm_shape_supported(f, g, lp), kErrArgument, "grouped_gemm: unsupported shape");
  if (!a.nitems || !grid) return;
  switch (f) {
    case Form::FP64: return launch_grouped_shape<Form::FP64>(g, lp, a, grid, s, nullptr);
    case Form::FP32: return launch_grouped_shape<Form::FP32>(g, lp, a, grid, s, nullptr);
    case Form::BF16: return launch_grouped_shape<Form::BF16>(g, lp, a, grid, s, nullptr);
    case Form::BF16X1: return launch_grouped_shape<Form::BF16X1>(g, lp, a, grid, s, nullptr);
    case Form::FP16: return launch_grouped_shape<Form::FP16>(g, lp, a, grid, s, nullptr);
    case Form::I8: return launch_grouped_shape<Form::I8>(g, lp, a, grid, s, nullptr);
    case Form::F8: return launch_grouped_shape<Form::F8>(g, lp, a, grid, s, nullptr);
  }
}

int grouped_gemm_blocks_per_cu(Form f, const GemmShape& g, int lp) {
  require(gemm_shape_supported(f, g, lp), kErrArgument, "grouped_gemm: unsupported shape");
  int occ = 0;
  const GroupedGemmArgs none{};
  switch (f) {
    case Form::FP64: launch_grouped_shape<Form::FP64>(g, lp, none, 0, nullptr, &occ); break;
    case Form::FP32: launch_grouped_shape<Form::FP32>(g, lp, none, 0, nullptr, &occ); break;
    case Form::BF16: launch_grouped_shape<Form::BF16>(g, lp, none, 0, nullptr, &occ); break;
    case Form::BF16X1: launch_grouped_shape<Form::BF16X1>(g, lp, none, 0, nullptr, &occ); break;
    case Form::FP16: launch_grouped_shape<Form::FP16>(g, lp, none, 0, nullptr, &occ); break;
    case Form::I8: launch_grouped_shape<Form::I8>(g, lp, none, 0, nullptr, &occ); break;
    case Form::F8: launch_grouped_shape<Form::F8>(g, lp, none, 0, nullptr, &occ); break;
  }
  return occ > 0 ? occ : 1;
}

int gemm_blocks_per_cu(Form f, const GemmShape& g, int lp) {
  require(gemm_shape_supported(f, g, lp), kErrArgument, "packed_gemm: unsupported shape");
  static std::mutex mu;
  static std::map<int, int> cache;
  const int key = ((int)f * 16 + g.wm) * 16 * 16 + g.waves * 16 + lp / 16;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int occ = 0;
  GemmArgs none{};
  none.lp = lp;
  switch (f) {
    case Form::FP64: launch_shape<Form::FP64>(g, none, nullptr, &occ); break;
    case Form::FP32: launch_shape<Form::FP32>(g, none, nullptr, &occ); break;
    case Form::BF16: launch_shape<Form::BF16>(g, none, nullptr, &occ); break;
    case Form::BF16X1: launch_shape<Form::BF16X1>(g, none, nullptr, &occ); break;
    case Form::FP16: launch_shape<Form::FP16>(g, none, nullptr, &occ); break;
    case Form::I8: launch_shape<Form::I8>(g, none, nullptr, &occ); break;
    case Form::F8: launch_shape<Form::F8>(g, none, nullptr, &occ); break;
  }
  occ = occ > 0 ? occ : 1;
  cache[key] = occ;
  return occ;
}

namespace {
constexpr uint64_t kTargetWorkgroups = 2048;  // 8 per CU
constexpr uint64_t kMinStepsPerChunk = 4;
// Products of at most this many rows may take up to two more split-K chunks
// than the rules below ask for, when that fills the last workgroup round.
constexpr uint64_t kQuantMaxRows = 1ull << 22;
constexpr uint64_t kQuantExtraChunks = 2;

// Share of the launch's workgroup rounds that is filled: w workgroups run in
// ceil(w / slots) rounds of `slots` resident workgroups.
double round_fill(uint64_t w, uint64_t slots) {
  const uint64_t r = (w + slots - 1) / slots;
  return (double)w / (double)(r * slots);
}

GemmPlan plan_for(const GemmShape& sh, uint64_t nrows, uint64_t kext, uint64_t kchunk_user, uint64_t kcap,
                  uint64_t slots) {
  GemmPlan p;
  p.shape = sh;
  const uint64_t tile = (uint64_t)gemm_row_tile(sh);
  p.rows = (nrows + tile - 1) / tile * tile;
  const uint64_t nrt = p.rows / tile;
  auto chunk_of = [&](uint64_t nch) {
    const uint64_t kc = (kext + nch - 1) / nch;
    return std::max<uint64_t>(kKStep, (kc + kKStep - 1) / kKStep * kKStep);
  };
  uint64_t kc = kchunk_user;
  if (!kc) {
    uint64_t nch = nrt >= kTargetWorkgroups / 2 ? 1 : (kTargetWorkgroups + nrt - 1) / nrt;
    nch = std::max<uint64_t>(nch, (kext + kcap - 1) / kcap);
    const uint64_t nmax = std::max<uint64_t>(1, kext / (kMinStepsPerChunk * kKStep));
    nch = std::min<uint64_t>(nch, nmax);
    kc = chunk_of(nch);
    if (slots && p.rows <= kQuantMaxRows) {
      // A last round with few workgroups leaves most CUs idle (1,954 row
      // tiles on 256 CUs fill 7.63 of 8 rounds; three chunks fill 22.9 of 23).
      // Each extra chunk costs one more partial block: a point of fill.
      double best = round_fill(nrt * ((kext + kc - 1) / kc), slots);
      for (uint64_t n = nch + 1; n <= std::min(nch + kQuantExtraChunks, nmax); ++n) {
        const uint64_t c = chunk_of(n);
        if (c > kcap) continue;
        const double fill = round_fill(nrt * ((kext + c - 1) / c), slots) - 0.01 * (double)(n - nch);
        if (fill > best + 1e-9) best = fill, kc = c;
      }
    }
  }
  p.kchunk = std::max<uint64_t>(kKStep, (kc + kKStep - 1) / kKStep * kKStep);
  p.nchunks = std::max<uint64_t>(1, (kext + p.kchunk - 1) / p.kchunk);
  return p;
}
}  // namespace

GemmPlan plan_gemm(Form f, const GemmShape& preferred, bool fixed_shape, int lp, uint64_t nrows, uint64_t kext,
                   uint64_t kchunk_user, uint64_t kcap, uint64_t cus) {
  auto slots = [&](const GemmShape& sh) { return cus ? cus * (uint64_t)gemm_blocks_per_cu(f, sh, lp) : 0; };
  GemmPlan best = plan_for(preferred, nrows, kext, kchunk_user, kcap, slots(preferred));
  if (fixed_shape || best.workgroups() >= kTargetWorkgroups / 2) return best;
  static const GemmShape smaller[] = {{4, 8}, {4, 4}, {2, 8}, {2, 4}, {1, 4}};
  for (const GemmShape& sh : smaller) {
    if (gemm_row_tile(sh) >= gemm_row_tile(preferred) || !gemm_shape_supported(f, sh, lp)) continue;
    const GemmPlan p = plan_for(sh, nrows, kext, kchunk_user, kcap, slots(sh));
    if (p.workgroups() > best.workgroups()) best = p;
    if (best.workgroups() >= kTargetWorkgroups / 2) break;
  }
  return best;
}

uint64_t plan_partial_rows(uint64_t nrows_max, uint64_t kext_max, uint64_t kchunk_user, uint64_t kcap) {
  // nchunks <= max(target / row tiles + 1, kext / kcap + 1) when automatic, so
  // nchunks * rows <= target * 512 + rows or (kext / kcap + 1) * rows, plus
  // the extra chunks that fill the last round on products of up to
  // kQuantMaxRows rows.
  const uint64_t rows = nrows_max + 512;
  if (kchunk_user) return ((kext_max + kchunk_user - 1) / kchunk_user + 1) * rows;
  return std::max(kTargetWorkgroups * 512 + rows, ((kext_max + kcap - 1) / kcap + 1) * rows) +
         kQuantExtraChunks * std::min(rows, kQuantMaxRows + 512);
}

void launch_packed_gemm(Form f, const GemmShape& g, const GemmArgs& a, hipStream_t s) {
  require(a.lp >= 16 && a.lp <= 128 && a.lp % 16 == 0, kErrArgument, "packed_gemm: bad lp");
  require(a.kchunk > 0 && a.kchunk % kKStep == 0 && a.k_begin % kKStep == 0 && a.k_end % kKStep == 0,
          kErrArgument, "packed_gemm: K range must be multiples of 256");
  require(gemm_shape_supported(f, g, a.lp), kErrArgument, "packed_gemm: unsupported shape");
  require(a.rows % (uint64_t)gemm_row_tile(g) == 0, kErrArgument, "packed_gemm: rows not a tile multiple");
  require(a.xpitch % 8 == 0 && a.xpitch * 16 >= a.k_end, kErrArgument, "packed_gemm: X pitch too small");
  require(a.wpitch % 16 == 0 && a.wpitch >= a.k_end, kErrArgument, "packed_gemm: W pitch too small");
  require(a.wpiece % 16 == 0, kErrArgument, "packed_gemm: W piece stride not 16-byte aligned");
  require(f != Form::I8 || a.kchunk <= kI8ChunkCap, kErrArgument, "packed_gemm: I8 chunks are limited to 65536 codes");
  if (a.rows == 0 || a.k_end <= a.k_begin) return;
  switch (f) {
    case Form::FP64: return launch_shape<Form::FP64>(g, a, s);
    case Form::FP32: return launch_shape<Form::FP32>(g, a, s);
    case Form::BF16: return launch_shape<Form::BF16>(g, a, s);
    case Form::BF16X1: return launch_shape<Form::BF16X1>(g, a, s);
    case Form::FP16: return launch_shape<Form::FP16>(g, a, s);
    case Form::I8: return launch_shape<Form::I8>(g, a, s);
    case Form::F8: return launch_shape<Form::F8>(g, a, s);
  }
}

}  // namespace pcaone
