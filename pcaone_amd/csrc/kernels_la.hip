// FP64 kernels around the genotype products: MFMA operand pieces, the affine
// epilogues of the two products, column sums, Gram reductions (CholeskyQR2,
// B*B^T, mev), small right-multiplications, flipOmg norms and the
// missing-entry (and EMU imputation) corrections.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "bf16.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace pcaone {
namespace {

inline unsigned grid_for(uint64_t n, unsigned block = 256, unsigned cap = 65535u * 8) {
  uint64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  return (unsigned)(g ? g : 1);
}

#define GRID_STRIDE(t, n) \
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (n); t += (uint64_t)gridDim.x * blockDim.x)


// Operand pieces of one form for each value y = X[r][l] colsign[l] rowscale[r]
// (F8 and I8 also divided by the column scale), and the value they sum to.
template <Form F>
struct Pieces;
template <>
struct Pieces<Form::FP64> {
  using w_t = double;
  static constexpr int P = 1;
  __device__ static double make(double y, double, w_t* pc) { return pc[0] = y; }
};
template <>
struct Pieces<Form::FP32> {
  using w_t = float;
  static constexpr int P = 1;
  __device__ static double make(double y, double, w_t* pc) { return (double)(pc[0] = (float)y); }
};
template <>
struct Pieces<Form::FP16> {
  using w_t = uint16_t;
  static constexpr int P = 1;
  __device__ static double make(double y, double, w_t* pc) {
    pc[0] = fp16_bits(y);
    return fp16_value(pc[0]);
  }
};
template <>
struct Pieces<Form::BF16X1> {
  using w_t = uint16_t;
  static constexpr int P = 1;
  __device__ static double make(double y, double, w_t* pc) {
    pc[0] = bf16_bits((float)y);
    return (double)bf16_value(pc[0]);
  }
};
template <>
struct Pieces<Form::BF16> {
  using w_t = uint16_t;
  static constexpr int P = 2;
  __device__ static double make(double y, double, w_t* pc) {
    pc[0] = bf16_bits((float)y);
    const double back = (double)bf16_value(pc[0]);
    pc[1] = bf16_bits((float)(y - back));
    return back + (double)bf16_value(pc[1]);
  }
};
template <>
struct Pieces<Form::F8> {
  using w_t = uint8_t;
  static constexpr int P = 3;
  // y / s (exact: s is a power of two, |y / s| <= 448) as three E4M3
  // pieces, each the nearest E4M3 to what the previous ones left
  __device__ static double make(double y, double s, w_t* pc) {
    const double x1 = y / s;
    double v1, v2, v3;
    pc[0] = e4m3_bits(x1, &v1);
    pc[1] = e4m3_bits(x1 - v1, &v2);
    pc[2] = e4m3_bits(x1 - v1 - v2, &v3);
    return (v1 + v2 + v3) * s;
  }
};
template <>
struct Pieces<Form::I8> {
  using w_t = uint8_t;
  static constexpr int P = 2;
  // q = rint(y / s) within +-kI8Max; q = 256 q_hi + q_lo, q_lo in [-128, 127]
  __device__ static double make(double y, double s, w_t* pc) {
    const int q = (int)fmin(fmax(rint(y / s), -(double)kI8Max), (double)kI8Max);
    const int hi = (q + 128) >> 8, lo = q - 256 * hi;
    pc[0] = (uint8_t)(int8_t)hi;
    pc[1] = (uint8_t)(int8_t)lo;
    return (double)q * s;
  }
};

// Rows per workgroup of the row-tiled kernels below and the LDS row pitch of
// their transposed pieces (element size wsz bytes): padded so that the
// column groups of one store land in different banks.
constexpr int kEpiRows = 64;
constexpr int epi_pitch(size_t wsz) { return kEpiRows + (wsz == 2 ? 2 : 1); }

// Operand pieces of X (rows x lp, row-major FP64), transposed to
// Wt[p][l][r], for a tile of 64 rows per workgroup.  Each thread owns 4
// adjacent columns of every RG-th row of the tile (16-byte loads along the
// row); the pieces go through LDS so that the transposed stores are coalesced
// along the rows.  Optional outputs: signed_out (X colsign), rounded (the
// pieces' sum / rowscale; in place over X allowed: every element is read and
// written by the same thread), and the per-tile column sums of
// w_r rounded[r][l] (w_r = bsc[r] (1/2 - bcf[r]), or 1 without bsc) in bpart.
template <Form F>
__global__ void __launch_bounds__(256) k_make_pieces(const double* X, uint64_t nreal, uint64_t rows_pad, int lp,
                                                     const double* __restrict__ rowscale,
                                                     const double* __restrict__ colsign, void* Wt, uint64_t wpitch,
                                                     uint64_t wpiece, double* rounded, double* __restrict__ signed_out,
                                                     const double* __restrict__ colscale,
                                                     const double* __restrict__ bsc, const double* __restrict__ bcf,
                                                     double* __restrict__ bpart) {
  using PF = Pieces<F>;
  using w_t = typename PF::w_t;
  constexpr int P = PF::P;
  constexpr int PITCH = epi_pitch(sizeof(w_t));
  extern __shared__ __align__(16) unsigned char sm[];
  w_t* pt = reinterpret_cast<w_t*>(sm);  // P x lp rows of PITCH
  double* bred = reinterpret_cast<double*>(sm + ((size_t)P * lp * PITCH * sizeof(w_t) + 15) / 16 * 16);
  const int CG = lp / 4, RG = 256 / CG;
  const int tid = threadIdx.x, cg = tid % CG, rg = tid / CG;
  const bool act = rg < RG;
  const uint64_t r0 = (uint64_t)blockIdx.x * kEpiRows;
  const int nr = (int)(rows_pad - r0 < (uint64_t)kEpiRows ? rows_pad - r0 : kEpiRows);
  double b4[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
    double cs[4] = {1.0, 1.0, 1.0, 1.0}, sl[4] = {1.0, 1.0, 1.0, 1.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (colsign) cs[q] = colsign[4 * cg + q];
      if constexpr (F == Form::F8 || F == Form::I8) sl[q] = colscale[4 * cg + q];
    }
    for (int r = rg; r < nr; r += RG) {
      const uint64_t j = r0 + r;
      const uint64_t t = j * lp + 4 * cg;
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      if (j < nreal) {
        const double2 a = *reinterpret_cast<const double2*>(X + t);
        const double2 b = *reinterpret_cast<const double2*>(X + t + 2);
        x[0] = a.x * cs[0], x[1] = a.y * cs[1], x[2] = b.x * cs[2], x[3] = b.y * cs[3];
      }
      if (signed_out) {
        *reinterpret_cast<double2*>(signed_out + t) = double2{x[0], x[1]};
        *reinterpret_cast<double2*>(signed_out + t + 2) = double2{x[2], x[3]};
      }
      const double rs = rowscale ? rowscale[j] : 1.0;
      const double w = j < nreal ? (bsc ? bsc[j] * (0.5 - bcf[j]) : 1.0) : 0.0;
      double rd[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        w_t pc[P];
        const double back = PF::make(x[q] * rs, sl[q], pc);
        rd[q] = (j < nreal && rs != 0.0) ? back / rs : 0.0;
        b4[q] += w * rd[q];
#pragma unroll
        for (int p = 0; p < P; ++p) pt[(p * lp + 4 * cg + q) * PITCH + r] = pc[p];
      }
      if (rounded) {
        *reinterpret_cast<double2*>(rounded + t) = double2{rd[0], rd[1]};
        *reinterpret_cast<double2*>(rounded + t + 2) = double2{rd[2], rd[3]};
      }
    }
    if (bpart) {
#pragma unroll
      for (int q = 0; q < 4; ++q) bred[rg * lp + 4 * cg + q] = b4[q];
    }
  }
  __syncthreads();
  w_t* wt = static_cast<w_t*>(Wt);
  for (int v = threadIdx.x; v < P * lp * kEpiRows; v += 256) {
    const int r = v % kEpiRows, pl = v / kEpiRows, p = pl / lp, l = pl - p * lp;
    if (r < nr) wt[p * wpiece + l * wpitch + r0 + r] = pt[pl * PITCH + r];
  }
  if (bpart) {
    for (int l = threadIdx.x; l < lp; l += 256) {
      double sum = 0.0;
      for (int q = 0; q < RG; ++q) sum += bred[q * lp + l];
      bpart[(uint64_t)blockIdx.x * lp + l] = sum;
    }
  }
}

// colmul (nullable): per-column factor of the summed partials (the I8
// form's int32 partials are sum (v/2) q, so the factor is 2 s_l; the F8
// form's are sum v (w / s_l), so it is s_l).
template <class P>
__global__ void k_epi_g(const P* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                        uint64_t rows_pad, int lp, const double* __restrict__ sc, const double* __restrict__ cf,
                        const double* __restrict__ colsum, double* __restrict__ out, const double* __restrict__ colmul) {
  const uint64_t n = rows_pad * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t j = t / lp;
    const int l = (int)(t - j * lp);
    double v = 0.0;
    if (j < nreal) {
      double y = 0.0;
      for (uint64_t c = 0; c < nchunks; ++c) y += (double)part[c * stride + t];
      if (colmul) y *= colmul[l];
      const double s = sc[j];
      v = 0.25 * s * y + s * (0.5 - cf[j]) * colsum[l];
    }
    out[t] = v;
  }
}

// k_epi_g for 32-bit partials (FP32 or int32), 4 adjacent columns per thread
// (16-byte partial loads, 32-byte stores).
template <class PT>
__global__ void k_epi_g_v4(const PT* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                           uint64_t rows_pad, int lp, const double* __restrict__ sc, const double* __restrict__ cf,
                           const double* __restrict__ colsum, double* __restrict__ out,
                           const double* __restrict__ colmul) {
  using v4 = typename std::conditional<std::is_same<PT, float>::value, float4, int4>::type;
  const uint64_t n4 = rows_pad * (uint64_t)lp / 4;
  GRID_STRIDE(t4, n4) {
    const uint64_t t = 4 * t4;
    const uint64_t j = t / lp;
    const int l = (int)(t - j * lp);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < nreal) {
      double y[4] = {0.0, 0.0, 0.0, 0.0};
      for (uint64_t c = 0; c < nchunks; ++c) {
        const v4 a = *reinterpret_cast<const v4*>(part + c * stride + t);
        y[0] += (double)a.x, y[1] += (double)a.y, y[2] += (double)a.z, y[3] += (double)a.w;
      }
      const double s = sc[j], beta = s * (0.5 - cf[j]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (colmul) y[q] *= colmul[l + q];
        v[q] = 0.25 * s * y[q] + beta * colsum[l + q];
      }
    }
    double2* o = reinterpret_cast<double2*>(out + t);
    o[0] = double2{v[0], v[1]};
    o[1] = double2{v[2], v[3]};
  }
}

// Fused G epilogue for data without missing entries.  A workgroup takes 64
// SNP rows of the window: reduces the split-K partials, applies the affine
// correction (G_j = sc_j/4 y_j + sc_j (1/2 - cF_j) colsum), splits sc_j G_j
// into operand pieces, writes G~ = sum(pieces)/sc_j (FP64, row-major), the
// pieces transposed through LDS (coalesced along SNPs) and the per-block
// column sums of beta_j G~_j with beta_j = sc_j (1/2 - cF_j).  Each thread
// owns 4 adjacent columns (16-byte partial loads) of every RG-th row; its
// beta sums stay in registers and are combined across row groups in LDS.

template <Form F, class PT>
__global__ void __launch_bounds__(256) k_epi_g_fused(const PT* __restrict__ part, uint64_t nchunks, uint64_t stride,
                                                     uint64_t nwin, int lp, const double* __restrict__ sc,
                                                     const double* __restrict__ cf, const double* __restrict__ colsum,
                                                     double* __restrict__ gout, void* Wt, uint64_t wpitch,
                                                     uint64_t wpiece, double* __restrict__ bpart) {
  using w_t = typename std::conditional<F == Form::FP64, double,
                                        typename std::conditional<F == Form::FP32, float, uint16_t>::type>::type;
  constexpr int P = F == Form::BF16 ? 2 : 1;
  extern __shared__ __align__(16) unsigned char sm[];
  constexpr int PITCH = epi_pitch(sizeof(w_t));
  w_t* pt = reinterpret_cast<w_t*>(sm);  // P x lp rows of PITCH (kEpiRows used)
  double* bred = reinterpret_cast<double*>(sm + ((size_t)P * lp * PITCH * sizeof(w_t) + 15) / 16 * 16);
  const int CG = lp / 4, RG = 256 / CG;  // column groups of 4, row groups
  const int tid = threadIdx.x, cg = tid % CG, rg = tid / CG;
  const bool act = rg < RG;
  const uint64_t r0 = (uint64_t)blockIdx.x * kEpiRows;
  const int nr = (int)(nwin - r0 < (uint64_t)kEpiRows ? nwin - r0 : kEpiRows);
  double b4[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
    for (int r = rg; r < nr; r += RG) {
      const uint64_t j = r0 + r;
      const uint64_t t = j * lp + 4 * cg;
      double y[4] = {0.0, 0.0, 0.0, 0.0};
      for (uint64_t c = 0; c < nchunks; ++c) {
        const PT* pc4 = part + c * stride + t;
        if constexpr (sizeof(PT) == 4) {
          const float4 v = *reinterpret_cast<const float4*>(pc4);
          y[0] += (double)v.x, y[1] += (double)v.y, y[2] += (double)v.z, y[3] += (double)v.w;
        } else {
          const double2 v0 = *reinterpret_cast<const double2*>(pc4);
          const double2 v1 = *reinterpret_cast<const double2*>(pc4 + 2);
          y[0] += v0.x, y[1] += v0.y, y[2] += v1.x, y[3] += v1.y;
        }
      }
      const double s = sc[j];
      const double beta = s * (0.5 - cf[j]);
      double gr4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = 4 * cg + q;
        const double g = 0.25 * s * y[q] + beta * colsum[l];
        const double yv = g * s;
        w_t pc[P];
        double back;
        if constexpr (F == Form::FP64) {
          pc[0] = yv;
          back = yv;
        } else if constexpr (F == Form::FP32) {
          pc[0] = (float)yv;
          back = (double)pc[0];
        } else if constexpr (F == Form::FP16) {
          pc[0] = fp16_bits(yv);
          back = fp16_value(pc[0]);
        } else {
          pc[0] = bf16_bits((float)yv);
          back = (double)bf16_value(pc[0]);
          if constexpr (P == 2) {
            pc[1] = bf16_bits((float)(yv - back));
            back += (double)bf16_value(pc[1]);
          }
        }
        gr4[q] = s != 0.0 ? back / s : 0.0;
        b4[q] += beta * gr4[q];
#pragma unroll
        for (int p = 0; p < P; ++p) pt[(p * lp + l) * PITCH + r] = pc[p];
      }
      *reinterpret_cast<double2*>(gout + t) = double2{gr4[0], gr4[1]};
      *reinterpret_cast<double2*>(gout + t + 2) = double2{gr4[2], gr4[3]};
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) bred[rg * lp + 4 * cg + q] = b4[q];
  }
  __syncthreads();
  w_t* w = static_cast<w_t*>(Wt);
  for (int v = threadIdx.x; v < P * lp * kEpiRows; v += 256) {
    const int r = v % kEpiRows, pl = v / kEpiRows, p = pl / lp, l = pl - p * lp;
    if (r < nr) w[p * wpiece + l * wpitch + r0 + r] = pt[pl * PITCH + r];
  }
  for (int l = threadIdx.x; l < lp; l += 256) {
    double sum = 0.0;
    for (int q = 0; q < RG; ++q) sum += bred[q * lp + l];
    bpart[(uint64_t)blockIdx.x * lp + l] = sum;
  }
}

template <class P>
__global__ void k_epi_h(const P* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                        uint64_t rows_pad, int lp, const double* __restrict__ b, double* __restrict__ out, int mode,
                        const double* __restrict__ colmul) {
  const uint64_t n = rows_pad * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t i = t / lp;
    const int l = (int)(t - i * lp);
    if (i < nreal) {
      double y = 0.0;
      for (uint64_t c = 0; c < nchunks; ++c) y += (double)part[c * stride + t];
      if (colmul) y *= colmul[l];
      const double v = 0.25 * y + b[l];
      out[t] = mode ? out[t] + v : v;
    } else if (!mode) {
      out[t] = 0.0;
    }
  }
}

// The same for FP32 partials, 4 adjacent elements of one row per thread:
// 16-byte partial loads, four chunks loaded ahead of their additions, each
// element's chunks still added in chunk order.  Needs lp and the chunk
// stride to be multiples of 4 and 16-byte aligned partials.
__global__ void k_epi_h_f4(const float* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                           uint64_t rows_pad, int lp, const double* __restrict__ b, double* __restrict__ out, int mode,
                           const double* __restrict__ colmul) {
  const uint64_t n4 = rows_pad * (uint64_t)lp / 4;
  GRID_STRIDE(t4, n4) {
    const uint64_t t = 4 * t4;
    const uint64_t i = t / lp;
    const int l = (int)(t - i * lp);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < nreal) {
      double y[4] = {0.0, 0.0, 0.0, 0.0};
      const float* p = part + t;
      uint64_t c = 0;
      for (; c + 4 <= nchunks; c += 4) {
        float4 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = *reinterpret_cast<const float4*>(p + (c + u) * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          y[0] += (double)a[u].x, y[1] += (double)a[u].y, y[2] += (double)a[u].z, y[3] += (double)a[u].w;
      }
      for (; c < nchunks; ++c) {
        const float4 a = *reinterpret_cast<const float4*>(p + c * stride);
        y[0] += (double)a.x, y[1] += (double)a.y, y[2] += (double)a.z, y[3] += (double)a.w;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (colmul) y[q] *= colmul[l + q];
        v[q] = 0.25 * y[q] + b[l + q];
      }
    } else if (mode) {
      continue;
    }
    double2* o = reinterpret_cast<double2*>(out + t);
    if (mode) {
      const double2 o0 = o[0], o1 = o[1];
      v[0] += o0.x, v[1] += o0.y, v[2] += o1.x, v[3] += o1.y;
    }
    o[0] = double2{v[0], v[1]};
    o[1] = double2{v[2], v[3]};
  }
}

// Column maxima of |X[r][l] rowscale[r]| over rows [0, nreal): each block
// reduces its rows in LDS and raises maxbits[l] (the bit pattern of a
// non-negative double orders like the value) with one atomic per column.
__global__ void __launch_bounds__(256) k_colmax(const double* __restrict__ X, uint64_t nreal, int lp,
                                                const double* __restrict__ rowscale,
                                                unsigned long long* __restrict__ maxbits) {
  __shared__ double m[256];
  const int per = 256 / lp;  // rows per pass
  const int t = threadIdx.x, g = t / lp, l = t - g * lp;
  double mx = 0.0;
  if (g < per) {
    for (uint64_t r = (uint64_t)blockIdx.x * per + g; r < nreal; r += (uint64_t)gridDim.x * per) {
      const double rs = rowscale ? rowscale[r] : 1.0;
      mx = fmax(mx, fabs(X[r * lp + l] * rs));
    }
  }
  m[t] = mx;
  __syncthreads();
  if (t < lp) {
    for (int q = 1; q < per; ++q) mx = fmax(mx, m[q * lp + t]);
    if (mx > 0.0) atomicMax(&maxbits[t], (unsigned long long)__double_as_longlong(mx));
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Column scales from the column maxima (1 for an all-zero column): I8
// s_l = max_l / kI8Max with epilogue factor 2 s_l; F8 s_l = 2^ceil(log2(max_l
// / 448)) (the column maximum lands in (224, 448]) with factor s_l.
__global__ void k_col_scales(const unsigned long long* __restrict__ maxbits, int lp, int f8, double* __restrict__ scale,
                             double* __restrict__ mul) {
  const int l = threadIdx.x;
  if (l >= lp) return;
  const double mx = __longlong_as_double((long long)maxbits[l]);
  double s = 1.0;
  if (mx > 0.0) {
    if (f8) {
      int e;
      const double f = frexp(mx / 448.0, &e);  // mx / 448 = f 2^e, f in [0.5, 1)
      s = ldexp(1.0, f == 0.5 ? e - 1 : e);    // smallest power of two >= mx / 448
    } else {
      s = mx / (double)kI8Max;
    }
  }
  scale[l] = s;
  mul[l] = f8 ? s : 2.0 * s;
}

// Stage 1 of a column reduction: block b sums rows [r0 + b*rpb, ...) for each column.
__global__ void k_colsum_partial(const double* __restrict__ X, uint64_t r0, uint64_t r1, int lp, uint64_t rpb,
                                 const double* __restrict__ sc, const double* __restrict__ cf,
                                 double* __restrict__ partial) {
  const uint64_t rb = r0 + (uint64_t)blockIdx.x * rpb;
  const uint64_t re = rb + rpb < r1 ? rb + rpb : r1;
  for (int l = threadIdx.x; l < lp; l += blockDim.x) {
    double s = 0.0;
    for (uint64_t r = rb; r < re; ++r) {
      const double w = sc ? sc[r] * (0.5 - cf[r]) : 1.0;
      s += w * X[r * lp + l];
    }
    partial[(uint64_t)blockIdx.x * lp + l] = s;
  }
}

// Sums of per-block partials, deterministic.  A workgroup owns 32
// consecutive outputs t of one slice of blocks: its 8 thread rows take every
// 8th block of the slice (coalesced 32-wide reads), combined in LDS in a
// fixed order.  Block b of the input sits at partial[b * bstride * width].
// With dst == partial (the first stage of a long reduction) the slice's sum
// replaces its first block; every workgroup reads its whole slice first.
__global__ void __launch_bounds__(256) k_sum_partials(double* partial, uint64_t nblocks, uint64_t width,
                                                      uint64_t bstride, uint64_t slice, double* dst) {
  __shared__ double red[8][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const uint64_t t = (uint64_t)blockIdx.x * 32 + tx;
  const uint64_t b0 = (uint64_t)blockIdx.y * slice;
  const uint64_t b1 = b0 + slice < nblocks ? b0 + slice : nblocks;
  const uint64_t step = bstride * width;
  double s0 = 0.0, s1 = 0.0;
  if (t < width) {
    uint64_t b = b0 + ty;
    for (; b + 8 < b1; b += 16) {
      s0 += partial[b * step + t];
      s1 += partial[(b + 8) * step + t];
    }
    if (b < b1) s0 += partial[b * step + t];
  }
  red[ty][tx] = s0 + s1;
  __syncthreads();
  if (ty == 0 && t < width) {
    double s = 0.0;
    for (int r = 0; r < 8; ++r) s += red[r][tx];
    if (dst == partial)
      partial[b0 * step + t] = s;
    else
      dst[t] = s;
  }
}

// out[t] = sum_b partial[b][t] for b < nblocks (partial is scratch: a long
// reduction first folds slices of 64 blocks in place).
inline void sum_partials(double* partial, uint64_t nblocks, uint64_t width, double* out, hipStream_t s) {
  const unsigned gx = (unsigned)((width + 31) / 32);
  constexpr uint64_t kSlice = 64;
  if (nblocks > 4 * kSlice) {
    const uint64_t ns = (nblocks + kSlice - 1) / kSlice;
    hipLaunchKernelGGL(k_sum_partials, dim3(gx, (unsigned)ns), dim3(256), 0, s, partial, nblocks, width, (uint64_t)1,
                       kSlice, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(gx, 1), dim3(256), 0, s, partial, ns, width, kSlice, ns, out);
    return;
  }
  hipLaunchKernelGGL(k_sum_partials, dim3(gx, 1), dim3(256), 0, s, partial, nblocks, width, (uint64_t)1, nblocks, out);
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));
constexpr int kTileRows = 32;

// Upper 16 x 16 blocks (i <= j) of an NB x NB block grid, numbered row by
// row: block t is (sym_block_i(t), sym_block_j(t)).
__host__ __device__ constexpr int sym_block_i(int t, int nb) {
  int i = 0;
  while (t >= nb - i) t -= nb - i, ++i;
  return i;
}
__host__ __device__ constexpr int sym_block_j(int t, int nb) {
  int i = 0;
  while (t >= nb - i) t -= nb - i, ++i;
  return i + t;
}
template <int NB>
constexpr int kSymPerWave = (NB * (NB + 1) / 2 + NB - 1) / NB;

// FP64-MFMA Gram partials: block b sums X[r]^T Y[r] over rows [b*rpb, ...).
// Cross products: one wave per 16-row block a of the output computes all NB
// blocks (a, :).  X^T X: the NB (NB + 1) / 2 upper blocks are dealt round
// robin to the NB waves and each is also stored mirrored (the products are
// commutative, so the result is exactly symmetric).  Row tiles of X (and Y) are
// staged in LDS.  v_mfma_f64_16x16x4: A = X^T (lane l: X[k0 + l/16][16i +
// l%16]), B = Y (lane l: Y[k0 + l/16][16j + l%16]), accumulator lane group q,
// register r = output row q + 4r.
template <int NB, bool CROSS>
__global__ void __launch_bounds__(64 * NB) k_gram_mfma(const double* __restrict__ X, const double* __restrict__ Y,
                                                       uint64_t nrows, uint64_t rpb, double* __restrict__ partial) {
  constexpr int LP = NB * 16;
  __shared__ double xs[kTileRows * LP];
  __shared__ double ys[CROSS ? kTileRows * LP : 1];
  const int tid = threadIdx.x, lane = tid & 63, a = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  f64x4_t acc[NB];
  for (int b = 0; b < NB; ++b) acc[b] = f64x4_t{0, 0, 0, 0};
  constexpr int S = kSymPerWave<NB>, NT = NB * (NB + 1) / 2;
  int ci[S], cj[S];  // this wave's upper blocks in X^T X (unused for cross products)
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int t = __builtin_amdgcn_readfirstlane(a + NB * s);
    ci[s] = t < NT ? 16 * sym_block_i(t, NB) : 0;
    cj[s] = t < NT ? 16 * sym_block_j(t, NB) : 0;
  }
  const uint64_t rb = (uint64_t)blockIdx.x * rpb;
  const uint64_t re = rb + rpb < nrows ? rb + rpb : nrows;
  // The next row tile is loaded into registers while the current one is
  // multiplied; rows past the block's end read as zero.
  constexpr int PER = kTileRows * LP / (64 * NB);
  static_assert(PER * 64 * NB == kTileRows * LP, "whole tiles per thread");
  double px[PER], py[CROSS ? PER : 1];
  auto fetch = [&](uint64_t r0) {
    const uint64_t nr = re - r0 < (uint64_t)kTileRows ? re - r0 : kTileRows;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int v = tid + e * 64 * NB;
      const bool in = (uint64_t)(v / LP) < nr;
      px[e] = in ? X[r0 * LP + v] : 0.0;
      if constexpr (CROSS) py[e] = in ? Y[r0 * LP + v] : 0.0;
    }
  };
  fetch(rb);
  for (uint64_t r0 = rb; r0 < re; r0 += kTileRows) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      xs[tid + e * 64 * NB] = px[e];
      if constexpr (CROSS) ys[tid + e * 64 * NB] = py[e];
    }
    __syncthreads();
    if (r0 + kTileRows < re) fetch(r0 + kTileRows);
    if constexpr (CROSS) {
#pragma unroll
      for (int k0 = 0; k0 < kTileRows; k0 += 4) {
        const double av = xs[(k0 + q) * LP + 16 * a + r16];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const double bv = ys[(k0 + q) * LP + 16 * b + r16];
          acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[b], 0, 0, 0);
        }
      }
    } else {
      // X^T X is symmetric: wave a takes the upper blocks (i <= j) numbered
      // a, a + NB, a + 2 NB, ... (NB (NB + 1) / 2 blocks over NB waves)
#pragma unroll
      for (int k0 = 0; k0 < kTileRows; k0 += 4) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
          if (a + NB * s < NT) {
            const double av = xs[(k0 + q) * LP + ci[s] + r16];
            const double bv = xs[(k0 + q) * LP + cj[s] + r16];
            acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[s], 0, 0, 0);
          }
        }
      }
    }
  }
  double* o = partial + (uint64_t)blockIdx.x * LP * LP;
  if constexpr (CROSS) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) o[(16 * a + q + 4 * r) * LP + 16 * b + r16] = acc[b][r];
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (a + NB * s < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          o[(ci[s] + q + 4 * r) * LP + cj[s] + r16] = acc[s][r];
          if (ci[s] != cj[s]) o[(cj[s] + r16) * LP + ci[s] + q + 4 * r] = acc[s][r];  // the mirror block
        }
      }
    }
  }
}

// FP64-MFMA right multiplication Y = X T for rows [0, nrows): one wave per 16
// rows computes all NB column blocks; A = X (lane l: X[16i + l%16][k0 + l/16]),
// B = T (lane l: T[k0 + l/16][16b + l%16]).  T is staged in LDS.
template <int NB>
__global__ void __launch_bounds__(256) k_apply_mfma(const double* __restrict__ X, const double* __restrict__ T,
                                                    uint64_t nrows, double* __restrict__ Y) {
  constexpr int LP = NB * 16;
  extern __shared__ double tl[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  for (int v = tid; v < LP * LP; v += 256) tl[v] = T[v];
  __syncthreads();
  const uint64_t nblk = (nrows + 15) / 16;
  for (uint64_t blk = (uint64_t)blockIdx.x * 4 + wave; blk < nblk; blk += (uint64_t)gridDim.x * 4) {
    const uint64_t row = blk * 16 + r16;
    const double* x = X + row * LP;
    const bool in = row < nrows;
    f64x4_t acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = f64x4_t{0, 0, 0, 0};
#pragma unroll 4
    for (int k0 = 0; k0 < LP; k0 += 4) {
      const double av = in ? x[k0 + q] : 0.0;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double bv = tl[(k0 + q) * LP + 16 * b + r16];
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[b], 0, 0, 0);
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint64_t orow = blk * 16 + q + 4 * r;
        if (orow < nrows) Y[orow * LP + 16 * b + r16] = acc[b][r];
      }
  }
}

__global__ void k_flip_partial(const double* __restrict__ O2, const double* __restrict__ Q, uint64_t nrows, int lp,
                               uint64_t rpb, double* __restrict__ partial) {
  const uint64_t rb = (uint64_t)blockIdx.x * rpb;
  const uint64_t re = rb + rpb < nrows ? rb + rpb : nrows;
  for (int l = threadIdx.x; l < lp; l += blockDim.x) {
    double a = 0.0, b = 0.0;
    for (uint64_t r = rb; r < re; ++r) {
      const double o = O2[r * lp + l], q = Q[r * lp + l];
      a += fabs(o - q);
      b += fabs(o + q);
    }
    partial[(uint64_t)blockIdx.x * 2 * lp + l] = a;
    partial[(uint64_t)blockIdx.x * 2 * lp + lp + l] = b;
  }
}

// Column-major n x L (leading dimension n) -> row-major rows_pad x lp, zero
// outside the n x L block.
__global__ void k_colmajor_to_rows(const double* __restrict__ src, uint64_t n, int L, int lp, uint64_t rows_pad,
                                   double* __restrict__ dst) {
  GRID_STRIDE(t, rows_pad * (uint64_t)lp) {
    const uint64_t r = t / lp;
    const int l = (int)(t - r * lp);
    dst[t] = (r < n && l < L) ? src[(uint64_t)l * n + r] : 0.0;
  }
}

// Row-major n x lp -> column-major n x k.
__global__ void k_rows_to_colmajor(const double* __restrict__ src, uint64_t n, int lp, int k, double* __restrict__ dst) {
  GRID_STRIDE(t, n * (uint64_t)k) {
    const uint64_t j = t / n, r = t - j * n;
    dst[t] = src[r * lp + j];
  }
}

// Zero bytes [b0, b0 + n0) and [b1, b1 + n1) of each of `rows` rows (pitch bytes).
__global__ void k_zero_ranges(unsigned char* __restrict__ base, uint64_t pitch, uint64_t rows, uint64_t b0,
                              uint64_t n0, uint64_t b1, uint64_t n1) {
  const uint64_t per = n0 + n1;
  GRID_STRIDE(t, rows * per) {
    const uint64_t r = t / per, c = t - r * per;
    base[r * pitch + (c < n0 ? b0 + c : b1 + (c - n0))] = 0;
  }
}

// flip_UV with V as the base (src/Utils.cpp:118-154): per column c < k of X
// (rows [0, n), pitch lp) the value at its first maximum of |x|.  Stage 1:
// block b scans rows [b rpb, ...) in order; stage 2 takes the blocks in
// order, so the first maximum wins ties as Eigen's maxCoeff does.
__global__ void k_col_absmax_partial(const double* __restrict__ X, uint64_t n, int lp, int k, uint64_t rpb,
                                     double* __restrict__ pval) {
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
  const uint64_t r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    double best = -1.0, v = 0.0;
    for (uint64_t r = r0; r < r1; ++r) {
      const double x = X[r * lp + c];
      if (fabs(x) > best) best = fabs(x), v = x;
    }
    pval[(uint64_t)blockIdx.x * lp + c] = v;
  }
}

__global__ void k_col_absmax_sign(const double* __restrict__ pval, uint64_t nb, int lp, int k,
                                  double* __restrict__ sign) {
  const int c = threadIdx.x;
  if (c >= lp) return;
  double s = 1.0;
  if (c < k) {
    double best = -1.0, v = 0.0;
    for (uint64_t b = 0; b < nb; ++b) {
      const double x = pval[b * lp + c];
      if (fabs(x) > best) best = fabs(x), v = x;
    }
    s = v < 0.0 ? -1.0 : 1.0;
  }
  sign[c] = s;
}

__global__ void k_scale_cols(double* __restrict__ X, uint64_t n, int lp, const double* __restrict__ sign) {
  GRID_STRIDE(t, n * (uint64_t)lp) X[t] *= sign[t % lp];
}

// dst (rows x kp) = the first k columns of src (rows x lp), times scale[t]
// (null = 1), zero padded.
__global__ void k_impute_rows(const double* __restrict__ src, uint64_t rows, int lp, int k, int kp,
                              const double* __restrict__ scale, double* __restrict__ dst) {
  GRID_STRIDE(t, rows * (uint64_t)kp) {
    const uint64_t r = t / kp;
    const int c = (int)(t - r * kp);
    dst[t] = c < k ? src[r * lp + c] * (scale ? scale[c] : 1.0) : 0.0;
  }
}

// Per-SNP column state as read: cF = F and sc = 1 for j < m, zero padding to rows.
__global__ void k_pristine_cols(const double* __restrict__ F, uint64_t m, uint64_t rows, double* __restrict__ cf,
                                double* __restrict__ sc) {
  GRID_STRIDE(j, rows) {
    cf[j] = j < m ? F[j] : 0.0;
    sc[j] = j < m ? 1.0 : 0.0;
  }
}

// standardize_E's scale (src/Data.cpp:351-362): sc_j *= sqrt(ploidy) / sd_j with
// sd_j = sqrt(F_j (1 - F_j)) when sd_j > tol.
__global__ void k_standardize(const double* __restrict__ F, uint64_t m, double sqrt_ploidy, double tol,
                              double* __restrict__ sc) {
  GRID_STRIDE(j, m) {
    const double f = F[j];
    const double sd = sqrt(f * (1.0 - f));
    if (sd > tol) sc[j] *= sqrt_ploidy / sd;
  }
}

__global__ void k_gather_f64(const double* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                             double* __restrict__ dst) {
  GRID_STRIDE(i, n) dst[i] = src[idx[i]];
}

__global__ void k_add(const double* __restrict__ X, const double* __restrict__ Z, double* __restrict__ Y, uint64_t n) {
  GRID_STRIDE(t, n) Y[t] = X[t] + Z[t];
}

// First list position in [lo, hi) whose index is >= key (lists are sorted).
__device__ __forceinline__ uint64_t lower_bound_ix(const uint32_t* ix, uint64_t lo, uint64_t hi, uint64_t key) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (ix[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Entry range of row r restricted to indices [klo, khi).
__device__ __forceinline__ void entry_range(const uint64_t* rp, const uint32_t* ix, uint64_t r, uint64_t klo,
                                            uint64_t khi, uint64_t& e0, uint64_t& e1) {
  e0 = rp[r];
  e1 = rp[r + 1];
  if (klo > 0) e0 = lower_bound_ix(ix, e0, e1, klo);
  e1 = lower_bound_ix(ix, e0, e1, khi);
}

// Per-entry weights of the missing-entry corrections without imputation:
// w = -sc_c (1/2 - cF_c), which cancels the centering term the affine
// epilogue gave the entry.  One wave per row of the lists and index chunk
// (grid.y), lanes over the row's entries.
__global__ void k_missing_weights(MissWeightArgs a) {
  const uint64_t r = blockIdx.x * (uint64_t)(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= a.nrows) return;
  const uint64_t klo = (uint64_t)blockIdx.y * a.kspan, khi = klo + a.kspan;
  uint64_t e0, e1;
  entry_range(a.rp, a.ix, r, klo, khi, e0, e1);
  for (uint64_t e = e0 + lane; e < e1; e += 64) {
    const uint64_t c = a.by_snp ? r : a.ix[e];  // SNP column
    a.w[e] = -a.sc[c] * (0.5 - a.cf[c]);
  }
}

// EMU weights (src/Data.cpp:318-348): w = -sc_c (1/2 - cF_c) + sc_c clamp(p, -F_c,
// 1 - F_c) with p = (U S)_i . V_c.  One wave per row of the lists (a SNP c for
// the by-SNP list, a sample i for the by-sample list) and index chunk
// (grid.y; chunks keep the gathered rows within L2).  The k-vectors are rows
// of pitch kp (whole 128-byte lines) read as 16-byte pairs, and a V row also
// carries its SNP's sc, cF and F, so an entry of the by-sample list touches
// one line.  The row's own vector sits in registers; the wave is split into
// 64 / LPE groups of LPE lanes (a power of two >= k / 2), each group taking
// one entry per step, reading the entry's vector as one contiguous access
// and reducing the dot product by butterfly shuffles.  The entries' indexes
// come 64 at a time (one coalesced load, the next batch's fetched during this
// one).
__global__ void __launch_bounds__(256) k_missing_weights_emu(MissWeightArgs a, int lpe) {
  const uint64_t r = blockIdx.x * 4ull + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= a.nrows) return;  // wave-uniform
  const uint64_t klo = (uint64_t)blockIdx.y * a.kspan, khi = klo + a.kspan;
  uint64_t e0, e1;
  entry_range(a.rp, a.ix, r, klo, khi, e0, e1);
  const int kk = a.kk, np = a.kp / 2;  // 16-byte pairs per row
  const int nk = (kk + 1) / 2;          // pairs holding loadings
  const int G = 64 / lpe, g = lane / lpe, cl = lane - g * lpe;
  const bool has = cl < nk;
  const double* ownrow = (a.by_snp ? a.V : a.US) + r * (uint64_t)a.kp;
  const double2* other = reinterpret_cast<const double2*>(a.by_snp ? a.US : a.V);
  // the row's own loadings (a zero past k in the last pair)
  double2 rv = double2{0.0, 0.0};
  if (has) {
    rv.x = ownrow[2 * cl];
    rv.y = 2 * cl + 1 < kk ? ownrow[2 * cl + 1] : 0.0;
  }
  double sc_r = 0.0, cf_r = 0.0, f_r = 0.0;  // the SNP's state for the by-SNP list
  if (a.by_snp) {
    sc_r = a.sc[r];
    cf_r = a.cf[r];
    f_r = a.Fclamp[r];
  }
  constexpr int U = 4;
  uint32_t nix = e0 + lane < e1 ? a.ix[e0 + lane] : 0u;
  for (uint64_t eb = e0; eb < e1; eb += 64) {
    const int nb = (int)min<uint64_t>(64, e1 - eb);
    const uint32_t myix = nix;
    nix = eb + 64 + lane < e1 ? a.ix[eb + 64 + lane] : 0u;
    const int steps = (nb + G - 1) / G;
    for (int st = 0; st < steps; st += U) {
      uint32_t kx[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = (st + u) * G + g;
        kx[u] = __shfl(myix, j < 63 ? j : 63);
        ok[u] = j < nb;
      }
      double2 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        x[u] = has ? other[(ok[u] ? (uint64_t)kx[u] : 0) * np + cl] : double2{0.0, 0.0};
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double p = rv.x * x[u].x + rv.y * x[u].y;
        for (int off = lpe / 2; off > 0; off >>= 1) p += __shfl_xor(p, off);
        const int j = (st + u) * G + g;
        if (cl == 0 && ok[u]) {
          double sc = sc_r, cf = cf_r, f = f_r;
          if (!a.by_snp) {  // the entry's SNP row carries its state after the loadings
            const double* vr = a.V + (uint64_t)kx[u] * a.kp + kk;
            sc = vr[0];
            cf = vr[1];
            f = vr[2];
          }
          double w = -sc * (0.5 - cf);
          w += sc * fmin(fmax(p, -f), 1.0 - f);
          a.w[eb + j] = w;
        }
      }
    }
  }
}

__global__ void k_pack_col_state(double* __restrict__ V, uint64_t m, int kp, int kk, const double* __restrict__ sc,
                                 const double* __restrict__ cf, const double* __restrict__ F) {
  GRID_STRIDE(j, m) {
    double* vr = V + j * (uint64_t)kp + kk;
    vr[0] = sc[j];
    vr[1] = cf[j];
    vr[2] = F[j];
  }
}

// out (rows x pitch, T) = the first L columns of src (rows x lp, FP64), zero padded.
template <class T>
__global__ void k_compact_rows(const double* __restrict__ src, uint64_t rows, int lp, int L, int pitch,
                               T* __restrict__ out) {
  GRID_STRIDE(t, rows * (uint64_t)pitch) {
    const uint64_t r = t / pitch;
    const int c = (int)(t - r * pitch);
    out[t] = c < L ? (T)src[r * lp + c] : (T)0;
  }
}

// Sparse correction out[r][c] (+)= sum_e w[e] src[ix[e] - src_k0][c], c < ncols,
// one wave per (target row, index chunk = grid.z).  The row's entries are
// taken 64 at a time (one coalesced load of their indexes and weights, the
// next batch's fetched during this one).  The wave is split into G groups of
// LPE = cp / VEC lanes, each lane holding VEC columns (16 bytes): per step
// every group takes one entry, receives its index and weight by lane permutes
// and reads the entry's compact source row as one contiguous access (16
// bytes per lane), accumulating in FP64.  The groups' sums are combined in a
// fixed order at the end.  With several index chunks each writes its own
// partial row (summed later in chunk order), else it adds to out.
template <class T>
__global__ void __launch_bounds__(256) k_missing_spmm(MissSpmmArgs a) {
  constexpr int VEC = 16 / sizeof(T);
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  const uint64_t r = blockIdx.x * 4ull + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= a.nrows) return;  // wave-uniform
  const int lpe = a.cp / VEC;  // lanes per entry (cp is a multiple of VEC)
  const int G = 64 / lpe;      // entries per step
  const uint64_t rpv = a.cpitch / VEC;  // 16-byte vectors per source row (whole cache lines)
  const int g = lane / lpe, cl = lane - g * lpe;
  const bool act = g < G;
  const uint64_t klo = a.k_lo + (uint64_t)blockIdx.z * a.kspan;
  const uint64_t khi = min(a.k_hi, klo + a.kspan);
  uint64_t e0, e1;
  entry_range(a.rp, a.ix, r, klo, khi, e0, e1);
  double acc[VEC];
#pragma unroll
  for (int q = 0; q < VEC; ++q) acc[q] = 0.0;
  const V* src = static_cast<const V*>(a.src);
  constexpr int U = 4;  // steps per pass: their gathers are issued together
  uint32_t nix = 0;
  double nw = 0.0;
  if (e0 + lane < e1) {
    nix = a.ix[e0 + lane];
    nw = a.w[e0 + lane];
  }
  for (uint64_t eb = e0; eb < e1; eb += 64) {
    const int nb = (int)min<uint64_t>(64, e1 - eb);
    const uint32_t myix = nix;
    const double myw = nw;
    if (eb + 64 + lane < e1) {
      nix = a.ix[eb + 64 + lane];
      nw = a.w[eb + 64 + lane];
    }
    const int steps = (nb + G - 1) / G;
    // an entry past the batch (or an idle lane) reads source row 0 and is
    // left out of the sum
    for (int st = 0; st < steps; st += U) {
      uint32_t kk[U];
      double ww[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = (st + u) * G + g;
        const int jl = j < 63 ? j : 63;  // every lane takes part in the permutes
        kk[u] = __shfl(myix, jl);
        ww[u] = __shfl(myw, jl);
        ok[u] = act && j < nb;
      }
      V xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = src[(ok[u] ? (uint64_t)kk[u] - a.src_k0 : 0) * rpv + cl];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const T* x = reinterpret_cast<const T*>(&xv[u]);
        if (ok[u]) {
#pragma unroll
          for (int q = 0; q < VEC; ++q) acc[q] += ww[u] * (double)x[q];
        }
      }
    }
  }
  // group 0 lane cl += the same lane of groups 1 .. G-1, in order
#pragma unroll
  for (int q = 0; q < VEC; ++q) {
    const double v = acc[q];
    double tot = v;
    for (int gg = 1; gg < G; ++gg) {
      const int sl = cl + gg * lpe;
      tot += __shfl(v, sl < 63 ? sl : 63);
    }
    acc[q] = tot;
  }
  if (g != 0) return;
  const bool partial = gridDim.z > 1;
  double* o = partial ? a.partial + ((uint64_t)blockIdx.z * a.nrows + r) * a.ncols : a.out + r * a.lp;
#pragma unroll
  for (int q = 0; q < VEC; ++q) {
    const int c = cl * VEC + q;
    if (c < a.ncols) o[c] = partial ? acc[q] : o[c] + acc[q];
  }
}

// ---- L x L steps of the thin-Q of H on the device (one workgroup each) ----
// Register-blocked: 256 threads form a 16 x 16 grid and thread (ty, tx) holds
// the elements (ty + 16 p, tx + 16 q), p, q < NB = lp / 16, of each L x L
// matrix in registers.  The elimination steps exchange one row and one
// column per step through double-buffered LDS vectors, one barrier per step.
// The elimination order is the host routines' (linalg_host.cpp): right-
// looking Cholesky, back substitution for R^-1, modified LU for the
// Householder signs (Ballard et al. 2014); divisions by a step's pivot are
// multiplications by its reciprocal.
constexpr int kSmallThreads = 256;
constexpr int kThinqMaxNB = 5;  // lp <= 80

template <int NB>
struct RegMat {
  double v[NB][NB];
};

template <int NB>
__device__ __forceinline__ void rm_zero(RegMat<NB>& m) {
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q) m.v[p][q] = 0.0;
}

// Rows and columns >= n read as zero.
template <int NB>
__device__ __forceinline__ void rm_load(RegMat<NB>& m, const double* src, int n, int lp, int ty, int tx) {
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int i = ty + 16 * p, k = tx + 16 * q;
      m.v[p][q] = (i < n && k < n) ? src[i * lp + k] : 0.0;
    }
}

template <int NB>
__device__ __forceinline__ void rm_store(const RegMat<NB>& m, double* dst, int lp, int ty, int tx) {
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q) dst[(ty + 16 * p) * lp + tx + 16 * q] = m.v[p][q];
}

// In place: A (n x n, symmetric positive definite, upper triangle used) ->
// upper R with A = R^T R; rows and columns >= n and the lower triangle end
// zero.  Returns false (in every thread) on breakdown.  rowbuf: 2 lp.
template <int NB>
__device__ bool rb_chol_upper(RegMat<NB>& a, int n, int lp, double* rowbuf, int ty, int tx,
                              double* min_pivot = nullptr) {
  double dmin = INFINITY;
  if (ty == 0) {
#pragma unroll
    for (int q = 0; q < NB; ++q) rowbuf[tx + 16 * q] = a.v[0][q];
  }
  __syncthreads();
  bool ok = true;
  for (int j = 0; j < n; ++j) {
    const double* rb = rowbuf + (j & 1) * lp;
    const double d = rb[j];
    if (!(d > 0.0) || !isfinite(d)) {  // the same value in every thread
      ok = false;
      break;
    }
    dmin = fmin(dmin, d);
    const double rjj = sqrt(d), inv = 1.0 / rjj;
    double si[NB], sk[NB];
#pragma unroll
    for (int p = 0; p < NB; ++p) si[p] = rb[ty + 16 * p] * inv;
#pragma unroll
    for (int q = 0; q < NB; ++q) sk[q] = rb[tx + 16 * q] * inv;
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int i = ty + 16 * p;
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int k = tx + 16 * q;
        if (i > j && k >= i && k < n) a.v[p][q] -= si[p] * sk[q];
        else if (i == j) a.v[p][q] = k > j ? sk[q] : (k == j ? rjj : 0.0);  // row j of R
      }
    }
    if (j + 1 < n && ty == ((j + 1) & 15)) {  // publish row j + 1 for the next step
      const int p1 = (j + 1) >> 4;
#pragma unroll
      for (int p = 0; p < NB; ++p)
        if (p == p1) {
#pragma unroll
          for (int q = 0; q < NB; ++q) rowbuf[((j + 1) & 1) * lp + tx + 16 * q] = a.v[p][q];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int i = ty + 16 * p, k = tx + 16 * q;
      if (!ok || k < i || i >= n || k >= n) a.v[p][q] = 0.0;
    }
  if (min_pivot) *min_pivot = dmin;
  return ok;
}

// X = R^-1 for upper R (n x n, zero outside), by rows from the bottom:
// X[i][c] = (delta_ic - sum_{k>i} R[i][k] X[k][c]) / R[i][i], the sums
// accumulated as each row of X is published.  buf: 5 lp (diag, 2 x row, 2 x column).
template <int NB>
__device__ void rb_inv_upper(const RegMat<NB>& r, RegMat<NB>& x, int n, int lp, double* buf, int ty, int tx) {
  double* diag = buf;
  double* xrow = buf + lp;
  double* rcol = buf + 3 * lp;
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (ty + 16 * p == tx + 16 * q) diag[ty + 16 * p] = r.v[p][q];
  rm_zero(x);
  __syncthreads();
  for (int i = n - 1; i >= 0; --i) {
    const int b = i & 1;
    if (ty == (i & 15)) {
      const int pi = i >> 4;
      const double inv = 1.0 / diag[i];
#pragma unroll
      for (int p = 0; p < NB; ++p)
        if (p == pi) {
#pragma unroll
          for (int q = 0; q < NB; ++q) {
            const int c = tx + 16 * q;
            const double v = (c >= i && c < n) ? ((c == i ? 1.0 : 0.0) - x.v[p][q]) * inv : 0.0;
            x.v[p][q] = v;
            xrow[b * lp + c] = v;
          }
        }
    }
    if (tx == (i & 15)) {
      const int qi = i >> 4;
#pragma unroll
      for (int q = 0; q < NB; ++q)
        if (q == qi) {
#pragma unroll
          for (int p = 0; p < NB; ++p) rcol[b * lp + ty + 16 * p] = r.v[p][q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      if (ty + 16 * p < i) {
        const double rv = rcol[b * lp + ty + 16 * p];
#pragma unroll
        for (int q = 0; q < NB; ++q) x.v[p][q] += rv * xrow[b * lp + tx + 16 * q];
      }
    }
  }
  __syncthreads();  // the buffers are free again
}

// C = A B over k < n (A, B in LDS, pitch lp, zero outside n x n).
template <int NB>
__device__ void rb_matmul(const double* A, const double* B, RegMat<NB>& c, int n, int lp, int ty, int tx) {
  rm_zero(c);
  for (int k = 0; k < n; ++k) {
    double av[NB], bv[NB];
#pragma unroll
    for (int p = 0; p < NB; ++p) av[p] = A[(ty + 16 * p) * lp + k];
#pragma unroll
    for (int q = 0; q < NB; ++q) bv[q] = B[k * lp + tx + 16 * q];
#pragma unroll
    for (int p = 0; p < NB; ++p)
#pragma unroll
      for (int q = 0; q < NB; ++q) c.v[p][q] += av[p] * bv[q];
  }
}

// Householder diagonal signs of the thin QR whose Q has top block W (n x n,
// destroyed): modified LU, s_i = -sign(u_ii).  buf: 4 lp.
template <int NB>
__device__ void rb_householder_signs(RegMat<NB>& w, int n, int lp, double* buf, double* sgn, int ty, int tx) {
  double* rowb = buf;
  double* colb = buf + 2 * lp;
  for (int i = 0; i < n; ++i) {
    const int b = i & 1;
    if (ty == (i & 15) && tx == (i & 15)) {  // the pivot: w[i][i] -= s_i
      const int pi = i >> 4;
#pragma unroll
      for (int p = 0; p < NB; ++p)
        if (p == pi) {
          const double si = w.v[p][p] >= 0.0 ? -1.0 : 1.0;
          w.v[p][p] -= si;
          sgn[i] = si;
        }
    }
    if (ty == (i & 15)) {
      const int pi = i >> 4;
#pragma unroll
      for (int p = 0; p < NB; ++p)
        if (p == pi) {
#pragma unroll
          for (int q = 0; q < NB; ++q) rowb[b * lp + tx + 16 * q] = w.v[p][q];
        }
    }
    if (tx == (i & 15)) {
      const int qi = i >> 4;
#pragma unroll
      for (int q = 0; q < NB; ++q)
        if (q == qi) {
#pragma unroll
          for (int p = 0; p < NB; ++p) colb[b * lp + ty + 16 * p] = w.v[p][q];
        }
    }
    __syncthreads();
    const double inv = 1.0 / rowb[b * lp + i];  // the pivot
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int r = ty + 16 * p;
      if (r > i && r < n) {
        const double l = colb[b * lp + r] * inv;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int c = tx + 16 * q;
          if (c > i && c < n) w.v[p][q] -= l * rowb[b * lp + c];
        }
      }
    }
  }
  __syncthreads();
}

// First pass: R1 = chol(G1), T1 = R1^-1 (lp x lp, zero padded).
template <int NB>
__global__ void __launch_bounds__(kSmallThreads) k_thinq_first(const double* __restrict__ G1, int n, int lp,
                                                               double* __restrict__ R1, double* __restrict__ T1,
                                                               int* __restrict__ status, double min_ratio) {
  __shared__ double buf[5 * 16 * kThinqMaxNB];
  __shared__ double dmax_s;
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  RegMat<NB> a, x;
  rm_load(a, G1, n, lp, ty, tx);
  if (min_ratio > 0.0) {
    // largest diagonal entry of the Gram (before the factorization overwrites it)
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int i = 0; i < n; ++i) m = fmax(m, G1[i * lp + i]);
      dmax_s = m;
    }
  }
  double dmin = 0.0;
  const bool ok = rb_chol_upper(a, n, lp, buf, ty, tx, &dmin);
  if (ok) rb_inv_upper(a, x, n, lp, buf, ty, tx);
  else rm_zero(x);
  if (!ok && threadIdx.x == 0) *status = 1;
  // the host's safety test of this Gram (smallest pivot / largest diagonal),
  // when the caller defers it to the device: a near rank-deficient H is marked
  if (ok && min_ratio > 0.0 && threadIdx.x == 0 && !(dmin > min_ratio * dmax_s)) *status = 2;
  rm_store(a, R1, lp, ty, tx);
  rm_store(x, T1, lp, ty, tx);
}

// Second pass: R2 = chol(G2), R^-1 = (R2 R1)^-1, the Householder signs s of
// Xtop R^-1 (Xtop = the first n rows of the thinned matrix) and
// T = R^-1 diag(s) (lp x lp, zero padded).  Dynamic LDS: 2 lp x lp matrices.
template <int NB>
__global__ void __launch_bounds__(kSmallThreads) k_thinq_second(const double* __restrict__ G2,
                                                                const double* __restrict__ R1,
                                                                const double* __restrict__ Xtop, int n, int lp,
                                                                double* __restrict__ T, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char sm[];
  __shared__ double buf[5 * 16 * kThinqMaxNB];
  __shared__ double sgn[16 * kThinqMaxNB];
  double* MA = reinterpret_cast<double*>(sm);
  double* MB = MA + lp * lp;
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  RegMat<NB> a, x;
  rm_load(a, G2, n, lp, ty, tx);
  const bool ok = rb_chol_upper(a, n, lp, buf, ty, tx);  // R2
  if (ok) {
    RegMat<NB> t;
    rm_store(a, MA, lp, ty, tx);
    rm_load(t, R1, n, lp, ty, tx);
    rm_store(t, MB, lp, ty, tx);
    __syncthreads();
    rb_matmul(MA, MB, a, n, lp, ty, tx);  // R = R2 R1
    rb_inv_upper(a, x, n, lp, buf, ty, tx);  // R^-1 (ends with a barrier: MA, MB are free)
    rm_store(x, MB, lp, ty, tx);
    rm_load(t, Xtop, n, lp, ty, tx);
    rm_store(t, MA, lp, ty, tx);
    __syncthreads();
    rb_matmul(MA, MB, t, n, lp, ty, tx);  // Xtop R^-1
    rb_householder_signs(t, n, lp, buf, sgn, ty, tx);
  } else {
    rm_zero(x);
    if (threadIdx.x < 16 * NB) sgn[threadIdx.x] = 1.0;
    if (threadIdx.x == 0) *status = 1;
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < NB; ++p)
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int c = tx + 16 * q;
      T[(ty + 16 * p) * lp + c] = c < n ? x.v[p][q] * sgn[c] : 0.0;
    }
}

// flipOmg's decision (src/RSVD.hpp:80-89): column l flips when
// sum |Omg2 - Q| > 2 sum |Omg2 + Q| (ab = the two norms, 2 x lp).
__global__ void k_flip_signs(const double* __restrict__ ab, int n, int lp, double* __restrict__ sign) {
  const int l = threadIdx.x;
  if (l < lp) sign[l] = (l < n && ab[l] > 2.0 * ab[lp + l]) ? -1.0 : 1.0;
}

// out[r][l] += sum_z partial[z][r][l] (z in order), l < ncols.
__global__ void k_sum_chunks(const double* __restrict__ partial, uint64_t nz, uint64_t nrows, int ncols,
                             double* __restrict__ out, int lp) {
  GRID_STRIDE(t, nrows * (uint64_t)ncols) {
    const uint64_t r = t / ncols;
    const int l = (int)(t - r * ncols);
    double v = 0.0;
    for (uint64_t z = 0; z < nz; ++z) v += partial[z * nrows * ncols + t];
    out[r * lp + l] += v;
  }
}

}  // namespace

void launch_colmajor_to_rows(const double* src, uint64_t n, int L, int lp, uint64_t rows_pad, double* dst,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_colmajor_to_rows, dim3(grid_for(rows_pad * lp)), dim3(256), 0, s, src, n, L, lp, rows_pad, dst);
  check_launch("colmajor_to_rows");
}

void launch_zero_ranges(void* base, uint64_t pitch, uint64_t rows, uint64_t b0, uint64_t n0, uint64_t b1, uint64_t n1,
                        hipStream_t s) {
  if (!rows || n0 + n1 == 0) return;
  hipLaunchKernelGGL(k_zero_ranges, dim3(grid_for(rows * (n0 + n1))), dim3(256), 0, s,
                     static_cast<unsigned char*>(base), pitch, rows, b0, n0, b1, n1);
  check_launch("zero_ranges");
}

void launch_flip_cols_by_max(double* V, uint64_t m, double* U, uint64_t n, int lp, int k, double* partial,
                             double* sign, hipStream_t s) {
  const uint64_t nb = std::min<uint64_t>(2048, std::max<uint64_t>(1, m));
  const uint64_t rpb = (m + nb - 1) / nb;
  const uint64_t nbu = (m + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_col_absmax_partial, dim3((unsigned)nbu), dim3(64), 0, s, V, m, lp, k, rpb, partial);
  hipLaunchKernelGGL(k_col_absmax_sign, dim3(1), dim3(128), 0, s, partial, nbu, lp, k, sign);
  hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(m * lp)), dim3(256), 0, s, V, m, lp, sign);
  hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(n * lp)), dim3(256), 0, s, U, n, lp, sign);
  check_launch("flip_cols_by_max");
}

void launch_impute_rows(const double* src, uint64_t rows, int lp, int k, int kp, const double* scale, double* dst,
                        hipStream_t s) {
  if (!rows) return;
  hipLaunchKernelGGL(k_impute_rows, dim3(grid_for(rows * (uint64_t)kp)), dim3(256), 0, s, src, rows, lp, k, kp, scale,
                     dst);
  check_launch("impute_rows");
}

void launch_pristine_cols(const double* F, uint64_t m, uint64_t rows, double* cf, double* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_pristine_cols, dim3(grid_for(rows)), dim3(256), 0, s, F, m, rows, cf, sc);
  check_launch("pristine_cols");
}

void launch_standardize(const double* F, uint64_t m, double sqrt_ploidy, double tol, double* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_standardize, dim3(grid_for(m)), dim3(256), 0, s, F, m, sqrt_ploidy, tol, sc);
  check_launch("standardize");
}

void launch_gather_f64(const double* src, const uint32_t* idx, uint64_t n, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_f64, dim3(grid_for(n)), dim3(256), 0, s, src, idx, n, dst);
  check_launch("gather_f64");
}

void launch_rows_to_colmajor(const double* src, uint64_t n, int lp, int k, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_rows_to_colmajor, dim3(grid_for(n * k)), dim3(256), 0, s, src, n, lp, k, dst);
  check_launch("rows_to_colmajor");
}

template <Form F>
static void make_pieces_t(const double* X, uint64_t nreal, uint64_t rows_pad, int lp, const double* rowscale,
                          const double* colsign, void* Wt, uint64_t wpitch, uint64_t wpiece, double* rounded,
                          double* signed_out, const double* colscale, const double* bsc, const double* bcf,
                          double* bpart, hipStream_t s) {
  using w_t = typename Pieces<F>::w_t;
  constexpr int P = Pieces<F>::P;
  const auto lds_for = [](int l) {
    return ((size_t)epi_pitch(sizeof(w_t)) * l * P * sizeof(w_t) + 15) / 16 * 16 + (size_t)1024 * 8;
  };
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_make_pieces<F>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_for(128)));
    attr = true;
  }
  const uint64_t nb = (rows_pad + kEpiRows - 1) / kEpiRows;
  hipLaunchKernelGGL(k_make_pieces<F>, dim3((unsigned)nb), dim3(256), lds_for(lp), s, X, nreal, rows_pad, lp, rowscale,
                     colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bpart);
}

void launch_make_pieces(Form f, const double* X, uint64_t nreal, uint64_t rows_pad, int lp, const double* rowscale,
                        const double* colsign, void* Wt, uint64_t wpitch, uint64_t wpiece, double* rounded,
                        double* signed_out, hipStream_t s, const double* colscale, const double* bsc,
                        const double* bcf, double* bpart, uint64_t bpart_cap, double* bvec) {
  require(!form_colscaled(f) || colscale, kErrInternal, "make_pieces: this form needs column scales");
  require(lp >= 16 && lp <= 128 && lp % 4 == 0, kErrArgument, "make_pieces: bad lp");
  if (!rows_pad) return;
  const uint64_t nb = (rows_pad + kEpiRows - 1) / kEpiRows;
  // the column sums ride along when their per-tile partials fit the scratch
  const bool fuse = bvec && bpart && nb * (uint64_t)lp <= bpart_cap;
  double* bp = fuse ? bpart : nullptr;
  switch (f) {
    case Form::FP64: make_pieces_t<Form::FP64>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::FP32: make_pieces_t<Form::FP32>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::BF16: make_pieces_t<Form::BF16>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::BF16X1: make_pieces_t<Form::BF16X1>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::FP16: make_pieces_t<Form::FP16>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::I8: make_pieces_t<Form::I8>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
    case Form::F8: make_pieces_t<Form::F8>(X, nreal, rows_pad, lp, rowscale, colsign, Wt, wpitch, wpiece, rounded, signed_out, colscale, bsc, bcf, bp, s); break;
  }
  check_launch("make_pieces");
  if (!bvec) return;
  if (fuse) {
    sum_partials(bpart, nb, (uint64_t)lp, bvec, s);
    check_launch("make_pieces_colsum");
  } else {
    require(rounded != nullptr && bpart != nullptr, kErrInternal, "make_pieces: column sums need rounded and scratch");
    launch_colsum(rounded, 0, nreal, lp, bsc, bcf, bpart, bvec, s);
  }
}

void launch_epi_g(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* sc, const double* cf, const double* colsum, double* out, hipStream_t s,
                  const double* colmul) {
  const unsigned g = grid_for(rows_pad * lp);
  require(!form_colscaled(f) || colmul, kErrInternal, "epi_g: this form needs column factors");
  if (f == Form::FP64)
    hipLaunchKernelGGL(k_epi_g<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, sc, cf, colsum, out, colmul);
  else if (lp % 4 == 0 && chunk_stride % 4 == 0 && reinterpret_cast<uintptr_t>(P) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(out) % 16 == 0) {
    const unsigned g4 = grid_for(rows_pad * lp / 4);
    if (f == Form::I8)
      hipLaunchKernelGGL(k_epi_g_v4<int32_t>, dim3(g4), dim3(256), 0, s, static_cast<const int32_t*>(P), nchunks,
                         chunk_stride, nreal, rows_pad, lp, sc, cf, colsum, out, colmul);
    else
      hipLaunchKernelGGL(k_epi_g_v4<float>, dim3(g4), dim3(256), 0, s, static_cast<const float*>(P), nchunks,
                         chunk_stride, nreal, rows_pad, lp, sc, cf, colsum, out, colmul);
  } else if (f == Form::I8)
    hipLaunchKernelGGL(k_epi_g<int32_t>, dim3(g), dim3(256), 0, s, static_cast<const int32_t*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, sc, cf, colsum, out, colmul);
  else
    hipLaunchKernelGGL(k_epi_g<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(P), nchunks, chunk_stride,
                       nreal, rows_pad, lp, sc, cf, colsum, out, colmul);
  check_launch("epi_g");
}

void launch_col_scales(Form f, const double* X, uint64_t nreal, int lp, const double* rowscale,
                       unsigned long long* maxbits, double* scale, double* mul, hipStream_t s) {
  require(form_colscaled(f), kErrInternal, "col_scales: not a column-scaled form");
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, s, maxbits, lp);
  const uint64_t per = 256 / lp;
  hipLaunchKernelGGL(k_colmax, dim3(grid_for(nreal, (unsigned)per, 2048)), dim3(256), 0, s, X, nreal, lp, rowscale,
                     maxbits);
  hipLaunchKernelGGL(k_col_scales, dim3(1), dim3(128), 0, s, maxbits, lp, f == Form::F8 ? 1 : 0, scale, mul);
  check_launch("col_scales");
}

uint64_t epi_g_fused_partial_elems(uint64_t nwin, int lp) { return (nwin + kEpiRows - 1) / kEpiRows * lp; }

template <Form F, class PT>
static void epi_fused_t(const void* P, uint64_t nchunks, uint64_t stride, uint64_t nwin, int lp, const double* sc,
                        const double* cf, const double* colsum, double* gout, void* Wt, uint64_t wpitch,
                        uint64_t wpiece, double* bpart, hipStream_t s) {
  constexpr int NP = F == Form::BF16 ? 2 : 1;
  constexpr size_t wsz = F == Form::FP64 ? 8 : (F == Form::FP32 ? 4 : 2);
  // pieces (transposed) + the row groups' beta sums (256 / (lp / 4) groups x lp)
  const auto lds_for = [](int l) { return ((size_t)epi_pitch(wsz) * l * NP * wsz + 15) / 16 * 16 + (size_t)1024 * 8; };
  const size_t lds = lds_for(lp);
  static bool attr = false;
  if (!attr) {
    const size_t lds_max = lds_for(128);
    PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_epi_g_fused<F, PT>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
    attr = true;
  }
  const uint64_t nb = (nwin + kEpiRows - 1) / kEpiRows;
  hipLaunchKernelGGL((k_epi_g_fused<F, PT>), dim3((unsigned)nb), dim3(256), lds, s, static_cast<const PT*>(P), nchunks,
                     stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart);
}

void launch_epi_g_fused(Form f, const void* P, uint64_t nchunks, uint64_t stride, uint64_t nwin, int lp,
                        const double* sc, const double* cf, const double* colsum, double* gout, void* Wt,
                        uint64_t wpitch, uint64_t wpiece, double* bpart, double* bvec, hipStream_t s) {
  if (!nwin) return;
  require(lp >= 16 && lp <= 128, kErrArgument, "epi_g_fused: bad lp");
  switch (f) {
    case Form::FP64: epi_fused_t<Form::FP64, double>(P, nchunks, stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart, s); break;
    case Form::FP32: epi_fused_t<Form::FP32, float>(P, nchunks, stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart, s); break;
    case Form::BF16: epi_fused_t<Form::BF16, float>(P, nchunks, stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart, s); break;
    case Form::BF16X1: epi_fused_t<Form::BF16X1, float>(P, nchunks, stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart, s); break;
    case Form::FP16: epi_fused_t<Form::FP16, float>(P, nchunks, stride, nwin, lp, sc, cf, colsum, gout, Wt, wpitch, wpiece, bpart, s); break;
    case Form::I8:
    case Form::F8: fail(kErrInternal, "epi_g_fused: the column-scaled forms take the unfused epilogue");
  }
  check_launch("epi_g_fused");
  const uint64_t nb = (nwin + kEpiRows - 1) / kEpiRows;
  sum_partials(bpart, nb, (uint64_t)lp, bvec, s);
  check_launch("epi_g_fused_sum");
}

void launch_epi_h(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* b, double* out, int mode, hipStream_t s, const double* colmul) {
  const unsigned g = grid_for(rows_pad * lp);
  require(!form_colscaled(f) || colmul, kErrInternal, "epi_h: this form needs column factors");
  if (f == Form::FP64)
    hipLaunchKernelGGL(k_epi_h<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, b, out, mode, colmul);
  else if (f == Form::I8)
    hipLaunchKernelGGL(k_epi_h<int32_t>, dim3(g), dim3(256), 0, s, static_cast<const int32_t*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, b, out, mode, colmul);
  else if (lp % 4 == 0 && chunk_stride % 4 == 0 && reinterpret_cast<uintptr_t>(P) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(out) % 16 == 0)
    hipLaunchKernelGGL(k_epi_h_f4, dim3(grid_for(rows_pad * lp / 4)), dim3(256), 0, s, static_cast<const float*>(P),
                       nchunks, chunk_stride, nreal, rows_pad, lp, b, out, mode, colmul);
  else
    hipLaunchKernelGGL(k_epi_h<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(P), nchunks, chunk_stride,
                       nreal, rows_pad, lp, b, out, mode, colmul);
  check_launch("epi_h");
}

static uint64_t colsum_rpb(uint64_t rows) {
  uint64_t rpb = (rows + 1023) / 1024;
  return rpb < 64 ? 64 : rpb;
}

void launch_colsum(const double* X, uint64_t r0, uint64_t r1, int lp, const double* w_sc, const double* w_cf,
                   double* partial, double* out, hipStream_t s) {
  if (r1 <= r0) {
    PCAONE_HIP(hipMemsetAsync(out, 0, lp * sizeof(double), s));
    return;
  }
  const uint64_t rpb = colsum_rpb(r1 - r0);
  const uint64_t nb = (r1 - r0 + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_colsum_partial, dim3((unsigned)nb), dim3(128), 0, s, X, r0, r1, lp, rpb, w_sc, w_cf, partial);
  check_launch("colsum_partial");
  sum_partials(partial, nb, (uint64_t)lp, out, s);
  check_launch("colsum_sum");
}

static uint64_t gram_rpb(uint64_t nrows) {
  uint64_t rpb = (nrows + 1023) / 1024;
  rpb = (rpb + kTileRows - 1) / kTileRows * kTileRows;
  return rpb < 128 ? 128 : rpb;
}

uint64_t gram_partial_elems(uint64_t nrows, int lp) {
  const uint64_t rpb = gram_rpb(nrows ? nrows : 1);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  return (nb ? nb : 1) * (uint64_t)lp * lp;
}

template <int NB>
static void gram_nb(const double* X, const double* Y, uint64_t nrows, uint64_t rpb, uint64_t nb, double* partial,
                    hipStream_t s) {
  if (Y)
    hipLaunchKernelGGL((k_gram_mfma<NB, true>), dim3((unsigned)nb), dim3(64 * NB), 0, s, X, Y, nrows, rpb, partial);
  else
    hipLaunchKernelGGL((k_gram_mfma<NB, false>), dim3((unsigned)nb), dim3(64 * NB), 0, s, X, X, nrows, rpb, partial);
}

void launch_gram(const double* X, const double* T, const double* Y0, uint64_t nrows, int lp, double* partial,
                 uint64_t partial_cap, double* out, hipStream_t s) {
  require(T == nullptr, kErrInternal, "gram: apply the transform first");
  require(lp <= 128 && lp % 16 == 0 && lp >= 16, kErrArgument, "gram: bad lp");
  if (nrows == 0) {
    PCAONE_HIP(hipMemsetAsync(out, 0, (size_t)lp * lp * sizeof(double), s));
    return;
  }
  const uint64_t rpb = gram_rpb(nrows);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  require(nb * lp * lp <= partial_cap, kErrInternal, "gram: partial buffer too small");
  switch (lp / 16) {
    case 1: gram_nb<1>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 2: gram_nb<2>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 3: gram_nb<3>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 4: gram_nb<4>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 5: gram_nb<5>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 6: gram_nb<6>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 7: gram_nb<7>(X, Y0, nrows, rpb, nb, partial, s); break;
    case 8: gram_nb<8>(X, Y0, nrows, rpb, nb, partial, s); break;
  }
  check_launch("gram_mfma");
  const uint64_t w = (uint64_t)lp * lp;
  sum_partials(partial, nb, w, out, s);
  check_launch("gram_sum");
}

template <int NB>
static void apply_nb(const double* X, const double* T, uint64_t nrows, double* Y, hipStream_t s) {
  constexpr size_t lds = (size_t)NB * 16 * NB * 16 * sizeof(double);
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_apply_mfma<NB>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  const uint64_t nblk = (nrows + 15) / 16;
  uint64_t g = (nblk + 3) / 4;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL((k_apply_mfma<NB>), dim3((unsigned)g), dim3(256), lds, s, X, T, nrows, Y);
}

void launch_apply(const double* X, const double* T, uint64_t nrows, int lp, double* Y, hipStream_t s) {
  if (!nrows) return;
  require(lp <= 128 && lp % 16 == 0 && lp >= 16, kErrArgument, "apply: bad lp");
  switch (lp / 16) {
    case 1: apply_nb<1>(X, T, nrows, Y, s); break;
    case 2: apply_nb<2>(X, T, nrows, Y, s); break;
    case 3: apply_nb<3>(X, T, nrows, Y, s); break;
    case 4: apply_nb<4>(X, T, nrows, Y, s); break;
    case 5: apply_nb<5>(X, T, nrows, Y, s); break;
    case 6: apply_nb<6>(X, T, nrows, Y, s); break;
    case 7: apply_nb<7>(X, T, nrows, Y, s); break;
    case 8: apply_nb<8>(X, T, nrows, Y, s); break;
  }
  check_launch("apply_mfma");
}

void launch_flip_norms(const double* O2, const double* Q, uint64_t nrows, int lp, double* partial, double* out,
                       hipStream_t s) {
  const uint64_t rpb = colsum_rpb(nrows);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_flip_partial, dim3((unsigned)nb), dim3(128), 0, s, O2, Q, nrows, lp, rpb, partial);
  check_launch("flip_partial");
  sum_partials(partial, nb, (uint64_t)2 * lp, out, s);
  check_launch("flip_sum");
}

void launch_add(const double* X, const double* Z, double* Y, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(256), 0, s, X, Z, Y, n);
  check_launch("add");
}

void launch_missing_weights(const MissWeightArgs& a, hipStream_t s) {
  if (!a.nrows) return;
  require(a.kspan > 0, kErrArgument, "missing_weights: zero index span");
  const uint64_t nz = (a.kext + a.kspan - 1) / a.kspan;
  require(nz < 65536, kErrArgument, "missing_weights: too many index chunks");
  const dim3 grid((unsigned)((a.nrows + 3) / 4), (unsigned)std::max<uint64_t>(nz, 1));
  if (!a.Fclamp) {
    hipLaunchKernelGGL(k_missing_weights, grid, dim3(256), 0, s, a);
  } else {
    require(a.kk >= 1 && a.kk <= 128 && a.kp % 16 == 0 && a.kp >= a.kk + 3, kErrArgument,
            "missing_weights: imputation rank must be in [1, 128]");
    int lpe = 1;
    while (lpe < (a.kk + 1) / 2) lpe *= 2;
    hipLaunchKernelGGL(k_missing_weights_emu, grid, dim3(256), 0, s, a, lpe);
  }
  check_launch("missing_weights");
}

void launch_pack_col_state(double* V, uint64_t m, int kp, int kk, const double* sc, const double* cf, const double* F,
                           hipStream_t s) {
  if (!m) return;
  hipLaunchKernelGGL(k_pack_col_state, dim3(grid_for(m)), dim3(256), 0, s, V, m, kp, kk, sc, cf, F);
  check_launch("pack_col_state");
}

void launch_compact_rows(const double* src, uint64_t rows, int lp, int L, int pitch, bool f32, void* out,
                         hipStream_t s) {
  if (!rows) return;
  const unsigned g = grid_for(rows * (uint64_t)pitch);
  if (f32)
    hipLaunchKernelGGL(k_compact_rows<float>, dim3(g), dim3(256), 0, s, src, rows, lp, L, pitch,
                       static_cast<float*>(out));
  else
    hipLaunchKernelGGL(k_compact_rows<double>, dim3(g), dim3(256), 0, s, src, rows, lp, L, pitch,
                       static_cast<double*>(out));
  check_launch("compact_rows");
}

void launch_missing_spmm(const MissSpmmArgs& a, hipStream_t s) {
  if (!a.nrows || a.k_hi <= a.k_lo) return;
  const int vec = a.f32 ? 4 : 2;
  require(a.ncols >= 1 && a.ncols <= a.lp && a.cp >= a.ncols && a.cp % vec == 0 && a.cp <= 128 &&
              a.cpitch >= a.cp && a.cpitch % vec == 0,
          kErrArgument, "missing_spmm: bad column counts");
  const uint64_t span = a.k_hi - a.k_lo;
  const uint64_t nz = a.kspan ? (span + a.kspan - 1) / a.kspan : 1;
  require(nz < 65536 && (nz == 1 || (a.partial && nz * a.nrows * (uint64_t)a.ncols <= a.partial_cap)), kErrInternal,
          "missing_spmm: bad index chunking");
  const dim3 block(256);
  const unsigned gx = (unsigned)((a.nrows + 3) / 4), gz = (unsigned)nz;
  if (a.f32) hipLaunchKernelGGL((k_missing_spmm<float>), dim3(gx, 1, gz), block, 0, s, a);
  else hipLaunchKernelGGL((k_missing_spmm<double>), dim3(gx, 1, gz), block, 0, s, a);
  check_launch("missing_spmm");
  if (nz > 1) launch_sum_chunks(a.partial, nz, a.nrows, a.ncols, a.out, a.lp, s);
}

void launch_sum_chunks(const double* partial, uint64_t nz, uint64_t nrows, int ncols, double* out, int lp,
                       hipStream_t s) {
  if (!nrows || !ncols) return;
  hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for(nrows * (uint64_t)ncols)), dim3(256), 0, s, partial, nz, nrows,
                     ncols, out, lp);
  check_launch("sum_chunks");
}

bool thinq_on_device(int lp) { return lp % 16 == 0 && lp <= 16 * kThinqMaxNB; }

void launch_thinq_first(const double* G1, int n, int lp, double* R1, double* T1, int* status, hipStream_t s,
                        double min_ratio) {
  require(thinq_on_device(lp) && n <= lp, kErrInternal, "thinq: lp too large for the device path");
  switch (lp / 16) {
    case 1: hipLaunchKernelGGL(k_thinq_first<1>, dim3(1), dim3(kSmallThreads), 0, s, G1, n, lp, R1, T1, status, min_ratio); break;
    case 2: hipLaunchKernelGGL(k_thinq_first<2>, dim3(1), dim3(kSmallThreads), 0, s, G1, n, lp, R1, T1, status, min_ratio); break;
    case 3: hipLaunchKernelGGL(k_thinq_first<3>, dim3(1), dim3(kSmallThreads), 0, s, G1, n, lp, R1, T1, status, min_ratio); break;
    case 4: hipLaunchKernelGGL(k_thinq_first<4>, dim3(1), dim3(kSmallThreads), 0, s, G1, n, lp, R1, T1, status, min_ratio); break;
    default: hipLaunchKernelGGL(k_thinq_first<5>, dim3(1), dim3(kSmallThreads), 0, s, G1, n, lp, R1, T1, status, min_ratio); break;
  }
  check_launch("thinq_first");
}

template <int NB>
void thinq_second_nb(const double* G2, const double* R1, const double* Xtop, int n, int lp, double* T, int* status,
                     hipStream_t s) {
  const size_t lds = (size_t)2 * lp * lp * sizeof(double);
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_thinq_second<NB>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(2 * 256 * NB * NB * sizeof(double))));
    attr = true;
  }
  hipLaunchKernelGGL(k_thinq_second<NB>, dim3(1), dim3(kSmallThreads), lds, s, G2, R1, Xtop, n, lp, T, status);
}

void launch_thinq_second(const double* G2, const double* R1, const double* Xtop, int n, int lp, double* T,
                         int* status, hipStream_t s) {
  require(thinq_on_device(lp) && n <= lp, kErrInternal, "thinq: lp too large for the device path");
  switch (lp / 16) {
    case 1: thinq_second_nb<1>(G2, R1, Xtop, n, lp, T, status, s); break;
    case 2: thinq_second_nb<2>(G2, R1, Xtop, n, lp, T, status, s); break;
    case 3: thinq_second_nb<3>(G2, R1, Xtop, n, lp, T, status, s); break;
    case 4: thinq_second_nb<4>(G2, R1, Xtop, n, lp, T, status, s); break;
    default: thinq_second_nb<5>(G2, R1, Xtop, n, lp, T, status, s); break;
  }
  check_launch("thinq_second");
}

void launch_flip_signs(const double* ab, int n, int lp, double* sign, hipStream_t s) {
  hipLaunchKernelGGL(k_flip_signs, dim3(1), dim3(128), 0, s, ab, n, lp, sign);
  check_launch("flip_signs");
}

}  // namespace pcaone
