// FP64 kernels around the genotype products: MFMA operand pieces, the affine
// epilogues of the two products, column sums, Gram reductions (CholeskyQR2,
// B*B^T, mev), small right-multiplications, flipOmg norms and the
// missing-entry (and EMU imputation) corrections.
#include <hip/hip_runtime.h>

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

// Round-to-nearest-even FP32 -> BF16 bits (inputs are finite).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

template <Form F>
__global__ void k_make_pieces(const double* __restrict__ X, uint64_t nreal, uint64_t rows_pad, int lp,
                              const double* __restrict__ rowscale, const double* __restrict__ colsign, void* Wt,
                              uint64_t wpitch, uint64_t wpiece, double* __restrict__ rounded,
                              double* __restrict__ signed_out) {
  const uint64_t n = rows_pad * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t l = t / rows_pad, r = t - l * rows_pad;
    double x = 0.0;
    if (r < nreal) {
      x = X[r * lp + l];
      if (colsign) x *= colsign[l];
    }
    if (signed_out) signed_out[r * lp + l] = x;
    const double rs = rowscale ? rowscale[r] : 1.0;
    const double y = x * rs;
    double back;
    if constexpr (F == Form::FP64) {
      static_cast<double*>(Wt)[l * wpitch + r] = y;
      back = y;
    } else if constexpr (F == Form::FP32) {
      const float yf = (float)y;
      static_cast<float*>(Wt)[l * wpitch + r] = yf;
      back = (double)yf;
    } else {
      const uint16_t hi = bf16_bits((float)y);
      uint16_t* w = static_cast<uint16_t*>(Wt);
      w[l * wpitch + r] = hi;
      back = (double)bf16_value(hi);
      if constexpr (F == Form::BF16) {
        const uint16_t lo = bf16_bits((float)(y - back));
        w[wpiece + l * wpitch + r] = lo;
        back += (double)bf16_value(lo);
      }
    }
    if (rounded) rounded[r * lp + l] = (r < nreal && rs != 0.0) ? back / rs : 0.0;
  }
}

template <class P>
__global__ void k_epi_g(const P* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                        uint64_t rows_pad, int lp, const double* __restrict__ sc, const double* __restrict__ cf,
                        const double* __restrict__ colsum, double* __restrict__ out) {
  const uint64_t n = rows_pad * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t j = t / lp;
    const int l = (int)(t - j * lp);
    double v = 0.0;
    if (j < nreal) {
      double y = 0.0;
      for (uint64_t c = 0; c < nchunks; ++c) y += (double)part[c * stride + t];
      const double s = sc[j];
      v = 0.25 * s * y + s * (0.5 - cf[j]) * colsum[l];
    }
    out[t] = v;
  }
}

template <class P>
__global__ void k_epi_h(const P* __restrict__ part, uint64_t nchunks, uint64_t stride, uint64_t nreal,
                        uint64_t rows_pad, int lp, const double* __restrict__ b, double* __restrict__ out, int mode) {
  const uint64_t n = rows_pad * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t i = t / lp;
    const int l = (int)(t - i * lp);
    if (i < nreal) {
      double y = 0.0;
      for (uint64_t c = 0; c < nchunks; ++c) y += (double)part[c * stride + t];
      const double v = 0.25 * y + b[l];
      out[t] = mode ? out[t] + v : v;
    } else if (!mode) {
      out[t] = 0.0;
    }
  }
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

__global__ void k_sum_partials(const double* __restrict__ partial, uint64_t nblocks, uint64_t width,
                               double* __restrict__ out) {
  GRID_STRIDE(t, width) {
    double s = 0.0;
    for (uint64_t b = 0; b < nblocks; ++b) s += partial[b * width + t];
    out[t] = s;
  }
}

constexpr int kGramRows = 16;

// out_partial[blk][a][b] = sum_{r in block} X[r][a] * Y[r][b]
__global__ void __launch_bounds__(256) k_gram_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                                      uint64_t nrows, int lp, uint64_t rpb,
                                                      double* __restrict__ partial) {
  __shared__ double xs[kGramRows][128];
  __shared__ double ys[kGramRows][128];
  const int t = threadIdx.x, ar = t >> 4, bc = t & 15;
  const int nb = lp / 16;
  double acc[8][8];
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
  const uint64_t rb = (uint64_t)blockIdx.x * rpb;
  const uint64_t re = rb + rpb < nrows ? rb + rpb : nrows;
  for (uint64_t r0 = rb; r0 < re; r0 += kGramRows) {
    __syncthreads();
    for (int v = t; v < kGramRows * lp; v += 256) {
      const int rr = v / lp, c = v - rr * lp;
      const uint64_t r = r0 + rr;
      xs[rr][c] = r < re ? X[r * lp + c] : 0.0;
      ys[rr][c] = r < re ? Y[r * lp + c] : 0.0;
    }
    __syncthreads();
    for (int rr = 0; rr < kGramRows; ++rr) {
      double xv[8], yv[8];
      for (int i = 0; i < 8; ++i) {
        xv[i] = i < nb ? xs[rr][ar + 16 * i] : 0.0;
        yv[i] = i < nb ? ys[rr][bc + 16 * i] : 0.0;
      }
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) acc[i][j] += xv[i] * yv[j];
    }
  }
  double* o = partial + (uint64_t)blockIdx.x * lp * lp;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) o[(ar + 16 * i) * lp + bc + 16 * j] = acc[i][j];
}

__global__ void k_apply(const double* __restrict__ X, const double* __restrict__ T, uint64_t nrows, int lp,
                        double* __restrict__ Y) {
  const uint64_t n = nrows * (uint64_t)lp;
  GRID_STRIDE(t, n) {
    const uint64_t r = t / lp;
    const int b = (int)(t - r * lp);
    const double* x = X + r * lp;
    double s = 0.0;
    for (int a = 0; a < lp; ++a) s += x[a] * T[a * lp + b];
    Y[t] = s;
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

__global__ void k_add(const double* __restrict__ X, const double* __restrict__ Z, double* __restrict__ Y, uint64_t n) {
  GRID_STRIDE(t, n) Y[t] = X[t] + Z[t];
}

// One wave per target row.
__global__ void k_missing_correction(MissArgs a) {
  const uint64_t r = blockIdx.x * (uint64_t)(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= a.nrows) return;
  uint64_t e = a.rp[r];
  const uint64_t e1 = a.rp[r + 1];
  // entries are sorted by k: skip those below k_lo
  {
    uint64_t lo = e, hi = e1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (a.ix[mid] < a.k_lo) lo = mid + 1; else hi = mid;
    }
    e = lo;
  }
  double acc0 = 0.0, acc1 = 0.0;
  const int l0 = lane, l1 = lane + 64;
  for (; e < e1; ++e) {
    const uint64_t k = a.ix[e];
    if (k >= a.k_hi) break;
    const uint64_t c = a.by_snp ? r : k;  // SNP column
    const uint64_t i = a.by_snp ? k : r;  // sample
    const double s = a.sc[c];
    double w = -s * (0.5 - a.cf[c]);
    if (a.Fclamp) {
      double p = 0.0;
      for (int t = 0; t < a.kk; ++t) p += a.US[i * a.kk + t] * a.V[c * a.kk + t];
      const double f = a.Fclamp[c];
      p = fmin(fmax(p, -f), 1.0 - f);
      w += s * p;
    }
    const double* src = a.src + k * a.lp;
    if (l0 < a.lp) acc0 += w * src[l0];
    if (l1 < a.lp) acc1 += w * src[l1];
  }
  double* o = a.out + r * a.lp;
  if (l0 < a.lp) o[l0] += acc0;
  if (l1 < a.lp) o[l1] += acc1;
}

}  // namespace

void launch_make_pieces(Form f, const double* X, uint64_t nreal, uint64_t rows_pad, int lp, const double* rowscale,
                        const double* colsign, void* Wt, uint64_t wpitch, uint64_t wpiece, double* rounded,
                        double* signed_out, hipStream_t s) {
  const unsigned g = grid_for(rows_pad * lp);
  switch (f) {
    case Form::FP64:
      hipLaunchKernelGGL(k_make_pieces<Form::FP64>, dim3(g), dim3(256), 0, s, X, nreal, rows_pad, lp, rowscale,
                         colsign, Wt, wpitch, wpiece, rounded, signed_out);
      break;
    case Form::FP32:
      hipLaunchKernelGGL(k_make_pieces<Form::FP32>, dim3(g), dim3(256), 0, s, X, nreal, rows_pad, lp, rowscale,
                         colsign, Wt, wpitch, wpiece, rounded, signed_out);
      break;
    case Form::BF16:
      hipLaunchKernelGGL(k_make_pieces<Form::BF16>, dim3(g), dim3(256), 0, s, X, nreal, rows_pad, lp, rowscale,
                         colsign, Wt, wpitch, wpiece, rounded, signed_out);
      break;
    case Form::BF16X1:
      hipLaunchKernelGGL(k_make_pieces<Form::BF16X1>, dim3(g), dim3(256), 0, s, X, nreal, rows_pad, lp, rowscale,
                         colsign, Wt, wpitch, wpiece, rounded, signed_out);
      break;
  }
  check_launch("make_pieces");
}

void launch_epi_g(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* sc, const double* cf, const double* colsum, double* out, hipStream_t s) {
  const unsigned g = grid_for(rows_pad * lp);
  if (f == Form::FP64)
    hipLaunchKernelGGL(k_epi_g<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, sc, cf, colsum, out);
  else
    hipLaunchKernelGGL(k_epi_g<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(P), nchunks, chunk_stride,
                       nreal, rows_pad, lp, sc, cf, colsum, out);
  check_launch("epi_g");
}

void launch_epi_h(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* b, double* out, int mode, hipStream_t s) {
  const unsigned g = grid_for(rows_pad * lp);
  if (f == Form::FP64)
    hipLaunchKernelGGL(k_epi_h<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(P), nchunks,
                       chunk_stride, nreal, rows_pad, lp, b, out, mode);
  else
    hipLaunchKernelGGL(k_epi_h<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(P), nchunks, chunk_stride,
                       nreal, rows_pad, lp, b, out, mode);
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
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(128), 0, s, partial, nb, (uint64_t)lp, out);
  check_launch("colsum_sum");
}

static uint64_t gram_rpb(uint64_t nrows) {
  uint64_t rpb = (nrows + 511) / 512;
  rpb = (rpb + kGramRows - 1) / kGramRows * kGramRows;
  return rpb < 256 ? 256 : rpb;
}

uint64_t gram_partial_elems(uint64_t nrows, int lp) {
  const uint64_t rpb = gram_rpb(nrows ? nrows : 1);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  return (nb ? nb : 1) * (uint64_t)lp * lp;
}

void launch_gram(const double* X, const double* T, const double* Y0, uint64_t nrows, int lp, double* partial,
                 uint64_t partial_cap, double* out, hipStream_t s) {
  require(T == nullptr, kErrInternal, "gram: apply the transform first");
  require(lp <= 128 && lp % 16 == 0, kErrArgument, "gram: bad lp");
  if (nrows == 0) {
    PCAONE_HIP(hipMemsetAsync(out, 0, (size_t)lp * lp * sizeof(double), s));
    return;
  }
  const uint64_t rpb = gram_rpb(nrows);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  require(nb * lp * lp <= partial_cap, kErrInternal, "gram: partial buffer too small");
  hipLaunchKernelGGL(k_gram_partial, dim3((unsigned)nb), dim3(256), 0, s, X, Y0 ? Y0 : X, nrows, lp, rpb, partial);
  check_launch("gram_partial");
  const uint64_t w = (uint64_t)lp * lp;
  hipLaunchKernelGGL(k_sum_partials, dim3(grid_for(w)), dim3(256), 0, s, partial, nb, w, out);
  check_launch("gram_sum");
}

void launch_apply(const double* X, const double* T, uint64_t nrows, int lp, double* Y, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_apply, dim3(grid_for(nrows * lp)), dim3(256), 0, s, X, T, nrows, lp, Y);
  check_launch("apply");
}

void launch_flip_norms(const double* O2, const double* Q, uint64_t nrows, int lp, double* partial, double* out,
                       hipStream_t s) {
  const uint64_t rpb = colsum_rpb(nrows);
  const uint64_t nb = (nrows + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_flip_partial, dim3((unsigned)nb), dim3(128), 0, s, O2, Q, nrows, lp, rpb, partial);
  check_launch("flip_partial");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, nb, (uint64_t)2 * lp, out);
  check_launch("flip_sum");
}

void launch_add(const double* X, const double* Z, double* Y, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(256), 0, s, X, Z, Y, n);
  check_launch("add");
}

void launch_missing_correction(const MissArgs& a, hipStream_t s) {
  if (!a.nrows) return;
  require(a.lp <= 128, kErrArgument, "missing_correction: lp > 128");
  hipLaunchKernelGGL(k_missing_correction, dim3((unsigned)((a.nrows + 3) / 4)), dim3(256), 0, s, a);
  check_launch("missing_correction");
}

}  // namespace pcaone
