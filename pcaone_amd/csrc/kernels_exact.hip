// Kernels of the exact-PCA (--svd 3) and IRAM (--svd 0) drivers: standardized
// sample columns or SNP rows of the resident genotypes as dense FP64 operands,
// and the placement of product blocks into an N x N Gram matrix.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace pcaone {

namespace {

#define GRID_STRIDE(t, n) \
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < (n); t += (uint64_t)gridDim.x * blockDim.x)

inline unsigned grid_for(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65535u * 4 ? 65535u * 4 : g));
}

__device__ __forceinline__ int ipos(int e) { return 8 * (e & 3) + 2 * (e >> 2); }

// The standardized value sc_j (BED2GENO[code] - cF_j) of (sample i, SNP j);
// missing entries are 0 (src/FilePlink.cpp:102-109, src/Data.cpp:351-362).
__device__ __forceinline__ double std_value(const uint32_t* xs, uint64_t pitch, uint64_t j, uint64_t i,
                                            const double* sc, const double* cf) {
  const uint32_t word = xs[j * pitch + i / 16];
  const int code = (int)((word >> ipos((int)(i & 15))) & 3u);
  if (code == 1) return 0.0;
  const double g = code == 0 ? 1.0 : (code == 2 ? 0.5 : 0.0);
  return sc[j] * (g - cf[j]);
}

// gt[j][c] = A_std(sample s0 + c, SNP j) for c < nc, 0 for the padding columns.
__global__ void k_sample_columns(const uint32_t* __restrict__ xs, uint64_t pitch, uint64_t m, uint64_t s0, int nc,
                                 int lp, const double* __restrict__ sc, const double* __restrict__ cf,
                                 double* __restrict__ gt) {
  GRID_STRIDE(t, m * (uint64_t)lp) {
    const uint64_t j = t / lp;
    const int c = (int)(t - j * lp);
    gt[t] = c < nc ? std_value(xs, pitch, j, s0 + c, sc, cf) : 0.0;
  }
}

// om[i + c n] = A_std(sample i, SNP t0 + c) for c < nc, 0 beyond (column-major n x L).
__global__ void k_snp_rows(const uint32_t* __restrict__ xs, uint64_t pitch, uint64_t n, uint64_t t0, int nc, int L,
                           const double* __restrict__ sc, const double* __restrict__ cf, double* __restrict__ om) {
  GRID_STRIDE(t, n * (uint64_t)L) {
    const uint64_t c = t / n, i = t - c * n;
    om[t] = (int)c < nc ? std_value(xs, pitch, t0 + c, i, sc, cf) : 0.0;
  }
}

// K[r + (col0 + c) ldk] = scale X[r lp + c] for r < rows, c < nc.
__global__ void k_rows_to_cols(const double* __restrict__ X, uint64_t rows, int lp, int nc, double scale,
                               double* __restrict__ K, uint64_t ldk, uint64_t col0) {
  GRID_STRIDE(t, rows * (uint64_t)nc) {
    const uint64_t c = t / rows, r = t - c * rows;
    K[r + (col0 + c) * ldk] = scale * X[r * lp + c];
  }
}

}  // namespace

void launch_sample_columns(const uint32_t* xs, uint64_t pitch, uint64_t m, uint64_t s0, int nc, int lp,
                           const double* sc, const double* cf, double* gt, hipStream_t s) {
  hipLaunchKernelGGL(k_sample_columns, dim3(grid_for(m * (uint64_t)lp)), dim3(256), 0, s, xs, pitch, m, s0, nc, lp, sc,
                     cf, gt);
  check_launch("sample_columns");
}

void launch_snp_rows(const uint32_t* xs, uint64_t pitch, uint64_t n, uint64_t t0, int nc, int L, const double* sc,
                     const double* cf, double* om, hipStream_t s) {
  hipLaunchKernelGGL(k_snp_rows, dim3(grid_for(n * (uint64_t)L)), dim3(256), 0, s, xs, pitch, n, t0, nc, L, sc, cf, om);
  check_launch("snp_rows");
}

void launch_rows_to_cols(const double* X, uint64_t rows, int lp, int nc, double scale, double* K, uint64_t ldk,
                         uint64_t col0, hipStream_t s) {
  hipLaunchKernelGGL(k_rows_to_cols, dim3(grid_for(rows * (uint64_t)nc)), dim3(256), 0, s, X, rows, lp, nc, scale, K,
                     ldk, col0);
  check_launch("rows_to_cols");
}

}  // namespace pcaone
