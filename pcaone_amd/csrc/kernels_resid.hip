// The residual matrix of Data::write_residuals (src/Data.cpp:242-291):
// R[:, j] = A[:, j] - U S V[j, :]^T, centered over the samples, as float32,
// with A the matrix the run multiplied (column state sc / cF, missing entries
// 0 or their EMU imputation).
//
// Tile: 32 SNP rows x 256 samples per workgroup, one sample per thread.  The
// SNP rows' S_t V_jt sit in LDS; each thread streams its sample's U_i in
// chunks of 16 (U stored k x N, so the loads coalesce) and keeps 32 FP64
// accumulators.  Pass 0 writes per-(sample chunk, SNP) sums of R; pass 1 takes
// each SNP's mean from them (chunks in order, so the result is deterministic),
// recomputes R and writes R - mean as float32, SNP-major (row j: N floats).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace pcaone {

namespace {

constexpr int kRTileSnps = 32;
constexpr int kRThreads = 256;
constexpr int kRChunkT = 16;

__device__ __forceinline__ int ipos(int e) { return 8 * (e & 3) + 2 * (e >> 2); }

__global__ void __launch_bounds__(kRThreads) k_residuals(ResidualArgs a, int pass) {
  __shared__ double sv[kRTileSnps * kResidMaxK];
  __shared__ double red[kRThreads / 64][kRTileSnps];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t jt = a.j0 + (uint64_t)blockIdx.x * kRTileSnps;  // first SNP of the tile
  const int nj = (int)min<uint64_t>(kRTileSnps, a.j1 - jt);
  const uint64_t i = (uint64_t)blockIdx.y * kRThreads + tid;
  const bool live = i < a.n;
  for (int e = tid; e < kRTileSnps * a.k; e += kRThreads) {
    const int s = e / a.k, t = e - s * a.k;
    sv[e] = s < nj ? a.sv[(jt + s) * (uint64_t)a.k + t] : 0.0;
  }
  __syncthreads();
  double acc[kRTileSnps];
#pragma unroll
  for (int s = 0; s < kRTileSnps; ++s) acc[s] = 0.0;
  // rank-k part: acc[s] = sum_t U_it S_t V_jt
  for (int t0 = 0; t0 < a.k; t0 += kRChunkT) {
    double u[kRChunkT];
#pragma unroll
    for (int c = 0; c < kRChunkT; ++c) u[c] = (live && t0 + c < a.k) ? a.ut[(uint64_t)(t0 + c) * a.n + i] : 0.0;
#pragma unroll
    for (int s = 0; s < kRTileSnps; ++s) {
      double d = 0.0;
#pragma unroll
      for (int c = 0; c < kRChunkT; ++c)
        if (t0 + c < a.k) d += u[c] * sv[s * a.k + t0 + c];
      acc[s] += d;
    }
  }
  // the matrix entry minus the rank-k part
  const double geno[4] = {1.0, 0.0, 0.5, 0.0};  // BED2GENO (src/Common.hpp:58-59); code 01 is missing
#pragma unroll
  for (int s = 0; s < kRTileSnps; ++s) {
    double r = 0.0;
    if (live && s < nj) {
      const uint64_t j = jt + s;
      const uint32_t word = a.xs[j * a.pitch + i / 16];
      const int code = (int)((word >> ipos((int)(i & 15))) & 3u);
      const double sc = a.sc ? a.sc[j] : 1.0;
      double x = 0.0;
      if (code != 1) {
        x = sc * (geno[code] - a.cf[j]);
      } else if (a.us) {  // EMU: the imputed value sc clamp((U S V^T)_ij, -F, 1 - F)
        double p = 0.0;
        for (int t = 0; t < a.kimp; ++t) p += a.us[i * a.kpimp + t] * a.vimp[j * a.kpimp + t];
        const double f = a.fclamp[j];
        x = sc * fmin(fmax(p, -f), 1.0 - f);
      }
      r = x - acc[s];
    }
    acc[s] = r;
  }
  if (pass == 0) {
    // sums over this chunk's samples, per SNP
#pragma unroll
    for (int s = 0; s < kRTileSnps; ++s) {
      double v = acc[s];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[wave][s] = v;
    }
    __syncthreads();
    if (tid < nj) {
      double v = 0.0;
      for (int w = 0; w < kRThreads / 64; ++w) v += red[w][tid];
      a.part[(uint64_t)blockIdx.y * a.part_stride + (jt - a.j0) + tid] = v;
    }
    return;
  }
  // pass 1: the means from the chunk sums (in chunk order), then R - mean
  if (tid < nj) {
    double v = 0.0;
    for (uint64_t c = 0; c < a.nchunks; ++c) v += a.part[c * a.part_stride + (jt - a.j0) + tid];
    red[0][tid] = v / (double)a.n;
  }
  __syncthreads();
  if (!live) return;
#pragma unroll
  for (int s = 0; s < kRTileSnps; ++s)
    if (s < nj) a.out[(jt - a.j0 + s) * a.n + i] = (float)(acc[s] - red[0][s]);
}

}  // namespace

void launch_residuals(const ResidualArgs& a0, hipStream_t s) {
  if (a0.j1 <= a0.j0 || !a0.n) return;
  ResidualArgs a = a0;
  a.nchunks = (a.n + kRThreads - 1) / kRThreads;
  a.part_stride = a.j1 - a.j0;
  const dim3 grid((unsigned)((a.j1 - a.j0 + kRTileSnps - 1) / kRTileSnps), (unsigned)a.nchunks);
  hipLaunchKernelGGL(k_residuals, grid, dim3(kRThreads), 0, s, a, 0);
  hipLaunchKernelGGL(k_residuals, grid, dim3(kRThreads), 0, s, a, 1);
  check_launch("residuals");
}

}  // namespace pcaone
