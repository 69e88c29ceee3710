// Data-layout kernels: PLINK rows -> interleaved words, code counts (allele
// frequencies), SNP-major -> sample-major transpose, row gathers and the
// missing-entry lists.  All are integer-only and run once per data set (or
// once per permutation).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"
#include "synth.hpp"

namespace pcaone {
namespace {

constexpr uint32_t kAllMissing = 0x55555555u;  // sixteen codes 01

// Interleaved position of code e (0..15) inside a word.
__device__ __forceinline__ int ipos(int e) { return 8 * (e & 3) + 2 * (e >> 2); }

__global__ void k_interleave(const uint8_t* __restrict__ raw, uint64_t nrows, uint64_t bps, uint32_t nsamples,
                             uint32_t* __restrict__ out, uint64_t pitch) {
  const uint64_t total = nrows * pitch;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = t / pitch, w = t - r * pitch;
    const uint8_t* row = raw + r * bps;
    uint32_t word = 0;
    for (int e = 0; e < 16; ++e) {
      const uint64_t i = w * 16 + e;  // sample index
      uint32_t code = 1u;
      if (i < nsamples) code = (row[i >> 2] >> (2 * (i & 3))) & 3u;
      word |= code << ipos(e);
    }
    out[t] = word;
  }
}

__global__ void k_fill(uint32_t* __restrict__ x, uint64_t n) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x)
    x[t] = kAllMissing;
}

// One wave per row.  Codes beyond `ncodes` are excluded by masking the last word.
__global__ void k_code_counts(const uint32_t* __restrict__ x, uint64_t pitch, uint64_t nrows, uint64_t ncodes,
                              uint32_t* __restrict__ cnt) {
  const uint64_t r = blockIdx.x * (uint64_t)(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  const uint32_t* row = x + r * pitch;
  const uint64_t nw = (ncodes + 15) / 16;
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  for (uint64_t w = lane; w < nw; w += 64) {
    const uint32_t v = row[w];
    uint32_t valid = 0xFFFFFFFFu;  // bit pairs of codes < ncodes
    if ((w + 1) * 16 > ncodes) {
      valid = 0;
      for (int e = 0; e < 16; ++e)
        if (w * 16 + e < ncodes) valid |= 3u << (8 * (e & 3) + 2 * (e >> 2));
    }
    const uint32_t lo = v & 0x55555555u & valid, hi = (v >> 1) & 0x55555555u & valid;
    const uint32_t m = 0x55555555u & valid;
    c0 += __popc(m & ~lo & ~hi);
    c1 += __popc(lo & ~hi);
    c2 += __popc(~lo & hi & m);
    c3 += __popc(lo & hi);
  }
  for (int off = 32; off > 0; off >>= 1) {
    c0 += __shfl_down(c0, off);
    c1 += __shfl_down(c1, off);
    c2 += __shfl_down(c2, off);
    c3 += __shfl_down(c3, off);
  }
  if (lane == 0) {
    cnt[4 * r + 0] = c0;
    cnt[4 * r + 1] = c1;
    cnt[4 * r + 2] = c2;
    cnt[4 * r + 3] = c3;
  }
}

// Tile: 64 SNP rows x 256 samples.  Thread t loads 4 input words (rows
// 16*(t/64)... covering the tile) into LDS, then writes output words 0..3 of
// sample t of the tile (SNPs 64*by + 16*u .. +15).
__global__ void __launch_bounds__(256) k_transpose(const uint32_t* __restrict__ xs, uint64_t pitch_s, uint64_t nsnps,
                                                   uint64_t nsamples, uint32_t* __restrict__ xt, uint64_t pitch_t) {
  __shared__ uint32_t tile[64][17];
  const uint64_t j0 = (uint64_t)blockIdx.y * 64;   // SNP base
  const uint64_t i0 = (uint64_t)blockIdx.x * 256;  // sample base
  const int t = threadIdx.x;
  for (int v = t; v < 64 * 16; v += 256) {
    const int jr = v >> 4, wc = v & 15;
    const uint64_t j = j0 + jr, w = i0 / 16 + wc;
    uint32_t val = kAllMissing;
    if (j < nsnps && w < pitch_s) val = xs[j * pitch_s + w];
    tile[jr][wc] = val;
  }
  __syncthreads();
  const uint64_t i = i0 + t;
  if (i >= nsamples) return;
  const int wc = t >> 4, e = t & 15;  // sample's word and position inside the SNP-major words
  const int sh = ipos(e);
  uint4 o;
  uint32_t* op = reinterpret_cast<uint32_t*>(&o);
  for (int u = 0; u < 4; ++u) {
    uint32_t word = 0;
    for (int f = 0; f < 16; ++f) {
      const uint32_t code = (tile[16 * u + f][wc] >> sh) & 3u;
      word |= code << ipos(f);
    }
    op[u] = word;
  }
  const uint64_t wout = j0 / 16;
  uint32_t* dst = xt + i * pitch_t + wout;
  if (wout + 4 <= pitch_t) {
    *reinterpret_cast<uint4*>(dst) = o;
  } else {
    for (int u = 0; u < 4 && wout + u < pitch_t; ++u) dst[u] = op[u];
  }
}

__global__ void k_gather_rows(const uint32_t* __restrict__ in, uint64_t pitch, const uint32_t* __restrict__ idx,
                              uint64_t nrows, uint32_t* __restrict__ out) {
  const uint64_t total = nrows * pitch;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = t / pitch, w = t - r * pitch;
    out[t] = in[(uint64_t)idx[r] * pitch + w];
  }
}

// One thread per row, in code order.
__global__ void k_missing_lists(const uint32_t* __restrict__ x, uint64_t pitch, uint64_t nrows, uint64_t ncodes,
                                const uint64_t* __restrict__ rp, uint32_t* __restrict__ ix) {
  const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const uint32_t* row = x + r * pitch;
  uint64_t o = rp[r];
  const uint64_t nw = (ncodes + 15) / 16;
  for (uint64_t w = 0; w < nw; ++w) {
    const uint32_t v = row[w];
    const uint32_t m = v & ~(v >> 1) & 0x55555555u;
    if (!m) continue;
    for (int e = 0; e < 16; ++e) {
      const uint64_t k = w * 16 + e;
      if (k < ncodes && ((m >> ipos(e)) & 1u)) ix[o++] = (uint32_t)k;
    }
  }
}

// Synthetic genotypes straight into interleaved rows: one thread per output
// word of row rows[r] (or r when rows is null).
__global__ void k_synth(uint64_t seed, uint32_t nsamples, uint32_t npops, const uint64_t* __restrict__ thr,
                        const uint32_t* __restrict__ attempts, const uint8_t* __restrict__ flags,
                        synth::MissSpec ms, const uint32_t* __restrict__ rows, uint64_t nrows,
                        uint32_t* __restrict__ out, uint64_t pitch) {
  const uint64_t total = nrows * pitch;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = t / pitch, w = t - r * pitch;
    const uint64_t j = rows ? rows[r] : r;
    const uint32_t att = attempts[j];
    uint32_t word = 0;
    for (int e = 0; e < 16; ++e) {
      const uint64_t i = w * 16 + e;
      uint32_t code = 1u;
      if (i < nsamples) {
        const uint32_t pop = (uint32_t)(i * npops / nsamples);
        code = synth::genotype_code(seed, att, j, i, thr[j * npops + pop], ms, flags[i] != 0);
      }
      word |= code << ipos(e);
    }
    out[j * pitch + w] = word;
  }
}

// Inverse of k_interleave: interleaved rows -> PLINK rows (padding bits zero).
__global__ void k_deinterleave(const uint32_t* __restrict__ x, uint64_t pitch, uint64_t nrows, uint32_t nsamples,
                               uint8_t* __restrict__ raw, uint64_t bps) {
  const uint64_t total = nrows * bps;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = t / bps, b = t - r * bps;
    uint8_t v = 0;
    for (int s = 0; s < 4; ++s) {
      const uint64_t i = b * 4 + s;
      if (i >= nsamples) break;
      const uint32_t word = x[r * pitch + i / 16];
      v |= (uint8_t)(((word >> ipos((int)(i & 15))) & 3u) << (2 * s));
    }
    raw[t] = v;
  }
}

inline unsigned grid_for(uint64_t n, unsigned block = 256, unsigned cap = 65535u * 4) {
  uint64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  return (unsigned)(g ? g : 1);
}

}  // namespace

void launch_interleave(const uint8_t* raw, uint64_t nrows, uint64_t bps, uint32_t nsamples, uint32_t* out,
                       uint64_t pitch, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_interleave, dim3(grid_for(nrows * pitch)), dim3(256), 0, s, raw, nrows, bps, nsamples, out,
                     pitch);
  check_launch("interleave");
}

void launch_fill_rows(uint32_t* x, uint64_t pitch, uint64_t r0, uint64_t r1, hipStream_t s) {
  if (r1 <= r0) return;
  const uint64_t n = (r1 - r0) * pitch;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, x + r0 * pitch, n);
  check_launch("fill_rows");
}

void launch_code_counts(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint64_t ncodes, uint32_t* cnt,
                        hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_code_counts, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, s, x, pitch, nrows, ncodes,
                     cnt);
  check_launch("code_counts");
}

void launch_transpose_codes(const uint32_t* xs, uint64_t pitch_s, uint64_t nsnps, uint64_t nsamples, uint32_t* xt,
                            uint64_t pitch_t, uint64_t rows_t, hipStream_t s) {
  require(pitch_t * 16 >= nsnps, kErrInternal, "transpose: sample-major pitch too small");
  // SNP columns beyond nsnps must read as 01; the tile loader substitutes 01
  // for rows >= nsnps, and rows (samples) >= nsamples are filled separately.
  launch_fill_rows(xt, pitch_t, 0, rows_t, s);
  const uint64_t gx = (nsamples + 255) / 256, gy = (nsnps + 63) / 64;
  require(gy < (1ull << 31) && gx < (1ull << 31), kErrArgument, "transpose: matrix too large");
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, xs, pitch_s, nsnps, nsamples,
                     xt, pitch_t);
  check_launch("transpose");
}

void launch_gather_rows(const uint32_t* in, uint64_t pitch, const uint32_t* idx, uint64_t nrows, uint32_t* out,
                        hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(nrows * pitch)), dim3(256), 0, s, in, pitch, idx, nrows, out);
  check_launch("gather_rows");
}

void launch_synth(uint64_t seed, uint32_t nsamples, uint32_t npops, const uint64_t* thr, const uint32_t* attempts,
                  const uint8_t* flags, const synth::MissSpec& ms, const uint32_t* rows, uint64_t nrows, uint32_t* out,
                  uint64_t pitch, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_synth, dim3(grid_for(nrows * pitch)), dim3(256), 0, s, seed, nsamples, npops, thr, attempts,
                     flags, ms, rows, nrows, out, pitch);
  check_launch("synth");
}

void launch_deinterleave(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint32_t nsamples, uint8_t* raw,
                         uint64_t bps, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_deinterleave, dim3(grid_for(nrows * bps)), dim3(256), 0, s, x, pitch, nrows, nsamples, raw,
                     bps);
  check_launch("deinterleave");
}

void launch_missing_lists(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint64_t ncodes, const uint64_t* rp,
                          uint32_t* ix, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_missing_lists, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, x, pitch, nrows, ncodes,
                     rp, ix);
  check_launch("missing_lists");
}

}  // namespace pcaone

namespace pcaone {
namespace {
// Plain streaming copy (16 bytes per thread per iteration), used only to
// calibrate HBM byte counters against a known traffic volume.
__global__ void k_copy_probe(const uint4* __restrict__ in, uint4* __restrict__ out, uint64_t n) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x)
    out[t] = in[t];
}
}  // namespace

void launch_copy_probe(const void* in, void* out, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / 16;
  if (!n) return;
  hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, s, static_cast<const uint4*>(in),
                     static_cast<uint4*>(out), n);
  check_launch("copy_probe");
}
}  // namespace pcaone
