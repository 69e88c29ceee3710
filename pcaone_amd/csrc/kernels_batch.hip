// Grouped kernels of the batched randomized SVD: operand pieces, epilogues,
// reductions and applications over the rows of every cohort, and the
// per-cohort L x L steps (Cholesky, triangular inverse, Jacobi
// eigen-decomposition, Householder signs, stop rule) in one workgroup each.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

#include "batch_kernels.hpp"
#include "bf16.hpp"
#include "device.hpp"

namespace pcaone {
namespace {

constexpr int LP = kBatchLp;
typedef double f64x4_t __attribute__((ext_vector_type(4)));

template <Form F>
using piece_t = typename std::conditional<
    F == Form::FP64, double,
    typename std::conditional<F == Form::FP32, float,
                              typename std::conditional<F == Form::F8, uint8_t, uint16_t>::type>::type>::type;
template <Form F>
constexpr int pieces() {
  return F == Form::BF16 ? 2 : (F == Form::F8 ? 3 : 1);
}

// FP8 column scale: the smallest power of two >= max / 448 (the column
// maximum lands in (224, 448]), 1 for an all-zero column.
__device__ __forceinline__ double f8_col_scale(unsigned long long bits) {
  const double mx = __longlong_as_double((long long)bits);
  if (!(mx > 0.0)) return 1.0;
  int e;
  const double f = frexp(mx / 448.0, &e);
  return ldexp(1.0, f == 0.5 ? e - 1 : e);
}

// x -> operand pieces; returns the value they sum to.  FP8: three E4M3
// pieces of x / cs (cs the column's power-of-two scale), each the nearest
// E4M3 to what the previous ones left.
template <Form F>
__device__ __forceinline__ double split(double x, piece_t<F> (&pc)[3], double cs = 1.0) {
  if constexpr (F == Form::F8) {
    const double x1 = x / cs;
    double v1, v2, v3;
    pc[0] = e4m3_bits(x1, &v1);
    pc[1] = e4m3_bits(x1 - v1, &v2);
    pc[2] = e4m3_bits(x1 - v1 - v2, &v3);
    return (v1 + v2 + v3) * cs;
  } else if constexpr (F == Form::FP64) {
    pc[0] = x;
    return x;
  } else if constexpr (F == Form::FP32) {
    pc[0] = (float)x;
    return (double)pc[0];
  } else if constexpr (F == Form::FP16) {
    pc[0] = fp16_bits(x);
    return fp16_value(pc[0]);
  } else {
    pc[0] = bf16_bits((float)x);
    double back = (double)bf16_value(pc[0]);
    if constexpr (F == Form::BF16) {
      pc[1] = bf16_bits((float)(x - back));
      back += (double)bf16_value(pc[1]);
    }
    return back;
  }
}

// ---------------------------------------------------------------- Omega pieces
// Rows of a piece tile in LDS: one element of padding keeps the transposed
// writes (consecutive columns) on distinct banks.
template <class T>
constexpr int tile_pitch(int rows) {
  return rows + (sizeof(T) == 2 ? 2 : 1);
}

// Omega (signed by the flipOmg signs or not) -> the FP64 copy, the value its
// pieces sum to and the transposed pieces, in tiles of 64 rows: row-major
// reads and writes, the pieces transposed through LDS.
template <Form F>
__global__ void __launch_bounds__(256) k_b_make_pieces(BatchArgs a, const RowItem* __restrict__ items, int srcsel,
                                                       int signed_) {
  constexpr int P = pieces<F>();
  constexpr int TR = 64;
  using w_t = piece_t<F>;
  constexpr int TRP = tile_pitch<w_t>(TR);
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != 1) return;
  __shared__ w_t pt[P * LP * TRP];
  const double* src = c.mat[srcsel];
  double* o2 = c.mat[kBufOmega];
  double* orr = c.mat[kBufOmegaR];
  w_t* w = static_cast<w_t*>(c.wg);
  const uint64_t wpiece = (uint64_t)LP * c.wg_pitch;
  const int l = threadIdx.x % LP, r0 = threadIdx.x / LP;
  const double sgn = (signed_ && l < a.L) ? c.st->colsign[l] : 1.0;
  const double cs = F == Form::F8 ? f8_col_scale(c.st->cmax_g[l]) : 1.0;
  for (uint32_t t0 = 0; t0 < it.nrows; t0 += TR) {
    const int nr = (int)min<uint32_t>(TR, it.nrows - t0);
    for (int r = r0; r < TR; r += 256 / LP) {
      w_t pc[3] = {w_t(0), w_t(0), w_t(0)};
      if (r < nr) {
        const uint64_t t = (it.row0 + t0 + r) * LP + l;
        double x = 0.0;
        if (l < a.L) {
          x = src[t];
          if (signed_) x *= sgn;
        }
        const double back = split<F>(x, pc, cs);
        o2[t] = x;
        orr[t] = back;
      }
#pragma unroll
      for (int p = 0; p < P; ++p) pt[(p * LP + l) * TRP + r] = pc[p];
    }
    __syncthreads();
    for (int v = threadIdx.x; v < P * LP * TR; v += 256) {
      const int r = v % TR, pl = v / TR, p = pl / LP, ll = pl - p * LP;
      if (r < nr) w[p * wpiece + ll * c.wg_pitch + it.row0 + t0 + r] = pt[pl * TRP + r];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- row reductions
// partial[item][l] = sum over the item's rows of X[r][l] (colsum) or the two
// flipOmg norms (width 2 LP).  8 row groups x 32 columns, combined in order.
template <bool FLIP>
__global__ void __launch_bounds__(256) k_b_rowsum(BatchArgs a, const RowItem* __restrict__ items,
                                                  double* __restrict__ partial) {
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != 1) return;
  __shared__ double red[2][8][LP];
  const int l = threadIdx.x & 31, g = threadIdx.x >> 5;
  double s0 = 0.0, s1 = 0.0;
  if constexpr (FLIP) {
    const double* o2 = c.mat[kBufOmega];
    const double* q = c.mat[kBufOmegaC];
    for (uint32_t rr = g; rr < it.nrows; rr += 8) {
      const uint64_t t = (it.row0 + rr) * LP + l;
      s0 += fabs(o2[t] - q[t]);
      s1 += fabs(o2[t] + q[t]);
    }
  } else {
    const double* x = c.mat[kBufOmegaR];
    for (uint32_t rr = g; rr < it.nrows; rr += 8) s0 += x[(it.row0 + rr) * LP + l];
  }
  red[0][g][l] = s0;
  red[1][g][l] = s1;
  __syncthreads();
  if (threadIdx.x < LP) {
    double t0 = 0.0, t1 = 0.0;
    for (int q = 0; q < 8; ++q) t0 += red[0][q][l], t1 += red[1][q][l];
    if constexpr (FLIP) {
      partial[(uint64_t)blockIdx.x * 2 * LP + l] = t0;
      partial[(uint64_t)blockIdx.x * 2 * LP + LP + l] = t1;
    } else {
      partial[(uint64_t)blockIdx.x * LP + l] = t0;
    }
  }
}

// Per cohort: dest = sum over its items (in order) of partial[item][0:width].
enum SumDest : int { kDestColsum = 0, kDestBvec, kDestGramA, kDestGramH, kDestAb };
enum ItemList : int { kListM = 0, kListN, kListE };

__global__ void __launch_bounds__(256) k_b_sum_items(BatchArgs a, const double* __restrict__ partial, int width,
                                                     int list, int dest, int state_req) {
  const BatchCohort& c = a.coh[blockIdx.x];
  if (c.st->state != state_req) return;
  const uint32_t b = list == kListM ? c.m_begin : (list == kListN ? c.n_begin : c.e_begin);
  const uint32_t e = list == kListM ? c.m_end : (list == kListN ? c.n_end : c.e_end);
  double* out = dest == kDestColsum ? c.colsum
                : dest == kDestBvec ? c.bvec
                : dest == kDestGramA ? c.st->gram_a
                : dest == kDestGramH ? c.st->gram_h
                                     : c.st->ab;
  for (int v = threadIdx.x; v < width; v += 256) {
    double s = 0.0;
    for (uint32_t i = b; i < e; ++i) s += partial[(uint64_t)i * width + v];
    out[v] = s;
  }
}

// ---------------------------------------------------------------- epilogues
// Split-K partial element 4 l4 .. 4 l4 + 3 of one row, summed in FP64.
__device__ __forceinline__ void add4(const float* p, double (&y)[4]) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  y[0] += (double)v.x, y[1] += (double)v.y, y[2] += (double)v.z, y[3] += (double)v.w;
}
__device__ __forceinline__ void add4(const double* p, double (&y)[4]) {
  const double2 v0 = *reinterpret_cast<const double2*>(p), v1 = *reinterpret_cast<const double2*>(p + 2);
  y[0] += v0.x, y[1] += v0.y, y[2] += v1.x, y[3] += v1.y;
}

// G epilogue for kBatchEpiRows rows of one cohort: sums the split-K partials
// in FP64, applies G_j = sc_j y / 4 + sc_j (1/2 - cF_j) colsum, writes G~ (the
// value the operand pieces sum to, over sc_j), the transposed pieces of
// diag(sc) G~ and this item's share of the b-vector.  Each thread owns 4
// adjacent columns of every (256 / (LP / 4))-th row; its b terms stay in
// registers and the row slots are combined in order at the end.
// MODE 0 (fused): G~ = sc (y / 4 + (1/2 - cF) colsum) from the partial
// products, split into the operand pieces of diag(sc) G~ (rounded G~ = the
// pieces' sum / sc) with the b-vector partials.  The FP8 form takes two
// passes around the column maxima of diag(sc) G~: MODE 1 writes G~ (y times
// the Omega column scale), MODE 2 splits it with the H-side column scales.
template <Form F, class PT, int MODE>
__global__ void __launch_bounds__(256) k_b_epi_g(BatchArgs a, const RowItem* __restrict__ items,
                                                 double* __restrict__ bpart) {
  constexpr int P = pieces<F>();
  constexpr int CG = LP / 4, RG = 256 / CG;
  using w_t = piece_t<F>;
  constexpr int TRP = tile_pitch<w_t>(kBatchEpiRows);
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != 1) return;
  __shared__ w_t pt[P * LP * TRP];
  __shared__ double bred[RG * LP];
  const PT* part = static_cast<const PT*>(c.part);
  double* gout = c.mat[kBufG];
  const int nr = (int)it.nrows;
  const int l0 = 4 * (threadIdx.x % CG), rs = threadIdx.x / CG;
  double cs[4], ms[4], bs[4] = {0.0, 0.0, 0.0, 0.0}, mx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cs[q] = c.colsum[l0 + q];
    ms[q] = 1.0;
    if (F == Form::F8) ms[q] = f8_col_scale(MODE == 2 ? c.st->cmax_h[l0 + q] : c.st->cmax_g[l0 + q]);
  }
  for (int r = rs; r < kBatchEpiRows; r += RG) {
    w_t pc[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) pc[q][0] = pc[q][1] = pc[q][2] = w_t(0);
    if (r < nr) {
      const uint64_t j = it.row0 + r;
      const uint64_t t = j * LP + l0;
      const double s = c.sc[j];
      const double beta = s * (0.5 - c.cf[j]);
      double g[4];
      if (MODE == 2) {
        const double2 u = *reinterpret_cast<const double2*>(gout + t);
        const double2 v = *reinterpret_cast<const double2*>(gout + t + 2);
        g[0] = u.x, g[1] = u.y, g[2] = v.x, g[3] = v.y;
      } else {
        double y[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint64_t ch = 0; ch < c.nch_g; ++ch) add4(part + ch * c.ychunk_g + t, y);
#pragma unroll
        for (int q = 0; q < 4; ++q) g[q] = 0.25 * s * (y[q] * (MODE == 1 ? ms[q] : 1.0)) + beta * cs[q];
      }
      double gr[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (MODE == 1) {
          gr[q] = g[q];
          mx[q] = fmax(mx[q], fabs(g[q] * s));
        } else {
          const double back = split<F>(g[q] * s, pc[q], MODE == 2 ? ms[q] : 1.0);
          gr[q] = s != 0.0 ? back / s : 0.0;
          bs[q] += beta * gr[q];
        }
      }
      *reinterpret_cast<double2*>(gout + t) = double2{gr[0], gr[1]};
      *reinterpret_cast<double2*>(gout + t + 2) = double2{gr[2], gr[3]};
    }
    if (MODE != 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < P; ++p) pt[(p * LP + l0 + q) * TRP + r] = pc[q][p];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) bred[rs * LP + l0 + q] = MODE == 1 ? mx[q] : bs[q];
  __syncthreads();
  if (MODE == 1) {  // the items' column maxima of diag(sc) G~, one atomic per column
    if (threadIdx.x < LP) {
      const int l = threadIdx.x;
      double m = 0.0;
      for (int q = 0; q < RG; ++q) m = fmax(m, bred[q * LP + l]);
      if (m > 0.0) atomicMax(&c.st->cmax_h[l], (unsigned long long)__double_as_longlong(m));
    }
    return;
  }
  w_t* w = static_cast<w_t*>(c.wh);
  const uint64_t wpiece = (uint64_t)LP * c.wh_pitch;
  if constexpr (sizeof(w_t) == 1) {
    // four rows' bytes per 32-bit store (rows start on 4-byte boundaries: the
    // pitch and row0 are multiples of 128)
    constexpr int R4 = kBatchEpiRows / 4;
    for (int v = threadIdx.x; v < P * LP * R4; v += 256) {
      const int r = 4 * (v % R4), pl = v / R4, p = pl / LP, l = pl - p * LP;
      if (r >= nr) continue;
      const w_t* src = pt + pl * TRP + r;
      w_t* dst = w + p * wpiece + l * c.wh_pitch + it.row0 + r;
      if (r + 4 <= nr)
        *reinterpret_cast<uint32_t*>(dst) = (uint32_t)src[0] | ((uint32_t)src[1] << 8) | ((uint32_t)src[2] << 16) |
                                            ((uint32_t)src[3] << 24);
      else
        for (int q = 0; q < nr - r; ++q) dst[q] = src[q];
    }
  } else {
    for (int v = threadIdx.x; v < P * LP * kBatchEpiRows; v += 256) {
      const int r = v % kBatchEpiRows, pl = v / kBatchEpiRows, p = pl / LP, l = pl - p * LP;
      if (r < nr) w[p * wpiece + l * c.wh_pitch + it.row0 + r] = pt[pl * TRP + r];
    }
  }
  if (threadIdx.x < LP) {
    const int l = threadIdx.x;
    double sm = 0.0;
    for (int q = 0; q < RG; ++q) sm += bred[q * LP + l];
    bpart[(uint64_t)blockIdx.x * LP + l] = sm;
  }
}

// FP8 form: column maxima of |X[r][l]| over the items' rows into cmax_g, one
// atomic per column and item.
__global__ void __launch_bounds__(256) k_b_colmax(BatchArgs a, const RowItem* __restrict__ items, int srcsel) {
  __shared__ double m[256];
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != 1) return;
  const double* X = c.mat[srcsel];
  const int l = threadIdx.x % LP, r0 = threadIdx.x / LP;
  double mx = 0.0;
  if (l < a.L)
    for (uint32_t r = r0; r < it.nrows; r += 256 / LP) mx = fmax(mx, fabs(X[(it.row0 + r) * LP + l]));
  m[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x < LP) {
    for (int q = 1; q < 256 / LP; ++q) mx = fmax(mx, m[q * LP + l]);
    if (mx > 0.0) atomicMax(&c.st->cmax_g[l], (unsigned long long)__double_as_longlong(mx));
  }
}

// Clears cmax_g or cmax_h of the running cohorts (one workgroup per cohort).
__global__ void k_b_zero_cmax(BatchArgs a, int h_side) {
  const BatchCohort& c = a.coh[blockIdx.x];
  if (c.st->state != 1 || threadIdx.x >= LP) return;
  (h_side ? c.st->cmax_h : c.st->cmax_g)[threadIdx.x] = 0ull;
}

// H = y / 4 + b over the item's rows, 4 adjacent elements per thread.
template <class PT>
__global__ void __launch_bounds__(256) k_b_epi_h(BatchArgs a, const RowItem* __restrict__ items) {
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != 1) return;
  const PT* part = static_cast<const PT*>(c.part);
  double* h = c.mat[kBufH];
  const int n4 = (int)it.nrows * (LP / 4);
  const int l0 = 4 * (threadIdx.x % (LP / 4));  // the same columns on every pass (256 is a multiple of LP / 4)
  double b[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) b[q] = c.bvec[l0 + q];
  // FP8 form: the products are of diag(sc) G~ / scale_l: times the H-side scale
  double m[4] = {0.25, 0.25, 0.25, 0.25};
  if (a.form == (int)Form::F8) {
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = 0.25 * f8_col_scale(c.st->cmax_h[l0 + q]);
  }
  for (int v = threadIdx.x; v < n4; v += 256) {
    const uint64_t t = it.row0 * LP + 4 * (uint64_t)v;
    double y[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t ch = 0; ch < c.nch_h; ++ch) add4(part + ch * c.ychunk_h + t, y);
    *reinterpret_cast<double2*>(h + t) = double2{m[0] * y[0] + b[0], m[1] * y[1] + b[1]};
    *reinterpret_cast<double2*>(h + t + 2) = double2{m[2] * y[2] + b[2], m[3] * y[3] + b[3]};
  }
}

// ---------------------------------------------------------------- Gram and apply
constexpr int kTileRows = 32;

// partial[item] = X^T Y over the item's rows (LP x LP), FP64 MFMA; wave wv
// owns output rows 16 wv .. 16 wv + 15.
template <bool CROSS>
__global__ void __launch_bounds__(128) k_b_gram(BatchArgs a, const RowItem* __restrict__ items, int xsel, int ysel,
                                                int state_req, double* __restrict__ partial) {
  constexpr int NB = LP / 16;
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != state_req) return;
  __shared__ double xs[kTileRows * LP];
  __shared__ double ys[CROSS ? kTileRows * LP : 1];
  const double* X = c.mat[xsel];
  const double* Y = CROSS ? c.mat[ysel] : X;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int q = lane >> 4, r16 = lane & 15;
  f64x4_t acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f64x4_t{0, 0, 0, 0};
  const uint64_t re = it.row0 + it.nrows;
  // The next row tile is loaded into registers while the current one is
  // multiplied; rows past the item's end read as zero.
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
  if (it.row0 < re) fetch(it.row0);
  for (uint64_t r0 = it.row0; r0 < re; r0 += kTileRows) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      xs[tid + e * 64 * NB] = px[e];
      if constexpr (CROSS) ys[tid + e * 64 * NB] = py[e];
    }
    __syncthreads();
    if (r0 + kTileRows < re) fetch(r0 + kTileRows);
    const double* yt = CROSS ? ys : xs;
#pragma unroll
    for (int k0 = 0; k0 < kTileRows; k0 += 4) {
      const double av = xs[(k0 + q) * LP + 16 * wv + r16];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double bv = yt[(k0 + q) * LP + 16 * b + r16];
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[b], 0, 0, 0);
      }
    }
  }
  double* o = partial + (uint64_t)blockIdx.x * LP * LP;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) o[(16 * wv + q + 4 * r) * LP + 16 * b + r16] = acc[b][r];
}

// Y = X T for the item's rows, T = t (tsel 0) or tv (tsel 1) of the cohort.
// Thread (row slot rs, column pair cp) keeps columns 2 cp, 2 cp + 1 of T in
// registers; X arrives in tiles of 64 rows staged in LDS with coalesced
// loads, and each output sums its 32 products in column order.
static_assert(sizeof(BatchSmall) % 16 == 0 && offsetof(BatchSmall, t) % 16 == 0 && offsetof(BatchSmall, tv) % 16 == 0,
              "the transforms are read as 16-byte pairs");

__global__ void __launch_bounds__(256) k_b_apply(BatchArgs a, const RowItem* __restrict__ items, int xsel, int ysel,
                                                 int tsel, int state_req) {
  constexpr int TR = 64;
  const RowItem it = items[blockIdx.x];
  const BatchCohort& c = a.coh[it.cohort];
  if (c.st->state != state_req) return;
  __shared__ double2 xs[TR * LP / 2];
  const double* src = tsel ? c.st->tv : c.st->t;
  const int cp = threadIdx.x & 15, rs = threadIdx.x >> 4;
  double t0[LP], t1[LP];
#pragma unroll
  for (int m = 0; m < LP; ++m) {
    const double2 tv = *reinterpret_cast<const double2*>(src + m * LP + 2 * cp);
    t0[m] = tv.x;
    t1[m] = tv.y;
  }
  const double* X = c.mat[xsel];
  double* Y = c.mat[ysel];
  for (uint32_t r0 = 0; r0 < it.nrows; r0 += TR) {
    const int nr = (int)min<uint32_t>(TR, it.nrows - r0);
    const double2* xg = reinterpret_cast<const double2*>(X + (it.row0 + r0) * LP);
    __syncthreads();  // the previous tile has been read
    for (int v = threadIdx.x; v < nr * LP / 2; v += 256) xs[v] = xg[v];
    __syncthreads();
    for (int r = rs; r < nr; r += 16) {
      const double2* xr = xs + r * (LP / 2);
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int m2 = 0; m2 < LP / 2; ++m2) {
        const double2 xv = xr[m2];
        y0 += xv.x * t0[2 * m2];
        y1 += xv.x * t1[2 * m2];
        y0 += xv.y * t0[2 * m2 + 1];
        y1 += xv.y * t1[2 * m2 + 1];
      }
      reinterpret_cast<double2*>(Y + (it.row0 + r0 + r) * LP)[cp] = double2{y0, y1};
    }
  }
}

__global__ void k_b_active(BatchArgs a, int* __restrict__ active) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.ncoh; i += gridDim.x * blockDim.x)
    active[i] = a.coh[i].st->state == 1 ? 1 : 0;
}

// ---------------------------------------------------------------- L x L steps
// Matrices in LDS, n x n in the top-left of LP x LP row-major storage.  Every
// helper is called by all threads of the workgroup and ends synchronized.
constexpr int kSmallThreads = 256;

__device__ void sm_zero(double* A) {
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) A[v] = 0.0;
  __syncthreads();
}

__device__ void sm_load(double* A, const double* g) {
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) A[v] = g[v];
  __syncthreads();
}

__device__ void sm_store(double* g, const double* A) {
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) g[v] = A[v];
  __syncthreads();
}

// C = op(A) B with op = transpose when ta (n x n), entries outside n x n zero.
__device__ void sm_mul(double* C, const double* A, const double* B, int n, bool ta) {
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
    const int i = v / LP, j = v % LP;
    double s = 0.0;
    if (i < n && j < n)
      for (int m = 0; m < n; ++m) s += (ta ? A[m * LP + i] : A[i * LP + m]) * B[m * LP + j];
    C[v] = s;
  }
  __syncthreads();
}

// Upper Cholesky factor R (A = R^T R).  Returns false on breakdown (a pivot
// that is not positive and finite), the criterion of the host factorization.
__device__ bool sm_chol(double* R, const double* A, int n, int* flag) {
  sm_zero(R);
  if (threadIdx.x == 0) *flag = 1;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      double d = A[j * LP + j];
      for (int k = 0; k < j; ++k) d -= R[k * LP + j] * R[k * LP + j];
      if (!(d > 0.0) || !isfinite(d)) {
        *flag = 0;
        d = 1.0;
      }
      R[j * LP + j] = sqrt(d);
    }
    __syncthreads();
    const double rjj = R[j * LP + j];
    for (int i = j + 1 + (int)threadIdx.x; i < n; i += kSmallThreads) {
      double s = A[j * LP + i];
      for (int k = 0; k < j; ++k) s -= R[k * LP + j] * R[k * LP + i];
      R[j * LP + i] = s / rjj;
    }
    __syncthreads();
  }
  const bool ok = *flag != 0;
  __syncthreads();
  return ok;
}

// X = R^-1 for upper triangular R, one column per thread.
__device__ void sm_inv_upper(double* X, const double* R, int n) {
  sm_zero(X);
  for (int j = threadIdx.x; j < n; j += kSmallThreads) {
    X[j * LP + j] = 1.0 / R[j * LP + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R[i * LP + k] * X[k * LP + j];
      X[i * LP + j] = -s / R[i * LP + i];
    }
  }
  __syncthreads();
}

// Symmetric eigen-decomposition of A (destroyed) by cyclic Jacobi rotations
// in round-robin order (n/2 disjoint pairs per step); w descending, V columns
// the eigenvectors in the same order.
__device__ void sm_sym_eig(double* A, double* V, double* w, double* tmp, int n, double* rot, int* perm) {
  const int ne = n + (n & 1);  // even number of players; index n (if odd) is a dummy
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) V[v] = (v / LP == v % LP && v / LP < n) ? 1.0 : 0.0;
  // symmetrize
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
    const int i = v / LP, j = v % LP;
    tmp[v] = (i < n && j < n) ? 0.5 * (A[i * LP + j] + A[j * LP + i]) : 0.0;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) A[v] = tmp[v];
  __syncthreads();
  __shared__ double offn;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (threadIdx.x == 0) {
      double off = 0.0, tot = 0.0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const double x = A[i * LP + j] * A[i * LP + j];
          tot += x;
          if (i != j) off += x;
        }
      offn = (off <= 1e-30 * tot || tot == 0.0) ? -1.0 : off;  // off-norm <= 1e-15 of the norm
    }
    __syncthreads();
    if (offn < 0.0) break;
    for (int step = 0; step < ne - 1; ++step) {
      // pairs of the round-robin schedule: player 0 fixed, others rotate
      for (int k = threadIdx.x; k < ne / 2; k += kSmallThreads) {
        int p = k == 0 ? 0 : 1 + (k - 1 + step) % (ne - 1);
        int q = 1 + (ne - 2 - k + step) % (ne - 1);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p * LP + q];
          if (apq != 0.0) {
            const double tau = (A[q * LP + q] - A[p * LP + p]) / (2.0 * apq);
            const double at = fabs(tau);
            // t = sign(tau) / (|tau| + sqrt(1 + tau^2)), ~ 1 / (2 tau) where tau^2 would overflow
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (at > 1e150 ? 2.0 * at : at + sqrt(1.0 + at * at));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        perm[2 * k] = p;
        perm[2 * k + 1] = q;
        rot[2 * k] = c;
        rot[2 * k + 1] = s;
      }
      __syncthreads();
      // A <- A J (column pairs), V <- V J
      for (int v = threadIdx.x; v < (ne / 2) * n; v += kSmallThreads) {
        const int k = v / n, i = v % n;
        const int p = perm[2 * k], q = perm[2 * k + 1];
        if (q >= n) continue;
        const double c = rot[2 * k], s = rot[2 * k + 1];
        const double aip = A[i * LP + p], aiq = A[i * LP + q];
        A[i * LP + p] = c * aip - s * aiq;
        A[i * LP + q] = s * aip + c * aiq;
        const double vip = V[i * LP + p], viq = V[i * LP + q];
        V[i * LP + p] = c * vip - s * viq;
        V[i * LP + q] = s * vip + c * viq;
      }
      __syncthreads();
      // A <- J^T A (row pairs)
      for (int v = threadIdx.x; v < (ne / 2) * n; v += kSmallThreads) {
        const int k = v / n, j = v % n;
        const int p = perm[2 * k], q = perm[2 * k + 1];
        if (q >= n) continue;
        const double c = rot[2 * k], s = rot[2 * k + 1];
        const double apj = A[p * LP + j], aqj = A[q * LP + j];
        A[p * LP + j] = c * apj - s * aqj;
        A[q * LP + j] = s * apj + c * aqj;
      }
      __syncthreads();
    }
  }
  // sort descending (stable selection on one thread), permute V's columns
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
      int b = i;
      for (int j = i + 1; j < n; ++j)
        if (A[perm[j] * LP + perm[j]] > A[perm[b] * LP + perm[b]]) b = j;
      const int t = perm[i];
      perm[i] = perm[b];
      perm[b] = t;
    }
    for (int i = 0; i < n; ++i) w[i] = A[perm[i] * LP + perm[i]];
  }
  __syncthreads();
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
    const int i = v / LP, j = v % LP;
    tmp[v] = (i < n && j < n) ? V[i * LP + perm[j]] : 0.0;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) V[v] = tmp[v];
  __syncthreads();
}

// Householder-convention column signs of the thin Q whose top n x n block is
// Y (destroyed): a modified LU without pivoting (Ballard et al., 2014).
__device__ void sm_householder_signs(double* W, int n, double* s) {
  for (int i = 0; i < n; ++i) {
    if (threadIdx.x == 0) {
      s[i] = W[i * LP + i] >= 0.0 ? -1.0 : 1.0;
      W[i * LP + i] -= s[i];
    }
    __syncthreads();
    const double piv = W[i * LP + i];
    for (int r = i + 1 + (int)threadIdx.x; r < n; r += kSmallThreads) {
      W[r * LP + i] /= piv;
      const double lr = W[r * LP + i];
      if (lr == 0.0) continue;
      for (int c = i + 1; c < n; ++c) W[r * LP + c] -= lr * W[i * LP + c];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kSmallThreads) k_b_small(BatchArgs a, int mode, int epoch) {
  extern __shared__ __align__(16) double smem[];
  double* M0 = smem;
  double* M1 = M0 + LP * LP;
  double* M2 = M1 + LP * LP;
  double* M3 = M2 + LP * LP;
  double* M4 = M3 + LP * LP;
  double* M5 = M4 + LP * LP;
  double* vec = M5 + LP * LP;   // 2 LP
  double* rot = vec + 2 * LP;   // LP
  __shared__ int flag;
  __shared__ int perm[LP];
  const BatchCohort& c = a.coh[blockIdx.x];
  BatchSmall* st = c.st;
  const int state = st->state;
  const int L = a.L, k = a.k;
  if (mode == kSmallFinish) {
    if (threadIdx.x == 0 && state == 2) st->state = 0;
    return;
  }
  if (mode == kSmallInit) {
    for (int l = threadIdx.x; l < LP; l += kSmallThreads) st->colsign[l] = 1.0;
    for (int e = threadIdx.x; e < kBatchMaxEpochs; e += kSmallThreads) st->trace[e] = 0.0;
    if (threadIdx.x == 0) {
      st->state = 1;
      st->epochs = 0;
      st->stop_epoch = 0;
      st->nsv = 0;
    }
    return;
  }
  if (state != 1) return;
  if (mode == kSmallEpoch) {
    // R = R2 R1 from the Gram of G~ (CholeskyQR2, second factor on the small side)
    sm_load(M0, st->gram_a);
    if (!sm_chol(M1, M0, L, &flag)) {
      if (threadIdx.x == 0) st->state = -5;
      return;
    }
    sm_inv_upper(M2, M1, L);
    sm_mul(M3, M0, M2, L, false);  // Ga R1^-1
    sm_mul(M4, M2, M3, L, true);   // R1^-T Ga R1^-1
    if (!sm_chol(M3, M4, L, &flag)) {
      if (threadIdx.x == 0) st->state = -5;
      return;
    }
    sm_mul(M4, M3, M1, L, false);  // R = R2 R1
    sm_inv_upper(M0, M4, L);       // R^-1
    // C = R^-T (H^T H) R^-1
    sm_load(M1, st->gram_h);
    sm_mul(M2, M1, M0, L, false);
    sm_mul(M1, M0, M2, L, true);
    sm_sym_eig(M1, M2, vec, M3, L, rot, perm);  // M2 = W, vec = eigenvalues (descending)
    for (int i = threadIdx.x; i < LP; i += kSmallThreads)
      st->sigma[i] = i < L ? sqrt(vec[i] > 0.0 ? vec[i] : 0.0) : 0.0;
    sm_mul(M3, M0, M2, L, false);  // R^-1 W
    for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
      const int i = v / LP, j = v % LP;
      const double rw = (i < L && j < k) ? M3[v] : 0.0;
      const double sg = j < L ? sqrt(vec[j] > 0.0 ? vec[j] : 0.0) : 0.0;
      st->tv[v] = rw;
      st->t[v] = sg > 0.0 ? rw / sg : 0.0;
    }
  } else if (mode == kSmallStop) {
    if (threadIdx.x == 0) {
      bool stop = false;
      if (epoch > 0) {
        // 1 - mev(Ucur, Upre), mev = (1/k) sum_i || Ucur^T upre_i ||
        double m = 0.0;
        for (int i = 0; i < k; ++i) {
          double s2 = 0.0;
          for (int r = 0; r < k; ++r) s2 += st->gram_a[r * LP + i] * st->gram_a[r * LP + i];
          m += sqrt(s2);
        }
        const double diff = 1.0 - m / k;
        if (epoch < kBatchMaxEpochs) st->trace[epoch] = diff;
        stop = diff < a.tol || epoch == a.p;
        if (stop) st->nsv = k;
      } else if (a.p == 0) {
        stop = true;
        st->nsv = 0;
      }
      st->epochs = epoch + 1;
      if (stop) {
        st->state = 2;
        st->stop_epoch = epoch;
      }
    }
  } else if (mode == kSmallOmega1) {
    sm_load(M0, st->gram_h);
    if (!sm_chol(M1, M0, L, &flag)) {
      if (threadIdx.x == 0) st->state = -5;
      return;
    }
    sm_store(st->r1h, M1);
    sm_inv_upper(M2, M1, L);
    sm_store(st->t, M2);
  } else if (mode == kSmallOmega2) {
    sm_load(M0, st->gram_a);  // Gram of Q1 = H R1^-1
    if (!sm_chol(M1, M0, L, &flag)) {
      if (threadIdx.x == 0) st->state = -5;
      return;
    }
    sm_load(M2, st->r1h);
    sm_mul(M3, M1, M2, L, false);  // R2 R1
    sm_inv_upper(M0, M3, L);       // Rinv
    // top L x L block of Q = H Rinv
    const double* H = c.mat[kBufH];
    for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
      const int i = v / LP, j = v % LP;
      M2[v] = (i < L && j < L) ? H[(uint64_t)i * LP + j] : 0.0;
    }
    __syncthreads();
    sm_mul(M4, M2, M0, L, false);
    sm_householder_signs(M4, L, vec);
    for (int v = threadIdx.x; v < LP * LP; v += kSmallThreads) {
      const int j = v % LP;
      st->t[v] = j < L ? M0[v] * vec[j] : 0.0;
    }
  } else if (mode == kSmallFlip) {
    for (int l = threadIdx.x; l < LP; l += kSmallThreads)
      st->colsign[l] = (l < L && st->ab[l] > 2.0 * st->ab[LP + l]) ? -1.0 : 1.0;
  }
}

unsigned nblocks(uint32_t n) { return n ? n : 1; }

}  // namespace

void launch_batch_small(const BatchArgs& a, int mode, int epoch, hipStream_t s) {
  constexpr size_t lds = (6 * LP * LP + 3 * LP) * sizeof(double);
  static bool attr = false;
  if (!attr) {
    PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_b_small),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  hipLaunchKernelGGL(k_b_small, dim3(a.ncoh), dim3(kSmallThreads), lds, s, a, mode, epoch);
  check_launch("batch_small");
}

void launch_batch_make_pieces(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, int src, bool signed_,
                              hipStream_t s) {
  if (!nitems) return;
  const int sg = signed_ ? 1 : 0;
  switch ((Form)a.form) {
    case Form::FP64: hipLaunchKernelGGL(k_b_make_pieces<Form::FP64>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::FP32: hipLaunchKernelGGL(k_b_make_pieces<Form::FP32>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::BF16: hipLaunchKernelGGL(k_b_make_pieces<Form::BF16>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::BF16X1: hipLaunchKernelGGL(k_b_make_pieces<Form::BF16X1>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::FP16: hipLaunchKernelGGL(k_b_make_pieces<Form::FP16>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::F8: hipLaunchKernelGGL(k_b_make_pieces<Form::F8>, dim3(nitems), dim3(256), 0, s, a, items_n, src, sg); break;
    case Form::I8: fail(kErrArgument, "batch: no int8 form");
  }
  check_launch("batch_make_pieces");
}

void launch_batch_omega_colmax(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, int src, hipStream_t s) {
  if (!nitems) return;
  hipLaunchKernelGGL(k_b_zero_cmax, dim3(nblocks(a.ncoh)), dim3(LP), 0, s, a, 0);
  hipLaunchKernelGGL(k_b_colmax, dim3(nitems), dim3(256), 0, s, a, items_n, src);
  check_launch("batch_omega_colmax");
}

void launch_batch_colsum(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, double* partial,
                         hipStream_t s) {
  if (!nitems) return;
  hipLaunchKernelGGL(k_b_rowsum<false>, dim3(nitems), dim3(256), 0, s, a, items_n, partial);
  hipLaunchKernelGGL(k_b_sum_items, dim3(nblocks(a.ncoh)), dim3(256), 0, s, a, partial, LP, (int)kListN,
                     (int)kDestColsum, 1);
  check_launch("batch_colsum");
}

void launch_batch_flip(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, double* partial,
                       hipStream_t s) {
  if (!nitems) return;
  hipLaunchKernelGGL(k_b_rowsum<true>, dim3(nitems), dim3(256), 0, s, a, items_n, partial);
  hipLaunchKernelGGL(k_b_sum_items, dim3(nblocks(a.ncoh)), dim3(256), 0, s, a, partial, 2 * LP, (int)kListN,
                     (int)kDestAb, 1);
  check_launch("batch_flip");
}

void launch_batch_epi_g(const BatchArgs& a, const RowItem* items_e, uint32_t nitems, double* partial,
                        hipStream_t s) {
  if (!nitems) return;
  const dim3 g(nitems), b(256);
  switch ((Form)a.form) {
    case Form::FP64: hipLaunchKernelGGL((k_b_epi_g<Form::FP64, double, 0>), g, b, 0, s, a, items_e, partial); break;
    case Form::FP32: hipLaunchKernelGGL((k_b_epi_g<Form::FP32, float, 0>), g, b, 0, s, a, items_e, partial); break;
    case Form::BF16: hipLaunchKernelGGL((k_b_epi_g<Form::BF16, float, 0>), g, b, 0, s, a, items_e, partial); break;
    case Form::BF16X1: hipLaunchKernelGGL((k_b_epi_g<Form::BF16X1, float, 0>), g, b, 0, s, a, items_e, partial); break;
    case Form::FP16: hipLaunchKernelGGL((k_b_epi_g<Form::FP16, float, 0>), g, b, 0, s, a, items_e, partial); break;
    case Form::F8:
      hipLaunchKernelGGL(k_b_zero_cmax, dim3(nblocks(a.ncoh)), dim3(LP), 0, s, a, 1);
      hipLaunchKernelGGL((k_b_epi_g<Form::F8, float, 1>), g, b, 0, s, a, items_e, partial);
      hipLaunchKernelGGL((k_b_epi_g<Form::F8, float, 2>), g, b, 0, s, a, items_e, partial);
      break;
    case Form::I8: fail(kErrArgument, "batch: no int8 form");
  }
  hipLaunchKernelGGL(k_b_sum_items, dim3(nblocks(a.ncoh)), dim3(256), 0, s, a, partial, LP, (int)kListE,
                     (int)kDestBvec, 1);
  check_launch("batch_epi_g");
}

void launch_batch_epi_h(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, hipStream_t s) {
  if (!nitems) return;
  if ((Form)a.form == Form::FP64)
    hipLaunchKernelGGL(k_b_epi_h<double>, dim3(nitems), dim3(256), 0, s, a, items_n);
  else
    hipLaunchKernelGGL(k_b_epi_h<float>, dim3(nitems), dim3(256), 0, s, a, items_n);
  check_launch("batch_epi_h");
}

void launch_batch_gram(const BatchArgs& a, const RowItem* items, uint32_t nitems, bool m_rows, int xsel, int ysel,
                       int dest, int state_req, double* partial, hipStream_t s) {
  if (!nitems) return;
  if (ysel < 0)
    hipLaunchKernelGGL(k_b_gram<false>, dim3(nitems), dim3(128), 0, s, a, items, xsel, xsel, state_req, partial);
  else
    hipLaunchKernelGGL(k_b_gram<true>, dim3(nitems), dim3(128), 0, s, a, items, xsel, ysel, state_req, partial);
  hipLaunchKernelGGL(k_b_sum_items, dim3(nblocks(a.ncoh)), dim3(256), 0, s, a, partial, LP * LP,
                     m_rows ? (int)kListM : (int)kListN, dest == 0 ? (int)kDestGramA : (int)kDestGramH, state_req);
  check_launch("batch_gram");
}

void launch_batch_apply(const BatchArgs& a, const RowItem* items, uint32_t nitems, int xsel, int ysel, int tsel,
                        int state_req, hipStream_t s) {
  if (!nitems) return;
  hipLaunchKernelGGL(k_b_apply, dim3(nitems), dim3(256), 0, s, a, items, xsel, ysel, tsel, state_req);
  check_launch("batch_apply");
}

void launch_batch_active(const BatchArgs& a, int* active, hipStream_t s) {
  hipLaunchKernelGGL(k_b_active, dim3((a.ncoh + 255) / 256), dim3(256), 0, s, a, active);
  check_launch("batch_active");
}

}  // namespace pcaone
