// Missing-entry corrections on tiles of the missing-entry lists (the reduced
// forms).  With missing genotypes the packed products run on the codes as for
// complete data (a missing code decodes to 0) and a sparse product adds, per
// target row r, sum over r's missing entries e = (r, k) of w_e * src[k]: the
// centering that the affine epilogue gave the entry, cancelled, plus its EMU
// imputation (src/Data.cpp:318-348, src/Halko.cpp:129-152,227-268).
//
// Gathering src[k] per entry from L2 is bound by the vector-memory request
// rate (one scattered row per entry).  Here a list is re-laid in tiles: a row
// block of kMissRB list rows times an index range whose entries fit one
// staging buffer.  A workgroup owns a row block; for each of its tiles it
// stages the row offsets, the entries (16-bit local index, FP32 weight) and
// the tile's source rows into LDS (double buffered; the next tile's loads are
// in flight in registers while the current one is processed), and each lane
// group of cp / 4 lanes walks the entries of the rows it owns, reading source
// rows from LDS.  Every sum runs in a fixed order (tiles, then entries, then the
// parts of a row block in part order), so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"

namespace pcaone {

namespace {

constexpr int kMissOffBytes = kMissRowStride * 2;  // row offsets at the start of a staging buffer

// Entry counts per (row block, granule); one wave per list row.  A row's
// indexes ascend, so each run of equal granules within 64 entries is counted
// by its first lane with one atomic.
__global__ void __launch_bounds__(256) k_miss_hist(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ ix,
                                                   uint64_t nrows, const uint32_t* __restrict__ rpos,
                                                   uint32_t ngran, uint32_t* __restrict__ cnt) {
  const uint64_t r = blockIdx.x * 4ull + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;  // wave-uniform
  uint32_t* c = cnt + ((rpos ? rpos[r] : r) / kMissRB) * (uint64_t)ngran;
  const uint64_t e1 = rp[r + 1];
  for (uint64_t eb = rp[r]; eb < e1; eb += 64) {
    const int nb = (int)min<uint64_t>(64, e1 - eb);
    const bool v = lane < nb;
    const uint32_t gi = v ? ix[eb + lane] / kMissKG : 0u;
    const uint32_t prev = __shfl_up(gi, 1);
    const bool head = v && (lane == 0 || prev != gi);
    const uint64_t hm = __ballot(head);
    if (head) {
      const uint64_t after = lane == 63 ? 0ull : (hm >> (lane + 1));
      const int next = after ? lane + 1 + __builtin_ctzll(after) : 64;
      atomicAdd(c + gi, (uint32_t)(min(next, nb) - lane));
    }
  }
}

// Row offsets and local indexes of every tile of a row block: one thread per
// row walks its CSR list through the block's tiles in order; the offsets are
// an exclusive scan of the rows' counts.  A tile whose rows hold more entries
// than planned is not written past its end and sets *status.
__global__ void __launch_bounds__(kMissRB) k_miss_fill(const uint64_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ ix, uint64_t nrows,
                                                      const uint32_t* __restrict__ rperm,
                                                      const MissTile* __restrict__ tiles,
                                                      const uint32_t* __restrict__ tb, uint16_t* __restrict__ trow,
                                                      uint16_t* __restrict__ eix, int* __restrict__ status) {
  __shared__ uint32_t wsum[kMissRB / 64];
  const uint64_t b = blockIdx.x;
  const int i = threadIdx.x, lane = i & 63, wave = i >> 6;
  const uint64_t r = rperm ? (uint64_t)rperm[b * kMissRB + i] : b * kMissRB + i;
  uint64_t cur = 0, end = 0;
  if (r < nrows) {
    cur = rp[r];
    end = rp[r + 1];
  }
  const uint32_t t1 = tb[b + 1];
  for (uint32_t t = tb[b]; t < t1; ++t) {
    const MissTile T = tiles[t];
    const uint64_t c0 = cur;
    while (cur < end && ix[cur] < T.k1) ++cur;
    const uint32_t n = (uint32_t)(cur - c0);
    uint32_t x = n;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    const uint32_t excl = base + x - n;
    uint16_t* ro = trow + (uint64_t)t * kMissRowStride;
    if (i == kMissRB - 1 && excl + n != T.ne) *status = 1;
    const uint32_t lo = min(excl, T.ne), hi = min(excl + n, T.ne);
    ro[i] = (uint16_t)lo;
    if (i == kMissRB - 1) ro[kMissRB] = (uint16_t)hi;
    for (uint32_t j = lo; j < hi; ++j) eix[T.e0 + j] = (uint16_t)(ix[c0 + (j - excl)] - T.k0);
    __syncthreads();  // wsum is reused by the next tile
  }
}

// Sparse correction on tiles.  Lane group g of a wave (LPE = cp / 4 lanes, 4
// FP32 columns each) owns rows RPG g' .. RPG g' + RPG - 1 of the row block (g'
// = the group's index in the workgroup).  Per tile the staging buffer holds,
// in 16-byte slots: the row offsets, the local indexes, the weights and the
// source rows of [max(k0, k_lo), min(k1, k_hi)) at their offset from k0.
template <int RPG>
__global__ void __launch_bounds__(kMissThreads) k_missing_spmm_tiled(MissTiledSpmmArgs a, int lpe) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int G = 64 / lpe, g = lane / lpe, cl = lane - g * lpe;
  const bool act = g < G;
  const int row0 = (wave * G + g) * RPG;
  const MissWork wk = a.work[blockIdx.x];
  const uint64_t b = wk.b;
  const uint32_t t0 = wk.t0, t1 = wk.t1;
  const int cpv = a.cp / 4;  // 16-byte slots per source row

  // A staging buffer is filled from registers: each thread loads its
  // 16-byte slots of the next tile before processing the current one and
  // writes them to LDS after it, so the loads' latency is covered by the
  // processing.  Tile descriptors are likewise read a step ahead.  The
  // buffer's regions sit at fixed slots (row offsets, indexes, weights,
  // source rows; sized for kMissEMax entries), so each thread decodes its
  // slots once: a region, a byte offset within the region's tile data and,
  // for a source row slot, the row it belongs to.
  constexpr int kSlots = (int)(kMissLdsBuf / 16 / kMissThreads);  // slots per thread
  constexpr uint32_t sB = kMissOffBytes / 16, sC = sB + miss_stored(kMissEMax) / 8,
                     sD = sC + miss_stored(kMissEMax) / 4;
  int reg_of[kSlots];      // 0 row offsets, 1 indexes, 2 weights, 3 source rows
  uint32_t off_of[kSlots]; // byte offset within the region's data
  uint32_t row_of[kSlots]; // source row (region 3), or the slot within its region
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    const uint32_t s = tid + i * kMissThreads;
    if (s < sB) {
      reg_of[i] = 0, row_of[i] = s, off_of[i] = s * 16;
    } else if (s < sC) {
      reg_of[i] = 1, row_of[i] = s - sB, off_of[i] = (s - sB) * 16;
    } else if (s < sD) {
      reg_of[i] = 2, row_of[i] = s - sC, off_of[i] = (s - sC) * 16;
    } else {
      const uint32_t d = s - sD, j = d / cpv, p = d - j * cpv;
      reg_of[i] = 3, row_of[i] = j, off_of[i] = (j * (uint32_t)a.cpitch + p * 4) * 4;
    }
  }
  uint4 reg[kSlots];
  auto load = [&](uint32_t t, const MissTile& T) {
    const uint32_t nb = miss_stored(T.ne) / 8, nc = miss_stored(T.ne) / 4;
    const uint32_t jlo = (uint32_t)(max<uint64_t>(T.k0, a.k_lo) - T.k0);
    const uint32_t jhi = (uint32_t)(min<uint64_t>(T.k1, a.k_hi) - T.k0);
    const uint64_t pa = reinterpret_cast<uint64_t>(a.t.trow + (uint64_t)t * kMissRowStride);
    const uint64_t pb = reinterpret_cast<uint64_t>(a.t.eix + T.e0);
    const uint64_t pc = reinterpret_cast<uint64_t>(a.ew + T.e0);
    // source row k of the tile is compact row k - src_k0 (>= 0 where staged);
    // the base is formed in integers, only rows at or above src_k0 are read
    const uint64_t pd = reinterpret_cast<uint64_t>(a.src) +
                        (uint64_t)((int64_t)T.k0 - (int64_t)a.src_k0) * (uint64_t)a.cpitch * 4ull;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
      const int rg = reg_of[i];
      const uint32_t r = row_of[i];
      const bool live = rg == 0 || (rg == 1 && r < nb) || (rg == 2 && r < nc) || (rg == 3 && r >= jlo && r < jhi);
      const uint64_t base = rg == 0 ? pa : (rg == 1 ? pb : (rg == 2 ? pc : pd));
      using gptr = const __attribute__((address_space(1))) uint32_t*;  // global loads, not flat ones
      uint4 v{0u, 0u, 0u, 0u};
      if (live) {
        const gptr g = reinterpret_cast<gptr>(base + off_of[i]);
        v = uint4{g[0], g[1], g[2], g[3]};
      }
      reg[i] = v;
    }
  };
  // The local indexes are stored as the 16-byte slot offsets of their source
  // rows (k cp / 4 < 65536 for the spans planned).  Slots the tile does not
  // fill are written too (zeros); nothing reads them.
  auto store = [&](unsigned char* buf) {
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
      uint4 v = reg[i];
      if (reg_of[i] == 1) {
        uint32_t* h = &v.x;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          h[j] = (__umul24(h[j] & 0xffffu, cpv) & 0xffffu) | (__umul24(h[j] >> 16, cpv) << 16);
      }
      *reinterpret_cast<uint4*>(buf + (uint64_t)(tid + i * kMissThreads) * 16) = v;
    }
  };

  double acc[RPG][4];
#pragma unroll
  for (int q = 0; q < RPG; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[q][c] = 0.0;

  // A row's entries in the tile are taken U at a time.  Reads past the row's
  // end meet the next rows' entries or the tile's zero padding, all indexes
  // of staged rows, so the last, partial step only zeroes their weights; in a
  // tile that the index range cuts, every step also moves entries outside the
  // range to the source row at `olo` (always staged).  The row's sum within
  // the tile is taken in FP32 and added to its FP64 total.
  auto process = [&](const unsigned char* buf, const MissTile& T, auto edge_tag) {
    constexpr bool EDGE = decltype(edge_tag)::value;
    const uint16_t* ro = reinterpret_cast<const uint16_t*>(buf);
    const uint16_t* ix = reinterpret_cast<const uint16_t*>(buf + sB * 16);
    const float* w = reinterpret_cast<const float*>(buf + sC * 16);
    const unsigned char* src = buf + sD * 16 + cl * 16;
    const uint32_t olo = EDGE ? (uint32_t)(max<uint64_t>(T.k0, a.k_lo) - T.k0) * cpv : 0u;
    const uint32_t ohi = EDGE ? (uint32_t)(min<uint64_t>(T.k1, a.k_hi) - T.k0) * cpv : 0u;
    constexpr int U = RPG >= 3 ? 4 : 8;  // three rows' sums leave registers for four entries a step
    auto step = [&](int e, int ee, float* f, bool masked) {
      uint32_t o[U];
      float wv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        o[u] = ix[e + u];
        wv[u] = w[e + u];
        if (EDGE) {
          const bool ok = e + u < ee && o[u] >= olo && o[u] < ohi;
          o[u] = ok ? o[u] : olo;
          wv[u] = ok ? wv[u] : 0.0f;
        } else if (masked) {
          wv[u] = e + u < ee ? wv[u] : 0.0f;
        }
      }
      float4 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = *reinterpret_cast<const float4*>(src + (o[u] << 4));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        f[0] = fmaf(wv[u], x[u].x, f[0]);
        f[1] = fmaf(wv[u], x[u].y, f[1]);
        f[2] = fmaf(wv[u], x[u].z, f[2]);
        f[3] = fmaf(wv[u], x[u].w, f[3]);
      }
    };
#pragma unroll
    for (int q = 0; q < RPG; ++q) {
      const int r = row0 + q;
      int e = 0, ee = 0;
      if (act && r < kMissRB) {
        e = ro[r];
        ee = ro[r + 1];
      }
      float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (!EDGE)
        for (; e + U <= ee; e += U) step(e, ee, f, false);
      for (; e < ee; e += U) step(e, ee, f, true);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[q][c] += (double)f[c];
    }
  };

  if (t0 < t1) {
    unsigned char* buf0 = lds;
    unsigned char* buf1 = lds + kMissLdsBuf;
    MissTile Tc = a.t.tiles[t0];
    MissTile Tn = a.t.tiles[min(t0 + 1, t1 - 1)];
    load(t0, Tc);
    store(buf0);
    __syncthreads();
    int cur = 0;
    for (uint32_t t = t0; t < t1; ++t) {
      const MissTile Tnn = a.t.tiles[min(t + 2, t1 - 1)];
      const bool more = t + 1 < t1;
      if (more) load(t + 1, Tn);
      if (Tc.k0 < a.k_lo || Tc.k1 > a.k_hi) process(cur ? buf1 : buf0, Tc, std::true_type{});
      else process(cur ? buf1 : buf0, Tc, std::false_type{});
      // the other buffer was last read before the previous barrier
      if (more) store(cur ? buf0 : buf1);
      __syncthreads();
      cur ^= 1;
      Tc = Tn;
      Tn = Tnn;
    }
  }

  const bool partial = a.slots > 1;
  const uint64_t nr = a.r_hi - a.r_lo;
#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const int r = row0 + q;
    if (!act || r >= kMissRB) continue;
    const uint64_t pos = b * kMissRB + r;
    const uint64_t gr = a.t.rperm ? (uint64_t)a.t.rperm[pos] : pos;
    if (gr < a.r_lo || gr >= a.r_hi) continue;
    const uint64_t orow = gr - a.r_lo;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = cl * 4 + c;
      if (col >= a.ncols) continue;
      if (partial) a.partial[((uint64_t)wk.slot * nr + orow) * a.ncols + col] = acc[q][c];
      else a.out[orow * a.lp + col] += acc[q][c];
    }
  }
}

// Weights of a tiled list, one thread per entry.  w = -sc_c (1/2 - cF_c) plus,
// with EMU, sc_c clamp((U S)_i . V_c, -F_c, 1 - F_c), where c is the entry's
// SNP and i its sample.  EMU: the row block's k-vectors (V rows with their
// SNP's sc, cF, F for the by-SNP list; U S rows for the by-sample list) are
// staged once, a tile's index-side vectors `sub` rows at a time.
template <bool EMU>
__global__ void __launch_bounds__(kMissThreads) k_missing_weights_tiled(MissTiledWeightArgs a, uint32_t sub) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const MissWork wk = a.work[blockIdx.x];
  const uint64_t b = wk.b;
  const int kk = a.kk;
  const int vr = a.by_snp ? kk + 3 : kk;  // doubles per row-side vector
  const int vi = a.by_snp ? kk : kk + 3;  // doubles per index-side vector
  double* rowvec = reinterpret_cast<double*>(lds);
  uint16_t* ro = reinterpret_cast<uint16_t*>(lds + (EMU ? (size_t)kMissRB * vr * 8 : 0));
  uint16_t* eix = ro + kMissRowStride;
  uint16_t* erow = eix + kMissEMax;
  double* ivec = reinterpret_cast<double*>(erow + kMissEMax);
  if (EMU) {
    const double* rs = a.by_snp ? a.V : a.US;
    for (int t = tid; t < kMissRB * vr; t += kMissThreads) {
      const int i = t / vr, c = t - i * vr;
      const uint64_t pos = b * kMissRB + i;
      const uint64_t r = a.t.rperm ? (uint64_t)a.t.rperm[pos] : pos;
      rowvec[t] = r < a.nrows ? rs[r * a.kp + c] : 0.0;
    }
  }
  for (uint32_t t = wk.t0; t < wk.t1; ++t) {
    const MissTile T = a.t.tiles[t];
    __syncthreads();  // the previous tile's buffers are free
    for (int i = tid; i <= kMissRB; i += kMissThreads) ro[i] = a.t.trow[(uint64_t)t * kMissRowStride + i];
    for (uint32_t e = tid; e < T.ne; e += kMissThreads) eix[e] = a.t.eix[T.e0 + e];
    __syncthreads();
    for (int i = tid; i < kMissRB; i += kMissThreads)
      for (int e = ro[i]; e < ro[i + 1]; ++e) erow[e] = (uint16_t)i;
    const uint32_t span = T.k1 - T.k0;
    const uint32_t step = EMU ? sub : span;
    for (uint32_t s0 = 0; s0 < span; s0 += step) {
      const uint32_t s1 = min(span, s0 + step);
      if (EMU) {
        __syncthreads();  // the previous pass has read ivec
        const double* is = a.by_snp ? a.US : a.V;
        for (uint32_t t2 = tid; t2 < (s1 - s0) * vi; t2 += kMissThreads) {
          const uint32_t j = t2 / vi, c = t2 - j * vi;
          ivec[t2] = is[(uint64_t)(T.k0 + s0 + j) * a.kp + c];
        }
      }
      __syncthreads();
      for (uint32_t e = tid; e < T.ne; e += kMissThreads) {
        const uint32_t x = eix[e];
        if (x < s0 || x >= s1) continue;
        const int i = erow[e];
        double w;
        if (EMU) {
          const double* rv = rowvec + i * vr;
          const double* iv = ivec + (x - s0) * vi;
          double p = 0.0;
          for (int j = 0; j < kk; ++j) p += rv[j] * iv[j];
          const double* st = a.by_snp ? rv + kk : iv + kk;
          const double sc = st[0], cf = st[1], f = st[2];
          w = -sc * (0.5 - cf) + sc * fmin(fmax(p, -f), 1.0 - f);
        } else {
          const uint64_t pos = b * kMissRB + i;
          const uint64_t c = a.by_snp ? (a.t.rperm ? (uint64_t)a.t.rperm[pos] : pos) : (uint64_t)T.k0 + x;
          w = -a.sc[c] * (0.5 - a.cf[c]);
        }
        a.ew[T.e0 + e] = (float)w;
      }
    }
  }
}

// out[r][l] += sum of the first nslot partial rows of r's block, in slot order.
__global__ void k_sum_slots(const double* __restrict__ partial, uint64_t nr, int ncols, uint64_t r_lo,
                            const uint32_t* __restrict__ rpos, const uint32_t* __restrict__ bslots,
                            double* __restrict__ out, int lp) {
  const uint64_t n = nr * (uint64_t)ncols;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = t / ncols;
    const int l = (int)(t - r * ncols);
    const uint64_t pos = rpos ? rpos[r_lo + r] : r_lo + r;
    const uint32_t ns = bslots[pos / kMissRB - r_lo / kMissRB];
    double v = 0.0;
    for (uint32_t z = 0; z < ns; ++z) v += partial[z * n + t];
    out[r * lp + l] += v;
  }
}

template <class K>
void allow_lds(K kernel, size_t bytes) {
  PCAONE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

}  // namespace

uint32_t miss_tile_span(int cp) {
  const uint32_t fixed = kMissOffBytes + miss_stored(kMissEMax) * 6;
  const uint32_t rows = (kMissLdsBuf - fixed) / (uint32_t)(cp * 4);
  return rows / kMissKG * kMissKG;
}

void launch_miss_hist(const uint64_t* rp, const uint32_t* ix, uint64_t nrows, const uint32_t* rpos, uint32_t ngran,
                      uint32_t* cnt, hipStream_t s) {
  if (!nrows) return;
  hipLaunchKernelGGL(k_miss_hist, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, s, rp, ix, nrows, rpos, ngran,
                     cnt);
  check_launch("miss_hist");
}

void launch_miss_fill(const uint64_t* rp, const uint32_t* ix, uint64_t nrows, const uint32_t* rperm,
                      const MissTile* tiles, const uint32_t* tb, uint64_t nblocks, uint16_t* trow, uint16_t* eix,
                      int* status, hipStream_t s) {
  if (!nblocks) return;
  hipLaunchKernelGGL(k_miss_fill, dim3((unsigned)nblocks), dim3(kMissRB), 0, s, rp, ix, nrows, rperm, tiles, tb,
                     trow, eix, status);
  check_launch("miss_fill");
}

void launch_missing_weights_tiled(const MissTiledWeightArgs& a, hipStream_t s) {
  if (!a.nblocks) return;
  const bool emu = a.Fclamp != nullptr;
  const size_t cap = kMissLdsBuf;  // two workgroups per CU: one stages a tile while the other computes
  size_t lds = kMissOffBytes + 2 * (size_t)kMissEMax * 2;
  uint32_t sub = 0;
  if (emu) {
    require(a.kk >= 1 && a.kk <= kMissMaxCp && a.kp >= a.kk + 3 && a.US && a.V, kErrArgument,
            "missing_weights_tiled: imputation rank must be in [1, 32]");
    const int vr = a.by_snp ? a.kk + 3 : a.kk, vi = a.by_snp ? a.kk : a.kk + 3;
    lds += (size_t)kMissRB * vr * 8;
    require(lds + (size_t)vi * 8 <= cap, kErrInternal, "missing_weights_tiled: k-vectors exceed LDS");
    sub = (uint32_t)((cap - lds) / ((size_t)vi * 8));
    lds += (size_t)sub * vi * 8;
  }
  if (!a.nwork) return;
  require(a.work != nullptr, kErrInternal, "missing_weights_tiled: no work items");
  const dim3 grid(a.nwork);
  if (emu) {
    static bool once = (allow_lds(k_missing_weights_tiled<true>, cap), true);
    (void)once;
    hipLaunchKernelGGL(k_missing_weights_tiled<true>, grid, dim3(kMissThreads), lds, s, a, sub);
  } else {
    hipLaunchKernelGGL(k_missing_weights_tiled<false>, grid, dim3(kMissThreads), lds, s, a, sub);
  }
  check_launch("missing_weights_tiled");
}

void launch_missing_spmm_tiled(const MissTiledSpmmArgs& a, hipStream_t s) {
  if (a.r_hi <= a.r_lo || a.k_hi <= a.k_lo || !a.nwork) return;
  require(a.cp >= 4 && a.cp % 4 == 0 && a.cp <= kMissMaxCp && a.ncols >= 1 && a.ncols <= a.cp &&
              a.cpitch >= a.cp && a.cpitch % 4 == 0 && a.ncols <= a.lp,
          kErrArgument, "missing_spmm_tiled: bad column counts");
  const uint64_t nr = a.r_hi - a.r_lo;
  require(a.slots >= 1 && a.work && a.bslots &&
              (a.slots == 1 || (a.partial && a.slots * nr * a.ncols <= a.partial_cap)),
          kErrInternal, "missing_spmm_tiled: bad work split");
  const int lpe = a.cp / 4, groups = 16 * (64 / lpe);
  const int rpg = (kMissRB + groups - 1) / groups;
  const dim3 grid(a.nwork);
  const size_t lds = 2 * (size_t)kMissLdsBuf;
  static bool once = (allow_lds(k_missing_spmm_tiled<1>, lds), allow_lds(k_missing_spmm_tiled<2>, lds),
                      allow_lds(k_missing_spmm_tiled<3>, lds), true);
  (void)once;
  switch (rpg) {
    case 1: hipLaunchKernelGGL(k_missing_spmm_tiled<1>, grid, dim3(kMissThreads), lds, s, a, lpe); break;
    case 2: hipLaunchKernelGGL(k_missing_spmm_tiled<2>, grid, dim3(kMissThreads), lds, s, a, lpe); break;
    default: hipLaunchKernelGGL(k_missing_spmm_tiled<3>, grid, dim3(kMissThreads), lds, s, a, lpe); break;
  }
  check_launch("missing_spmm_tiled");
  if (a.slots > 1) {
    const uint64_t n = nr * (uint64_t)a.ncols;
    const unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 65535u * 8);
    hipLaunchKernelGGL(k_sum_slots, dim3(g), dim3(256), 0, s, a.partial, nr, a.ncols, a.r_lo, a.rpos, a.bslots,
                       a.out, a.lp);
    check_launch("sum_slots");
  }
}

}  // namespace pcaone
