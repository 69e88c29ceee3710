// Launchers of the non-GEMM kernels (kernels_data.hip, kernels_la.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm.hpp"
#include "synth.hpp"

namespace pcaone {

// ---- data layout (kernels_data.hip) --------------------------------------
// raw PLINK rows (bps bytes each) -> interleaved rows of `pitch` words; codes of
// samples >= nsamples become 01.
void launch_interleave(const uint8_t* raw, uint64_t nrows, uint64_t bps, uint32_t nsamples, uint32_t* out,
                       uint64_t pitch, hipStream_t s);
// Fills rows [r0, r1) of an interleaved matrix with code 01.
void launch_fill_rows(uint32_t* x, uint64_t pitch, uint64_t r0, uint64_t r1, hipStream_t s);
// Per-row counts of codes 00/01/10/11 over the first `ncodes` codes: cnt[4*r + c].
void launch_code_counts(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint64_t ncodes, uint32_t* cnt,
                        hipStream_t s);
// SNP-major (rows = SNPs, codes = samples) -> sample-major (rows = samples,
// codes = SNPs).  Rows nsnps.. of `xt` get code 01.
void launch_transpose_codes(const uint32_t* xs, uint64_t pitch_s, uint64_t nsnps, uint64_t nsamples, uint32_t* xt,
                            uint64_t pitch_t, uint64_t rows_t, hipStream_t s);
// out row i = in row idx[i] (interleaved rows, whole words).
void launch_gather_rows(const uint32_t* in, uint64_t pitch, const uint32_t* idx, uint64_t nrows, uint32_t* out,
                        hipStream_t s);
// Synthetic genotype rows (rows[r] or r for r < nrows) from the shared
// generator; thr is nsnps x npops, attempts per SNP, flags per sample.
void launch_synth(uint64_t seed, uint32_t nsamples, uint32_t npops, const uint64_t* thr, const uint32_t* attempts,
                  const uint8_t* flags, const synth::MissSpec& ms, const uint32_t* rows, uint64_t nrows, uint32_t* out,
                  uint64_t pitch, hipStream_t s);
// Interleaved rows -> PLINK rows of bps bytes.
void launch_deinterleave(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint32_t nsamples, uint8_t* raw,
                         uint64_t bps, hipStream_t s);
// Plain device-to-device copy of `bytes` (multiple of 16), for counter calibration.
void launch_copy_probe(const void* in, void* out, uint64_t bytes, hipStream_t s);
// Missing-entry lists: for row r, the positions k < ncodes whose code is 01, in
// increasing order, written at ix[rp[r] ..].
void launch_missing_lists(const uint32_t* x, uint64_t pitch, uint64_t nrows, uint64_t ncodes, const uint64_t* rp,
                          uint32_t* ix, hipStream_t s);

// ---- linear algebra on tall FP64 row-major matrices (kernels_la.hip) -----
// Splits X (rows x lp, row-major FP64) into MFMA operand pieces, transposed:
// Wt[p][l][r] for r < rows_pad (rows >= nreal written as 0).  Row r is first
// multiplied by rowscale[r] (null = 1) and column l by colsign[l] (null = 1).
// rounded (nullable, row-major rows x lp) receives sum(pieces) / rowscale
// (in place over X allowed); signed_out (nullable) X with the column signs.
// The I8 and F8 forms take their column scales from colscale (launch_col_scales).
// With bvec: bvec[l] = sum_r w_r rounded[r][l], w_r = bsc[r] (1/2 - bcf[r])
// (1 without bsc), from per-tile partials in bpart (bpart_cap elements; a
// separate column sum over rounded when they do not fit).
void launch_make_pieces(Form f, const double* X, uint64_t nreal, uint64_t rows_pad, int lp, const double* rowscale,
                        const double* colsign, void* Wt, uint64_t wpitch, uint64_t wpiece, double* rounded,
                        double* signed_out, hipStream_t s, const double* colscale = nullptr,
                        const double* bsc = nullptr, const double* bcf = nullptr, double* bpart = nullptr,
                        uint64_t bpart_cap = 0, double* bvec = nullptr);
// Column scales of X for a column-scaled form (rows [0, nreal), each row
// times rowscale[r], null = 1), from m_l = max_r |X[r][l] rowscale[r]|:
// I8: scale[l] = m_l / kI8Max, mul[l] = 2 scale[l] (the genotype operand is
// v/2); F8: scale[l] = the power of two that brings m_l into (224, 448],
// mul[l] = scale[l].  An all-zero column has scale 1.  maxbits: lp words of
// scratch.
void launch_col_scales(Form f, const double* X, uint64_t nreal, int lp, const double* rowscale,
                       unsigned long long* maxbits, double* scale, double* mul, hipStream_t s);
void launch_epi_g(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* sc, const double* cf, const double* colsum, double* out, hipStream_t s,
                  const double* colmul = nullptr);
// G epilogue fused with the operand split, for data without missing entries
// (not for I8, whose column scales need the whole window first):
// gout (nwin x lp) = G~ window, Wt pieces of diag(sc) G~ (window offset
// already applied), bvec = sum_j sc_j (1/2 - cF_j) G~_j via bpart (scratch of
// epi_g_fused_partial_elems).
uint64_t epi_g_fused_partial_elems(uint64_t nwin, int lp);
void launch_epi_g_fused(Form f, const void* P, uint64_t nchunks, uint64_t stride, uint64_t nwin, int lp,
                        const double* sc, const double* cf, const double* colsum, double* gout, void* Wt,
                        uint64_t wpitch, uint64_t wpiece, double* bpart, double* bvec, hipStream_t s);
// H epilogue: t = sum_c P[c][i][l] / 4 + b[l]; mode 0: out = t, 1: out += t (i < nreal; rows above = 0 in mode 0).
void launch_epi_h(Form f, const void* P, uint64_t nchunks, uint64_t chunk_stride, uint64_t nreal, uint64_t rows_pad,
                  int lp, const double* b, double* out, int mode, hipStream_t s, const double* colmul = nullptr);
// Column sums out[l] = sum_{r in [r0,r1)} w_r * X[r][l] (w null = 1), two-stage, deterministic.
void launch_colsum(const double* X, uint64_t r0, uint64_t r1, int lp, const double* w_sc, const double* w_cf,
                   double* partial, double* out, hipStream_t s);
// Gram-type reduction: out = (X T)^T (Y) over rows [0, nrows), where Y = Y0 if
// given else X T; T null = identity.  out is lp x lp row-major.
void launch_gram(const double* X, const double* T, const double* Y0, uint64_t nrows, int lp, double* partial,
                 uint64_t partial_cap, double* out, hipStream_t s);
uint64_t gram_partial_elems(uint64_t nrows, int lp);
// Y = X T (rows x lp) for rows [0, nrows); T lp x lp row-major; rows above nrows untouched.
void launch_apply(const double* X, const double* T, uint64_t nrows, int lp, double* Y, hipStream_t s);
// flipOmg norms: a[l] = sum_i |O2 - Q|, b[l] = sum_i |O2 + Q|.
void launch_flip_norms(const double* O2, const double* Q, uint64_t nrows, int lp, double* partial, double* out,
                       hipStream_t s);
// Thin-Q of an N x L matrix X (CholeskyQR2 with Householder column signs) with
// its L x L steps on the device, for lp <= 80 (thinq_on_device):
//   first:  R1 = chol(G1 = X^T X), T1 = R1^-1;
//   second: R2 = chol(G2 = (X T1)^T (X T1)), T = (R2 R1)^-1 diag(s), s the
//           Householder signs of Xtop (R2 R1)^-1, Xtop the first n rows of X.
// A Cholesky breakdown sets *status = 1 (the outputs are then zero).
bool thinq_on_device(int lp);
// min_ratio > 0: also set *status = 2 when the smallest Cholesky pivot is at
// most min_ratio times the largest diagonal entry (H near rank deficient).
void launch_thinq_first(const double* G1, int n, int lp, double* R1, double* T1, int* status, hipStream_t s,
                        double min_ratio = 0.0);
void launch_thinq_second(const double* G2, const double* R1, const double* Xtop, int n, int lp, double* T,
                         int* status, hipStream_t s);
// flipOmg's column signs from the norms of launch_flip_norms: -1 where
// ab[l] > 2 ab[lp + l] (l < n), else 1 (lp entries).
void launch_flip_signs(const double* ab, int n, int lp, double* sign, hipStream_t s);
// Zero bytes [b0, b0 + n0) and [b1, b1 + n1) of each of `rows` rows of `pitch` bytes.
void launch_zero_ranges(void* base, uint64_t pitch, uint64_t rows, uint64_t b0, uint64_t n0, uint64_t b1, uint64_t n1,
                        hipStream_t s);
// flip_UV with V as the base (src/Utils.cpp:118-154) on device-resident
// V (m x lp) and U (n x lp), row-major: column c < k of both changes sign when
// V's entry of largest magnitude (the first such) is negative.  partial:
// 2048 x lp scratch; sign: lp.
void launch_flip_cols_by_max(double* V, uint64_t m, double* U, uint64_t n, int lp, int k, double* partial,
                             double* sign, hipStream_t s);
// dst (rows x kp) = first k columns of src (rows x lp) times scale[c] (null = 1), zero padded.
void launch_impute_rows(const double* src, uint64_t rows, int lp, int k, int kp, const double* scale, double* dst,
                        hipStream_t s);
// Per-SNP column state: cf = F, sc = 1 for j < m, 0 up to rows.
void launch_pristine_cols(const double* F, uint64_t m, uint64_t rows, double* cf, double* sc, hipStream_t s);
// sc_j *= sqrt_ploidy / sqrt(F_j (1 - F_j)) where that sd exceeds tol (j < m).
// ---- exact PCA / IRAM operands (kernels_exact.hip) -----------------------
// gt (m x lp, row-major) <- the standardized columns of samples [s0, s0 + nc)
void launch_sample_columns(const uint32_t* xs, uint64_t pitch, uint64_t m, uint64_t s0, int nc, int lp,
                           const double* sc, const double* cf, double* gt, hipStream_t s);
// om (n x L, column-major) <- the standardized rows of SNPs [t0, t0 + nc)
void launch_snp_rows(const uint32_t* xs, uint64_t pitch, uint64_t n, uint64_t t0, int nc, int L, const double* sc,
                     const double* cf, double* om, hipStream_t s);
// K[:, col0:col0 + nc] (column-major, leading dimension ldk) <- scale X[:, :nc] (rows x lp row-major)
void launch_rows_to_cols(const double* X, uint64_t rows, int lp, int nc, double scale, double* K, uint64_t ldk,
                         uint64_t col0, hipStream_t s);

// ---- residual matrix (kernels_resid.hip) --------------------------------
constexpr int kResidMaxK = 64;  // rank of the subtracted U S V^T
struct ResidualArgs {
  const uint32_t* xs;  // SNP-major interleaved codes
  uint64_t pitch;      // words per row
  uint64_t n;          // samples
  uint64_t j0, j1;     // SNP rows (current column order)
  const double* sc;    // per-SNP scale, null = 1
  const double* cf;    // per-SNP centering frequency
  int k;               // rank of the subtracted part (0: none)
  const double* ut;    // U, k x n
  const double* sv;    // S_t V_jt, M x k
  int kimp, kpimp;     // EMU imputation rank and row pitch (us null: no imputation)
  const double* us;    // U S of the imputation, n x kpimp
  const double* vimp;  // V of the imputation, M x kpimp
  const double* fclamp;
  double* part;        // nchunks x (j1 - j0) sums
  uint64_t nchunks, part_stride;
  float* out;          // (j1 - j0) x n
};
// Capacity of `part` needed for a launch over `snps` rows and n samples.
inline uint64_t residual_partial_elems(uint64_t n, uint64_t snps) { return (n + 255) / 256 * snps; }
void launch_residuals(const ResidualArgs& a, hipStream_t s);

void launch_standardize(const double* F, uint64_t m, double sqrt_ploidy, double tol, double* sc, hipStream_t s);
// dst[i] = src[idx[i]], i < n.
void launch_gather_f64(const double* src, const uint32_t* idx, uint64_t n, double* dst, hipStream_t s);
// Column-major n x L (leading dimension n) -> row-major rows_pad x lp (zero padded).
void launch_colmajor_to_rows(const double* src, uint64_t n, int L, int lp, uint64_t rows_pad, double* dst,
                             hipStream_t s);
// Row-major n x lp -> column-major n x k.
void launch_rows_to_colmajor(const double* src, uint64_t n, int lp, int k, double* dst, hipStream_t s);
// Y = X + Z elementwise (n elements).
void launch_add(const double* X, const double* Z, double* Y, uint64_t n, hipStream_t s);
// Missing-entry corrections, in two steps.
// (1) Weights, one per list entry e = (row r, index k), recomputed whenever the
//     column state or the imputation changes:
//     w[e] = -sc_c (1/2 - cF_c) + (emu ? sc_c * clamp(sum_t US[i][t] V[c][t], -F_c, 1-F_c) : 0)
//     where c is the SNP (r for by_snp lists, k otherwise) and i the sample.
//     The index range [0, kext) is processed in chunks of kspan.
struct MissWeightArgs {
  const uint64_t* rp;
  const uint32_t* ix;
  uint64_t nrows;
  bool by_snp;
  const double* sc;      // per SNP column
  const double* cf;      // per SNP column
  const double* Fclamp;  // per SNP column (EMU clamp bound), null = no EMU
  const double* US;      // N x kk row-major (U * S), row pitch kp, EMU only
  const double* V;       // M x kk row-major, row pitch kp, EMU only
  int kk;
  int kp;                // row pitch of US and V: whole 128-byte lines; V row j holds sc_j, cF_j, F_j at kk..kk+2
  double* w;             // one weight per list entry
  uint64_t kext, kspan;
};
void launch_missing_weights(const MissWeightArgs& a, hipStream_t s);
// V[j][kk + 0..2] = sc_j, cF_j, F_j (rows of pitch kp, j < m).
void launch_pack_col_state(double* V, uint64_t m, int kp, int kk, const double* sc, const double* cf, const double* F,
                           hipStream_t s);
// Compact copy of the first L columns of rows x lp FP64 rows into rows x cp
// (FP32 if f32, else FP64), padding columns zero.
void launch_compact_rows(const double* src, uint64_t rows, int lp, int L, int pitch, bool f32, void* out,
                         hipStream_t s);
// (2) out[r][:ncols] += sum_e w[e] * src[ix[e] - src_k0][:ncols] over the
//     entries of row r with index in [k_lo, k_hi).  src is a compact copy
//     (pitch cp, a multiple of 4 for FP32 / 2 for FP64).  The index range is
//     split into chunks of kspan (0 = one chunk); with several chunks the
//     per-chunk sums go to `partial` (chunks x nrows x ncols FP64) and are
//     added to out in chunk order.
struct MissSpmmArgs {
  const uint64_t* rp;  // row pointers (rows may start at an offset into the lists)
  const uint32_t* ix;
  const double* w;
  uint64_t nrows;
  uint64_t k_lo, k_hi;
  const void* src;
  bool f32;
  uint64_t src_k0;     // index of the first source row
  int cp;              // columns read per source row (a multiple of 16 bytes)
  int cpitch;          // elements between source rows (whole 128-byte lines)
  int ncols;
  uint64_t kspan;      // index chunk (rows of src that fit in L2)
  double* partial;     // [chunk][nrows][ncols] partial rows when the index range is split
  uint64_t partial_cap;  // elements of partial
  double* out;         // nrows x lp
  int lp;
};
// Adds the corrections to out; the source rows are gathered from L2 in index
// chunks of kspan rows (partial rows summed in chunk order).
void launch_missing_spmm(const MissSpmmArgs& a, hipStream_t s);
// out[r][l] += sum_z partial[z][r][l] in z order (nz x nrows x ncols partials; out pitch lp).
void launch_sum_chunks(const double* partial, uint64_t nz, uint64_t nrows, int ncols, double* out, int lp,
                       hipStream_t s);

// ---- tiled missing-entry corrections (kernels_miss.hip) -------------------
// A missing-entry list (CSR rows, sorted indexes) re-laid in tiles: row blocks
// of kMissRB list rows, each split along the index range into tiles of whole
// kMissKG-index granules that hold at most kMissEMax entries and span at most
// `span` indexes.  A tile's entries are stored together (row by row, indexes
// ascending within a row) as 16-bit indexes local to the tile and FP32 weights,
// with kMissRowStride 16-bit row offsets per tile.  One workgroup stages a
// whole tile - row offsets, entries and the source rows of its index range -
// into LDS and every lane group owns fixed rows of the block, so a source row
// is read from HBM or L2 once per tile rather than once per entry.
constexpr int kMissRB = 192;
constexpr int kMissKG = 32;
constexpr int kMissEMax = kMissRB * kMissKG;  // one granule always fits
constexpr int kMissRowStride = 200;           // kMissRB + 1 offsets, padded to 16 bytes
constexpr int kMissThreads = 1024;
constexpr int kMissMaxCp = 32;                // FP32 source rows of at most 32 columns
constexpr uint32_t kMissLdsBuf = 80 * 1024;   // bytes of one staging buffer (two per workgroup)

// Entries stored per tile: its ne entries, zeros up to a multiple of 8 and 8
// more zeros (index 0, weight 0), so that a step of 8 entries that starts
// inside the tile reads only initialized entries.
__host__ __device__ constexpr uint32_t miss_stored(uint32_t ne) { return (ne + 7) / 8 * 8 + 8; }

struct MissTile {
  uint64_t e0;      // first entry (a multiple of 8)
  uint32_t k0, k1;  // index range [k0, k1)
  uint32_t ne;      // entries
  uint32_t pad;
};

// Largest index span of a tile whose FP32 source rows of cp columns fit a
// staging buffer beside the row offsets and kMissEMax entries.
uint32_t miss_tile_span(int cp);
// The rows of a list may be placed in the row blocks in another order:
// position p = b kMissRB + i holds list row rperm[p] (>= nrows: none), and
// rpos[r] is row r's position (both null: row r at position r).
// Entry counts per (row block, granule): cnt[b * ngran + g] += entries of the
// block's rows with index in granule g (cnt zeroed by the caller).
void launch_miss_hist(const uint64_t* rp, const uint32_t* ix, uint64_t nrows, const uint32_t* rpos, uint32_t ngran,
                      uint32_t* cnt, hipStream_t s);
// Fills the tiles' row offsets and local indexes from the CSR list (tiles and
// tb = first tile of each row block, planned on the host from the counts);
// *status = 1 if a tile's rows hold other than its planned entry count.
void launch_miss_fill(const uint64_t* rp, const uint32_t* ix, uint64_t nrows, const uint32_t* rperm,
                      const MissTile* tiles, const uint32_t* tb, uint64_t nblocks, uint16_t* trow, uint16_t* eix,
                      int* status, hipStream_t s);

// A workgroup's share of a tiled correction: tiles [t0, t1) of row block b,
// the slot-th of the block's nslot shares.
struct MissWork {
  uint32_t b, t0, t1, slot;
};
struct MissTileList {
  const MissTile* tiles;
  const uint32_t* tb;     // nblocks + 1
  const uint16_t* trow;
  const uint16_t* eix;
  const uint32_t* rperm;  // list row of each position (null: the position itself)
};
// Weights of a tiled list, as MissWeightArgs describes them, in tile order
// (FP32).  EMU: the row side's k-vectors of a block and the index side's of a
// tile (`sub` rows at a time) sit in LDS.
struct MissTiledWeightArgs {
  MissTileList t;
  float* ew;
  uint64_t nrows, nblocks;
  bool by_snp;
  const double* sc;
  const double* cf;
  const double* Fclamp;  // null = no EMU
  const double* US;
  const double* V;
  int kk, kp;
  const MissWork* work;  // one workgroup per item
  uint32_t nwork;
};
void launch_missing_weights_tiled(const MissTiledWeightArgs& a, hipStream_t s);
// out[r - r_lo][:ncols] (+)= sum over the entries of list rows r in [r_lo,
// r_hi) with index k in [k_lo, k_hi) of w * src[k - src_k0][:ncols]; src are
// FP32 compact rows (cp columns, pitch cpitch).  One workgroup per work item;
// with several shares per block (slots > 1) each writes partial rows
// ([slot][r - r_lo][ncols]) that are summed in slot order, the block's nslot
// (bslots[b - r_lo / kMissRB]) of them.
struct MissTiledSpmmArgs {
  MissTileList t;
  const float* ew;
  uint64_t r_lo, r_hi, k_lo, k_hi;
  const float* src;
  uint64_t src_k0;
  int cp, cpitch, ncols;
  const MissWork* work;
  uint32_t nwork;
  uint32_t slots;           // largest share count of a block
  const uint32_t* bslots;   // shares of each block in the range
  const uint32_t* rpos;     // position of each list row (null: the row itself)
  double* partial;
  uint64_t partial_cap;
  double* out;
  int lp;
};
void launch_missing_spmm_tiled(const MissTiledSpmmArgs& a, hipStream_t s);

}  // namespace pcaone
