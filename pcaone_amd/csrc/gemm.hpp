// Packed-genotype MFMA GEMM: Y[chunk][r][:] = sum_{k in chunk} v(code[r][k]) * W[k][:]
//
// X holds 2-bit PLINK codes in the "interleaved" word layout (code e of a
// 32-bit word sits at bit 8*(e%4) + 2*(e/4), so (w >> 2t) & 0x03030303 puts
// codes 4t..4t+3 into the four bytes).  Rows are padded to a multiple of
// kKStep codes with code 01, which decodes to 0.
//
// v(code) = 2*(copies of A1 - 1): 00 -> +2, 10 -> 0, 11 -> -2, 01 (missing)
// -> 0.  These values and their products with BF16 pieces are exact, and the
// affine offset (and the missing entries) are corrected by the callers'
// epilogues, so the decode is two byte permutes per four codes.
//
// W is stored transposed per piece: Wt[piece][col][k], col < lp, pitch wpitch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pcaone {

constexpr int kKStep = 256;       // codes consumed per K-step (and row padding unit)

// Operand form of one product.  BF16 = hi/lo pieces, BF16X1 = one piece,
// FP16 = one FP16 piece (11 significand bits).  I8 = 16-bit fixed point per
// column, q = rint(w / s_col) with |q| <= kI8Max, split into two int8 pieces
// q = 256 q_hi + q_lo on the int8 matrix cores; the genotype operand is
// v/2 in {+1, 0, -1} and the products accumulate exactly in int32.  F8 =
// three FP8 E4M3 pieces of w / s_col (s_col a power of two that brings the
// column maximum into (224, 448]): w = s_col (p1 + p2 + p3) to about 2^-12
// of |w|, on the FP8 matrix cores (v_mfma_f32_16x16x128_f8f6f4); the
// genotype operand v in {+2, 0, -2} is exact in E4M3 and the pieces
// accumulate together in FP32.
enum class Form : int { FP64 = 0, FP32 = 1, BF16 = 2, BF16X1 = 3, FP16 = 4, I8 = 5, F8 = 6 };

// Largest |q| of the I8 form: q_hi = (q + 128) >> 8 stays within int8.
constexpr int kI8Max = 32639;
// K codes per I8 chunk: |256 sum a q_hi + sum a q_lo| <= K (127 * 256 + 128)
// must stay below 2^31.
constexpr uint64_t kI8ChunkCap = 65536;

// User-facing precision (PCAONE_PREC_*) -> forms of the two products
// (G = A Omega, H = A^T G).  4 = BF16 mixed: one BF16 piece of Omega, hi/lo of
// G; 5 = FP16, one piece of each; 6 = FP16 mixed: one FP16 piece of Omega,
// BF16 hi/lo of G; 7 = I8 for both; 8 = F8 for both.
struct ProductForms {
  Form g, h;
};
inline ProductForms forms_of(int precision) {
  switch (precision) {
    case 0: return {Form::FP64, Form::FP64};
    case 1: return {Form::FP32, Form::FP32};
    case 2: return {Form::BF16, Form::BF16};
    case 3: return {Form::BF16X1, Form::BF16X1};
    case 4: return {Form::BF16X1, Form::BF16};
    case 5: return {Form::FP16, Form::FP16};
    case 6: return {Form::FP16, Form::BF16};
    case 7: return {Form::I8, Form::I8};
    default: return {Form::F8, Form::F8};
  }
}
constexpr int kNumPrecisions = 9;

inline int form_pieces(Form f) { return f == Form::F8 ? 3 : (f == Form::BF16 || f == Form::I8 ? 2 : 1); }
inline size_t form_wbytes(Form f) {
  return f == Form::FP64 ? 8 : (f == Form::FP32 ? 4 : (f == Form::I8 || f == Form::F8 ? 1 : 2));
}
// Forms whose operand pieces carry a per-column scale (undone by the epilogues).
inline bool form_colscaled(Form f) { return f == Form::I8 || f == Form::F8; }
// Partial products: FP64, int32 (I8) or FP32.
inline size_t form_accbytes(Form f) { return f == Form::FP64 ? 8 : 4; }
// Largest split-K chunk a form accepts.
inline uint64_t form_kcap(Form f, uint64_t cap) { return f == Form::I8 && cap > kI8ChunkCap ? kI8ChunkCap : cap; }

struct GemmArgs {
  const uint32_t* X;   // interleaved packed rows
  uint64_t xpitch;     // 32-bit words per row
  uint64_t rows;       // rows to compute (multiple of the row tile)
  const void* Wt;      // pieces of W, transposed
  uint64_t wpitch;     // elements between consecutive cols of one piece
  uint64_t wpiece;     // elements between pieces
  void* Y;             // partials [chunk][rows][lp] (float; double for FP64, int32 for I8)
  uint64_t ychunk;     // elements between chunks of Y
  uint64_t k_begin;    // first code (multiple of kKStep)
  uint64_t k_end;      // one past the last code (multiple of kKStep)
  uint64_t kchunk;     // codes per chunk (multiple of kKStep)
  int lp;              // padded number of columns (multiple of 16, <= 128)
};

// Grouped (batched) form: many independent products in one persistent launch.
struct GemmCohort {
  const uint32_t* X;   // packed rows of this product
  uint64_t xpitch;
  const void* Wt;      // operand pieces, transposed
  uint64_t wpitch, wpiece;
  void* Y;             // partials [chunk][rows][lp]
  uint64_t ychunk;
};
struct GemmItem {
  uint32_t cohort;
  uint32_t chunk;      // index of the partial block written
  uint64_t row0;       // first row of the tile
  uint64_t k0, k1;     // code range
};
struct GroupedGemmArgs {
  const GemmItem* items;
  uint32_t nitems;
  const GemmCohort* cohorts;
  const int* active;   // per cohort, null = all active
};

// Workgroup shape: `waves` waves (4 or 8), each owning `wm` row blocks of 16.
struct GemmShape {
  int wm = 0;
  int waves = 0;
};
inline int gemm_row_tile(const GemmShape& g) { return g.waves * g.wm * 16; }
GemmShape gemm_default_shape(Form f, int lp);
bool gemm_shape_supported(Form f, const GemmShape& g, int lp);
void launch_packed_gemm(Form f, const GemmShape& g, const GemmArgs& a, hipStream_t s);
// `grid` workgroups walk the item list (rows of every item are a whole tile of g).
void launch_grouped_gemm(Form f, const GemmShape& g, int lp, const GroupedGemmArgs& a, unsigned grid,
                         hipStream_t s);
// Workgroups of the grouped kernel of this shape resident on one CU.
int grouped_gemm_blocks_per_cu(Form f, const GemmShape& g, int lp);
// The same for the single-product kernel (cached per form, shape and lp).
int gemm_blocks_per_cu(Form f, const GemmShape& g, int lp);

// Launch plan of one product: shape, padded rows and split-K chunk.
struct GemmPlan {
  GemmShape shape;
  uint64_t rows = 0;    // rows rounded up to the shape's tile
  uint64_t kchunk = 0;  // codes per chunk
  uint64_t nchunks = 0;
  uint64_t workgroups() const { return rows / (uint64_t)gemm_row_tile(shape) * nchunks; }
};
// Picks the split-K chunk (and, when the preferred shape leaves the chip
// under-filled and `fixed_shape` is false, a smaller row tile) for a product
// of `nrows` rows over `kext` codes: at least 2,048 workgroups where the
// matrix allows it, at least 4 K-steps per workgroup, chunks of at most
// `kcap` codes.  kchunk_user > 0 fixes the chunk.  With the device's CU count
// `cus` (0 = not known) a product of at most 4 Mi rows takes up to two more
// chunks when they fill its last round of resident workgroups.
GemmPlan plan_gemm(Form f, const GemmShape& preferred, bool fixed_shape, int lp, uint64_t nrows, uint64_t kext,
                   uint64_t kchunk_user, uint64_t kcap, uint64_t cus = 0);
// Upper bound of nchunks * rows over every plan_gemm result for these
// arguments' extremes (partial-product buffer sizing).
uint64_t plan_partial_rows(uint64_t nrows_max, uint64_t kext_max, uint64_t kchunk_user, uint64_t kcap);

}  // namespace pcaone
