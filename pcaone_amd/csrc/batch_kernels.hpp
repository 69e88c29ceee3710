// Device structures and launchers of the batched (many-cohort) randomized SVD.
//
// Every cohort is an independent genotype matrix with its own N (samples) and
// M (SNPs); all cohorts share L = k + oversamples (L <= 32) and the operand
// form.  Grouped kernels walk item lists that cover the rows of every cohort,
// so one launch advances all cohorts by one step; the per-cohort L x L work
// runs in one workgroup per cohort on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm.hpp"

namespace pcaone {

constexpr int kBatchMaxL = 32;      // largest L of the batched form
constexpr int kBatchLp = 32;        // row pitch (columns) of every FP64 matrix
constexpr int kBatchMaxEpochs = 64; // entries of the per-cohort trace

// Per-cohort state and the L x L matrices of the small steps (lp x lp, row-major).
struct BatchSmall {
  int state;       // 1 running, 2 stopping at this epoch, 0 done, < 0 failed (PCAONE_ERR_* negated)
  int epochs;      // epochs run
  int stop_epoch;  // epoch whose outputs are the result
  int nsv;         // singular values produced (0 when stopped at epoch 0 of --maxp 0)
  double gram_a[kBatchLp * kBatchLp];  // Gram of G~, of Q1, or the cross Gram Ucur^T Upre
  double gram_h[kBatchLp * kBatchLp];  // H^T H
  double r1h[kBatchLp * kBatchLp];     // R1 of H (CholeskyQR2 of H)
  double t[kBatchLp * kBatchLp];       // transform of the next apply
  double tv[kBatchLp * kBatchLp];      // R^-1 W[:, :k] of the stopping epoch (loadings)
  double sigma[kBatchLp];
  double colsign[kBatchLp];
  double ab[2 * kBatchLp];             // flipOmg column norms
  double trace[kBatchMaxEpochs];       // 1 - mev per epoch
  // FP8 form: column maxima (bit patterns of non-negative doubles) of the
  // Omega that is split (G side) and of diag(sc) G~ (H side); each column's
  // power-of-two scale follows from its maximum
  unsigned long long cmax_g[kBatchLp];
  unsigned long long cmax_h[kBatchLp];
};

// Which per-cohort FP64 matrix a grouped kernel reads or writes.
enum BatchBuf : int {
  kBufOmega = 0,  // Omg2: the current Omega, exact (N rows)
  kBufOmegaR,     // Omega as multiplied (sum of operand pieces) (N rows)
  kBufOmegaC,     // Q of H, the next Omega before flipOmg (N rows)
  kBufG,          // G~ (M rows)
  kBufH,          // H (N rows)
  kBufTmpN,       // Q1 of H (N rows)
  kBufU0,         // Ucur of even epochs (N rows)
  kBufU1,         // Ucur of odd epochs (N rows)
  kBufV,          // loadings (M rows)
  kBufOmega0,     // the initial Omega from the host draw (N rows), kept across runs
  kNumBatchBufs
};

struct BatchCohort {
  uint64_t N, M;
  double* mat[kNumBatchBufs];  // indexed by BatchBuf, lp = kBatchLp columns
  void* wg;              // Omega operand pieces [piece][col][sample]
  uint64_t wg_pitch;     // >= round_up(N, 128)
  void* wh;              // operand pieces of diag(sc) G~ [piece][col][snp]
  uint64_t wh_pitch;     // >= round_up(M, 128)
  const double* sc;      // per SNP scale
  const double* cf;      // per SNP centering frequency
  double* colsum;        // column sums of Omega~ (lp)
  double* bvec;          // sum_j sc_j (1/2 - cF_j) G~_j (lp)
  void* part;            // split-K partial products (FP32, FP64 for the FP64 form)
  uint64_t nch_g, ychunk_g, nch_h, ychunk_h;
  BatchSmall* st;
  // item ranges of this cohort in the row-item lists (reductions sum them in order)
  uint32_t m_begin, m_end;   // items over the M rows (reductions)
  uint32_t n_begin, n_end;   // items over the N rows
  uint32_t e_begin, e_end;   // epilogue items over the M rows (kBatchEpiRows rows each)
};

struct RowItem {
  uint32_t cohort;
  uint32_t nrows;
  uint64_t row0;
};

constexpr int kBatchEpiRows = 128;    // rows per G-epilogue item
constexpr int kBatchRowsPerItem = 1024;  // rows per reduction / elementwise item

struct BatchArgs {
  const BatchCohort* coh;
  uint32_t ncoh;
  int L, k, p;      // columns, components, last epoch index (maxp)
  double tol;
  int form;         // Form of the products
};

// Small steps (one workgroup per cohort).
enum SmallMode : int {
  kSmallEpoch = 0,   // R from Gram(G~), B B^T, eigen-decomposition, Ucur transform
  kSmallStop,        // mev diff and the stop rule (cross Gram in gram_a)
  kSmallOmega1,      // R1 = chol(H^T H), t = R1^-1
  kSmallOmega2,      // R2 = chol(Gram(Q1)), Householder signs, t = (R2 R1)^-1 diag(s)
  kSmallFlip,        // flipOmg decision from ab
  kSmallFinish,      // stopping cohorts become done
  kSmallInit,        // every cohort running, no flips, counters cleared
};

void launch_batch_small(const BatchArgs& a, int mode, int epoch, hipStream_t s);
// Omega operand pieces from src (kBufOmega0 or kBufOmegaC), with colsign when
// `signed_`: writes Wt, Omega~ (kBufOmegaR) and Omg2 (kBufOmega); running
// cohorts only.
void launch_batch_make_pieces(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, int src, bool signed_,
                              hipStream_t s);
// Column sums of Omega~ into colsum (partials per item, then per cohort).
void launch_batch_colsum(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, double* partial,
                         hipStream_t s);
// FP8 form: the column maxima of the Omega in buffer src (N rows) that the
// next make_pieces splits.
void launch_batch_omega_colmax(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, int src, hipStream_t s);
// G epilogue (fused with the operand split of diag(sc) G~ and the b-vector).
// FP8 form: G~ and the column maxima of diag(sc) G~ first, then the split
// with the column scales.
void launch_batch_epi_g(const BatchArgs& a, const RowItem* items_e, uint32_t nitems, double* partial,
                        hipStream_t s);
// H epilogue: H = sum_c partials / 4 + b.
void launch_batch_epi_h(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, hipStream_t s);
// Gram X^T Y (Y = X when ysel < 0) of every cohort in state `state_req` into
// gram_a (dest 0) or gram_h (dest 1).
void launch_batch_gram(const BatchArgs& a, const RowItem* items, uint32_t nitems, bool m_rows, int xsel, int ysel,
                       int dest, int state_req, double* partial, hipStream_t s);
// Y = X t (tsel 0) or X tv (tsel 1) for cohorts in state `state_req`.
void launch_batch_apply(const BatchArgs& a, const RowItem* items, uint32_t nitems, int xsel, int ysel, int tsel,
                        int state_req, hipStream_t s);
// flipOmg norms of Omg2 against Q (kBufOmegaC) into ab.
void launch_batch_flip(const BatchArgs& a, const RowItem* items_n, uint32_t nitems, double* partial,
                       hipStream_t s);
// Cohort table with `active[c] = (state == 1)` for the grouped GEMM.
void launch_batch_active(const BatchArgs& a, int* active, hipStream_t s);

}  // namespace pcaone
