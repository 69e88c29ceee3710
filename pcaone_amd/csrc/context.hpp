// Device context: the resident genotype matrix in two packed layouts, the
// per-SNP centering/scaling state, and the randomized-SVD workspace.
//
// HBM layout (N samples, M SNPs, codes = 2 bits):
//   xs  SNP-major interleaved words, rows_s x pitch_s   (product G = A * Omega)
//   xt  sample-major interleaved words, rows_t x pitch_t (product H = A^T * G)
//   cF, sc, F   per-column centering frequency, scale, and allele frequency
//               in the reference's (unpermuted) SNP order
//   miss_s/t    CSR lists of missing entries by SNP and by sample, and for the
//               reduced forms the same lists in LDS-sized tiles
//   omega       Omega (FP64, N x lp, row-major) == Omg2 of the reference
//   omega_r     the Omega actually multiplied (sum of operand pieces)
//   wg          Omega operand pieces, transposed [piece][col][sample]
//   gt          G~ (FP64, M x lp) = the G actually multiplied by A^T
//   wh          window operand pieces of diag(sc) G~, transposed [piece][col][snp]
//   H, H1, H2, Hsum   FP64 N x lp
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
#include "linalg_host.hpp"

namespace pcaone {

namespace synth {
struct Spec;
}
class PgenSource;

// Synthetic genotypes of `spec` straight into interleaved rows (row j = SNP j,
// `pitch` words per row) on `stream`, byte-identical to the host generator.
void synth_rows(const synth::Spec& spec, uint32_t* xs, uint64_t pitch, hipStream_t stream);

struct GemmTuning {
  int wm_g = 0;          // wave tile of the G product (0 = default)
  int wm_h = 0;          // wave tile of the H product
  int waves_g = 0;       // waves per workgroup of the G product (4 or 8; 0 = default)
  int waves_h = 0;       // waves per workgroup of the H product
  uint64_t kchunk_g = 0; // split-K chunk of the G product in codes (0 = auto)
  uint64_t kchunk_h = 0; // split-K chunk of the H product in codes (0 = auto)
  bool miss_gather = false;  // missing-entry corrections from the CSR lists even where the tiles apply
};

struct EpochResult {
  std::vector<double> sigma;  // L singular values of B, descending
  double diff = 0.0;          // 1 - mev(Ucur, Upre) (valid if have_prev)
};

constexpr uint64_t kRowPad = 512;  // largest GEMM row tile (4 waves x 8 x 16 rows)

class Context {
 public:
  explicit Context(int device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;

  // ---------------------------------------------------------------- data
  void reserve(uint32_t nsamples, uint64_t nsnps);
  // Streamed genotypes, for matrices beyond HBM (the reference's -m reads,
  // src/FilePlink.cpp:122-298): the packed .bed rows stay in host memory, a
  // caller's buffer (kept alive and unchanged while the context uses it) or
  // the mapped payload of a .bed file, and each product loads the block of
  // at most block_snps SNP rows it covers into block-sized device layouts.
  // Opening streams every block once to count the codes (F, missing entries).
  void stream_host(const uint8_t* rows, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps);
  void stream_bed(const std::string& bed, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps);
  // The same from a .pgen: every block load decodes its records again.
  void stream_pgen(std::shared_ptr<const PgenSource> src, uint64_t block_snps);
  bool streamed() const { return src_rows_ != nullptr || pgen_ != nullptr; }
  uint64_t stream_block_snps() const { return sblock_; }
  uint64_t stream_loads = 0;   // blocks moved host -> device
  double stream_load_ms = 0.0; // host time spent loading them
  void upload_rows(const uint8_t* rows, uint64_t j0, uint64_t n);  // raw .bed rows (bps bytes each)
  void generate(const synth::Spec& spec);                          // synthetic data straight into HBM
  void download_rows(uint8_t* out);  // current column order, PLINK row format
  // Counts codes, sets F (FilePlink.cpp:37-55), resets the column state
  // (identity permutation, cF = F, sc = 1) and builds xt and the missing lists.
  void finalize();
  void filter_snps(const std::vector<uint32_t>& keep);  // keep SNPs (Data.cpp:88-105)
  // column i <- column perm[i] (RSVD.hpp:61-78).  The allele frequencies used
  // by standardize and the imputation clamp stay indexed by position (the
  // in-core shuffle) unless with_freq, where they follow their SNPs (a file
  // rewritten in permute_plink order before it is read).
  void permute_columns(const std::vector<int>& perm, bool with_freq = false);
  void standardize(int ploidy);                         // Data.cpp:351-362 semantics
  void set_impute(const double* U, const double* S, const double* V, int kk);  // Data.cpp:318-348 (column order)
  // The same from the device's own U (ucur) and V (vbuf), S on the host.
  void set_impute_device(const double* S, int kk);
  // EM bookkeeping on device-resident U and V (src/Halko.cpp:295-312):
  // flip_UV with V as the base, the previous V kept, mev(V, Vpre).
  void flip_uv_device(int k);
  void save_vpre();
  double mev_vpre(int k);
  void clear_impute();
  // Back to the state right after reading: original SNP order, centered and
  // unscaled, no imputation (each run starts from freshly read data).
  void reset_columns();

  uint32_t N = 0;
  uint64_t M = 0;
  uint64_t nmiss = 0;
  bool pgen_input = false;     // read from a .pgen: the -m shuffle is compute_pgen_perm's order
  std::vector<double> F;       // unpermuted allele frequencies (after filtering)
  std::vector<int> perm;       // device column i holds SNP perm[i]
  double p_miss() const { return N && M ? (double)nmiss / ((double)N * (double)M) : 0.0; }

  // ---------------------------------------------------------------- rsvd
  void configure(int precision, int L, const GemmTuning& t = {});
  int precision = 1;
  Form form_g = Form::FP32, form_h = Form::FP32;  // operand forms of G = A Omega and H = A^T G
  GemmShape shape_g, shape_h;
  int L = 0, lp = 0;
  GemmTuning tuning;

  void set_omega(const double* omega_colmajor);  // N x L, also Omg2 := Omega
  // Pinned host space for an N x L column-major test matrix; set_omega_staged()
  // sends what was written there, as set_omega does, without a host copy.
  double* omega_staging();
  void set_omega_staged();
  // G rows [j0, j1) = A[j0:j1] Omega and the H contribution of those SNPs:
  // hmode 0: H = ..., 1: H1 += ..., 2: H2 += ..., 3: H += ...
  void sketch(uint64_t j0, uint64_t j1, int hmode);
  // H (hmode 0; H1 1, H2 2) = A^T G~ over all SNPs for the G~ (M x lp,
  // row-major) the caller wrote into gt() (the FP64, FP32 and BF16 forms).
  void h_from_g(int hmode);
  double* gt() { return gt_.get(); }
  double* h() { return H_.get(); }
  // the resident operands, for drivers that build their own dense blocks
  const uint32_t* xs() const { return xs_.get(); }
  uint64_t pitch_s() const { return pitch_s_; }
  const double* sc() const { return sc_.get(); }
  const double* cf() const { return cF_.get(); }
  // Omega from an N x L column-major matrix the caller wrote on the device
  // into omega_scratch(): set_omega without the host copy.
  double* omega_scratch() { return tmp_n_.get(); }
  void set_omega_scratch() { omega_pieces(); }
  // Omega <- thin-Q(src) with Householder signs, then flipOmg against Omg2
  // unless !flip (the out-of-core --svd 1 branch has none);
  // src 0: H, 1: H1 + H2 (also stored as the current H).
  void omega_from(int src, bool flip = true);
  void zero_h(int which);  // 0: H, 1: H1, 2: H2
  // QR of G twice, B = R^-T H^T, eigen-decomposition of B B^T, Ucur; mev
  // against the previous Ucur if have_prev.  finish_epoch = finish_begin +
  // finish_end; between them a caller may enqueue the next epoch's device work
  // (omega_from, hold, sketch): finish_begin has sent the Gram matrices of G~
  // and H to the host, hold keeps this epoch's G~ and H aside for finish_end
  // and compute_v, and unhold puts them back when that next epoch is not run.
  EpochResult finish_epoch(int k, bool have_prev);
  void finish_begin();
  EpochResult finish_end(int k, bool have_prev);
  void hold();
  void unhold();
  // The next epoch's device work can be enqueued before finish_end: the thin-Q
  // of H runs on the device and no kernel timing waits on the host.
  bool overlap_ok() const;
  void save_prev();  // Upre := Ucur
  // Outputs of the last finish_epoch: U (N x k), V (M x k), column-major FP64.
  void download_u(int k, double* U);
  void compute_v(int k);  // V := Q U_B[:, :k] on device
  void download_v(int k, double* V);
  // G of the last sketches, rows [j0, j1) (current column order), and H / H1 / H2:
  // column-major FP64 with L columns.
  void download_g(uint64_t j0, uint64_t j1, double* G);
  void download_h(int which, double* H);
  const double* device_v() const { return vbuf_.get(); }
  // Data::write_residuals (src/Data.cpp:242-291).  U (N x k), S (k) and V
  // (M x k, current column order) column-major, subtracted as U S V^T (k = 0:
  // --ld-stats 1, nothing subtracted).  as_read: the out-of-core branch, whose
  // blocks are read again centered and unscaled without imputation; otherwise
  // the matrix the last products multiplied (scale, imputation).
  void residuals_prepare(int k, const double* U, const double* S, const double* V, bool as_read);
  // Rows [j0, j1) (current column order) of the centered residual matrix:
  // (j1 - j0) x N float32 into host memory.
  void residuals_rows(uint64_t j0, uint64_t j1, float* out);

  // timing of the hot kernels (HIP events on the context stream)
  struct KernelTimes {
    double g_ms = 0, h_ms = 0;
    uint64_t g_launches = 0, h_launches = 0;
  } ktimes;
  bool time_kernels = false;
  // how often the orthonormalization safeguards ran (Householder thin-Q of H,
  // second pass over G~)
  uint64_t householder_fallbacks = 0, gram_second_passes = 0;
  // winSVD epochs: with deferred_thinq, every Omega update of the epoch runs the
  // device thin-Q with its safety test on the device (no host round trip); an
  // epoch in which one update failed the test is rolled back to its
  // checkpoint and run again with the host test before each update.
  bool deferred_thinq = false;
  bool deferred_ok() const;          // the device thin-Q serves this configuration
  void epoch_checkpoint();           // H1, H2 and the Omega state; clears the status
  bool epoch_thinq_failed();         // synchronizes
  void epoch_rollback();
  uint64_t epoch_rollbacks = 0;

 private:
  // streamed genotypes
  const uint8_t* src_rows_ = nullptr;  // packed rows in file order (host)
  void* map_ = nullptr;                // mapped .bed file (stream_bed)
  hipStream_t copy_stream_ = nullptr;  // upload and block DMAs beside the interleaves (created on first use)
  std::shared_ptr<const PgenSource> pgen_;  // .pgen records decoded per block (stream_pgen)
  size_t map_len_ = 0;
  uint64_t cus_ = 0;                   // compute units of the device (split-K planning)
  uint64_t sblock_ = 0;                // device block capacity in SNP rows
  uint64_t blk0_ = 0, blk1_ = 0;       // loaded rows [blk0_, blk1_) of the current column order; blk1_ == 0: none
  uint64_t rows_state_ = 0;            // rows of the per-SNP device state (M padded)
  uint64_t lnmiss_ = 0, lrows_ = 0;    // missing entries / SNP rows of the current missing lists
  void stream_setup(uint32_t nsamples, uint64_t nsnps, uint64_t block_snps);
  void stream_close();
  void load_block(uint64_t j0, uint64_t j1);
  void block_rows_to_device(uint64_t j0, uint64_t n);  // rows of the current order -> xs_ (device)
  void build_lists(uint64_t rows, uint64_t ncodes_t);  // missing lists of the loaded layouts
  uint64_t xoff() const { return streamed() ? blk0_ : 0; }
  // data layout
  uint64_t bps_ = 0;
  uint64_t ks_ = 0, kt_ = 0;           // K extents (codes) of xs and xt
  uint64_t pitch_s_ = 0, pitch_t_ = 0; // words per row
  uint64_t rows_s_ = 0, rows_t_ = 0;   // allocated rows
  DevBuf<uint32_t> xs_, xt_, tmpx_;
  DevBuf<uint8_t> raw_;
  PinnedBuf pinned_;
  PinnedBuf omega_pinned_;  // the run's test matrix, drawn straight into pinned memory
  // epoch overlap: this epoch's G~ / H set aside while the next one is computed
  DevBuf<double> gt_alt_, H_alt_;
  bool held_ = false;
  PinnedBuf gram_pinned_;  // Gram matrices of G~ and H, thin-Q status (finish_begin)
  hipEvent_t gram_ev_ = nullptr;
  bool gram_h_sent_ = false;
  uint64_t begin_h_version_ = 0;
  DevBuf<double> dF_, cF_, sc_;
  DevBuf<uint64_t> rp_s_, rp_t_;
  DevBuf<uint32_t> ix_s_, ix_t_;
  DevBuf<double> w_s_, w_t_;   // per-entry weights of the missing-entry corrections
  bool weights_dirty_ = true;
  void update_missing_weights();
  // The missing-entry lists in tiles (FP32 weights in tile order) for the
  // reduced forms with L <= kMissMaxCp: a tile's source rows are read once
  // per tile from LDS instead of once per entry from L2.
  struct MissTiles {
    std::vector<MissTile> h_tiles;
    std::vector<uint32_t> h_tb;
    DevBuf<MissTile> tiles;
    DevBuf<uint32_t> tb;
    DevBuf<uint16_t> trow, eix;
    DevBuf<float> ew;
    DevBuf<uint32_t> rperm, rpos;  // rows by descending entry count (balance), else empty
    uint64_t nrows = 0, kext = 0, nblocks = 0;
    uint32_t span = 0;
    bool built = false;
    MissTileList list() const {
      return {tiles.get(), tb.get(), trow.get(), eix.get(), rperm.size() ? rperm.get() : nullptr};
    }
  };
  MissTiles mt_s_, mt_t_;
  bool miss_tiled_ = false;     // this run's corrections use the tiles
  bool weights_tiled_ = false;  // the current weights are the tiles' (else the CSR lists')
  // balance: place the rows in the row blocks by descending entry count (the
  // by-sample list, whose corrections always take all rows)
  void build_tiles(MissTiles& t, const DevBuf<uint64_t>& rp, const DevBuf<uint32_t>& ix, uint64_t nrows,
                   uint64_t kext, uint32_t span, bool balance);
  // Work items of a tiled correction over list rows [r_lo, r_hi) and indexes
  // [k_lo, k_hi): each row block's tiles in shares of about equal tile
  // counts, kept per range (the windows of an epoch repeat every epoch).
  struct MissWorkSet {
    DevBuf<MissWork> work;
    DevBuf<uint32_t> bslots;
    uint32_t nwork = 0, slots = 1;
  };
  std::map<std::array<uint64_t, 6>, MissWorkSet> mwork_;
  const MissWorkSet& miss_work(const MissTiles& t, uint64_t r_lo, uint64_t r_hi, uint64_t k_lo, uint64_t k_hi);
  // compact sources of the sparse corrections (Omega~ rows, G~ window rows)
  DevBuf<unsigned char> csg_, csh_;
  DevBuf<double> mpart_;
  int cp_ = 0;
  int cpitch_ = 0;  // elements between compact source rows (whole 128-byte lines)
  bool miss_f32_ = true;
  uint64_t miss_span(uint64_t nrows, uint64_t kext) const;
  // EMU state
  int emu_k_ = 0;
  int emu_kp_ = 0;  // row pitch of us_ / vimp_ (whole 128-byte lines, >= emu_k_ + 3)
  DevBuf<double> us_, vimp_;
  // rsvd workspace
  DevBuf<double> omega_, omega_r_, omega_c_, colsum_, bvec_;
  // I8 form: column scales and epilogue factors of Omega (G product) and of
  // diag(sc) G~ (H product), [scale_g | mul_g | scale_h | mul_h] x 128, and
  // the column-maximum scratch.
  DevBuf<double> i8s_;
  DevBuf<unsigned long long> i8max_;
  DevBuf<unsigned char> wg_, wh_;
  uint64_t wg_pitch_ = 0, wh_pitch_ = 0;
  DevBuf<double> gt_, H_, H1_, H2_, tmp_n_, tmp_m_, ucur_, upre_, vbuf_, vpre_;
  DevBuf<unsigned char> part_;
  DevBuf<double> small_, red_;
  uint64_t red_cap_ = 0;
  // host copies of the last epoch's small factors
  la::Mat W_;         // Q = G~ W_ orthonormal (W_ = R^-1 of QR(G) when G~ has full rank), L x L
  la::Mat WB_;        // eigenvectors of B B^T, L x L
  std::vector<double> sigma_;
  uint64_t h_version_ = 0, gram_version_ = ~0ull;
  la::Mat gram_h_;    // Gram of H (L x L) for h_version_
  // device side of the thin-Q of H (lp <= 48): Gram of H for gram_dev_version_,
  // second Gram, R1, column signs (lp x lp each, signs lp); breakdown status
  DevBuf<double> res_ut_, res_sv_, res_part_;
  DevBuf<float> res_out_;
  int res_k_ = 0;
  bool res_as_read_ = false;
  DevBuf<double> ldev_;
  DevBuf<int> lstatus_;
  DevBuf<double> ck_H1_, ck_H2_, ck_omega_, ck_omega_r_, ck_colsum_, ck_i8s_;  // epoch checkpoint
  DevBuf<unsigned char> ck_wg_;
  uint64_t gram_dev_version_ = ~0ull;

  void build_derived();  // xt + missing lists from xs
  void pristine_columns();  // cF = F as read, sc = 1 (device)
  bool perm_identity_ = true;  // perm[i] == i
  bool freq_follows_ = false;  // the device allele frequencies are in the current column order
  // L x L block of X^T Y; with dev_out the lp x lp result also stays there
  la::Mat gram_l(const double* X, uint64_t rows, const double* Y = nullptr, double* dev_out = nullptr);
  void omega_pieces();  // tmp_n_ (column-major Omega) -> omega_c_, pieces, Omg2, column sums
  // I8 column scales / epilogue factors of the G (h = 0) or H (h = 1)
  // product's operand; null when that product is not in the I8 form.
  // column scales and epilogue factors of the column-scaled forms (null otherwise)
  double* i8_scale(int h) { return form_colscaled(h ? form_h : form_g) ? i8s_.get() + 256 * h : nullptr; }
  double* i8_mul(int h) { return form_colscaled(h ? form_h : form_g) ? i8s_.get() + 256 * h + 128 : nullptr; }
  void set_t(const la::Mat& T);
  void h_half(uint64_t j0, uint64_t j1, int hmode);  // the H product of sketch for G~ rows [j0, j1)
  // Orthonormalization safeguards (Cholesky pivots d_j over the largest diagonal entry).
  static constexpr double kOnePassMinPivot = 1e-6;  // one CholeskyQR pass of G~ is enough above this
  static constexpr double kThinqMinPivot = 1e-10;   // CholeskyQR2 is used above this, else the fallbacks
  static constexpr double kRankTol = 1e-12;         // eigenvalues of a Gram kept above this x the largest
  la::Mat from_pinned_gram(int which) const;        // 0: Gram of G~, 1: Gram of H (finish_begin)
  bool h_thinq_safe();                              // CholeskyQR2 of the current H is safe
  uint64_t h_safe_version_ = ~0ull;
  bool h_safe_ = false;
  void omega_householder(bool flip);                // Omega <- Householder thin Q of H (host)
  void orthonormal_factor(const la::Mat& G1, const double* g);  // sets W_
  void download_rowmajor(const double* src, uint64_t rows, double* out);  // rows x lp -> column-major rows x L
  // host <-> device copies of count doubles through the pinned staging buffer
  void stage_in(const double* host, uint64_t count, double* dev);
  void stage_out(const double* dev, uint64_t count, double* host);
  bool cols_pristine_ = false;  // column state equals the as-read state (identity order, cF = F, sc = 1)  // uploads an L x L matrix into small_ as lp x lp (zero padded)
  void run_gemm(bool g_product, const GemmShape& shape, uint64_t nchunks, const GemmArgs& a);
  bool fixed_g_ = false, fixed_h_ = false;  // shapes set by the caller (no per-launch change)
};

}  // namespace pcaone
