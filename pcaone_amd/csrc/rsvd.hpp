// Host-side mirror of PCAone's randomized-SVD operator interface, backed by
// the device context.  Names, flags and control flow follow the reference:
//   Param                          src/Cmd.hpp:16-98, derived rules src/Cmd.cpp:171-238
//   RsvdOpData / Normal / Fancy    src/Halko.hpp:6-93, src/Halko.cpp:15-269
//   run_pca_with_halko             src/Halko.cpp:271-345
// The genotype matrix and the G/H sketches live in HBM; only U, S (and V
// when needed) cross back to the host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "context.hpp"

namespace pcaone {

struct Param {
  int svd = 2;  // 1: single-pass RSVD (sSVD), 2: window-based RSVD (winSVD)
  uint32_t k = 10, oversamples = 10, maxp = 20, bands = 64, maxiter = 100;
  double tol = 1e-4, tolem = 1e-5, maf = 0.0;
  int seed = 112, rand = 1, scale = -9, ploidy = 2, verbose = 0;
  bool noshuffle = false, emu = false, printv = false;
  int precision = 1;  // PCAONE_PREC_*
  GemmTuning tuning;
  double memory = 0.0;  // -m/--memory (GB); > 0 selects the out-of-core schedule
  uint32_t buffer = 2;  // --buffer (GB) of permute_plink
  bool ld = false;      // -D/--ld: unscaled run, residual matrix written (src/Cmd.cpp:92)
  int ld_stats = 0;     // --ld-stats (src/Cmd.cpp:114)
  uint32_t imaxiter = 1000;  // --imaxiter (IRAM)
  double itol = 1e-6;        // --itol (IRAM)
  bool stream = false;       // run_bfile: stream the .bed blocks instead of loading the matrix
  bool hardcall = false;     // --hardcall: the hardcall track of a .pgen that also stores dosages
  uint32_t threads = 12;     // -n/--threads: the reader threads that shape a .pgen's -m shuffle
  uint32_t ncv = 20;         // derived: max(20, 2k + 1) (src/Cmd.cpp:215)
  // derived (src/Cmd.cpp:171-238)
  bool perm = false, missme = false, filterSNP = false, out_of_core = false;
  void derive();
};

// Data::prepare's out-of-core block plan (src/Data.cpp:42-84, the Halko
// branch with dopca): SNP blocks [start[b], stop[b]] and the winSVD band factor.
struct BlockPlan {
  uint64_t blocksize = 0, nblocks = 1, band_factor = 1;
  std::vector<uint64_t> start, stop;
};
BlockPlan plan_blocks(const Param& p, uint64_t nsamples, uint64_t nsnps);

// permute_plink's SNP order (src/FilePlink.cpp:303-408): position i of the
// permuted file holds original SNP order[i].
std::vector<int> permute_plink_order(uint64_t nsamples, uint64_t nsnps, uint32_t gb, uint32_t nbands);

struct RunInfo {
  int epochs = 0;                 // epochs of the last computeUSV
  int total_epochs = 0;           // over all computeUSV calls (EMU)
  int em_iters = 0;
  std::vector<double> trace;      // diff of every checked epoch (all computeUSV calls)
  std::vector<double> em_trace;   // EM convergence diffs
  double seconds = 0.0;           // wall time of run_pca_with_halko
  double omega_ms = 0.0;          // host time of the Omega draws (part of seconds)
};

class RsvdOpData {
 public:
  RsvdOpData(Context& ctx, const Param& p, int k, int os);
  virtual ~RsvdOpData() = default;

  uint64_t rows() const { return ctx_.M; }  // SNPs
  uint64_t cols() const { return ctx_.N; }  // samples
  int ranks() const { return nk_; }
  int oversamples() const { return os_; }
  int size() const { return nk_ + os_; }

  virtual void computeGandH(int pi) = 0;
  // Enqueues epoch pi's device work while the host steps of epoch pi - 1 are
  // still to come (the previous epoch's G~ and H held aside); false where the
  // operator cannot.
  virtual bool enqueue_next(int pi) {
    (void)pi;
    return false;
  }
  void setFlags(bool is_update, bool is_standardize) {
    update = is_update;
    standardize = is_standardize;
  }
  void computeUSV(int p, double tol);
  void initOmg();

  bool update = false, standardize = false;
  std::vector<double> U, S, V;  // N x k, k, M x k (column-major; V in device column order)
  bool need_v_host = false;
  double* u_out = nullptr;  // when set (and no EMU), U is written there instead of into U
  bool device_uv = false;   // U and V stay on the device (EM iterations); U, V left empty
  RunInfo* info = nullptr;

 protected:
  Context& ctx_;
  const Param& params_;
  int nk_, os_;
  void fit_with_pi();  // src/Data.cpp:293-349 via the device impute state
};

class NormalRsvdOpData : public RsvdOpData {
 public:
  using RsvdOpData::RsvdOpData;
  void computeGandH(int pi) override;
  bool enqueue_next(int pi) override;
  const BlockPlan* plan = nullptr;  // out-of-core block plan (null: in-core)

 private:
  void products();  // G and H over all SNPs (block by block on streamed genotypes)
};

class FancyRsvdOpData : public RsvdOpData {
 public:
  using RsvdOpData::RsvdOpData;
  void computeGandH(int pi) override;
  uint64_t omega_updates = 0;
  const BlockPlan* plan = nullptr;  // out-of-core block plan (null: in-core)

 private:
  uint64_t bandsize_ = 1, blocksize_ = 0;
  void windows(int pi);         // the epoch's window schedule (in-core)
  void windows_blocks(int pi);  // the same over the out-of-core plan's blocks
};

struct PcaResult {
  std::vector<double> U, S, E, V;  // V in the reference's SNP order (un-permuted)
  int k = 0;
};

// U_out (nullable, N x k column-major): receives U directly; res.U is then empty.
PcaResult run_pca_with_halko(Context& ctx, const Param& params, RunInfo* info, double* U_out = nullptr);
// The PCA the params select (src/Main.cpp:176-217): --svd 0 IRAM, 1/2 the
// randomized SVD, 3 the exact eigen-decomposition.
PcaResult run_pca(Context& ctx, const Param& params, RunInfo* info, double* U_out = nullptr);


}  // namespace pcaone
