// Batched randomized SVD: many independent cohorts (each its own genotype
// matrix) run --svd 1 (src/Halko.cpp:46-127, 271-345) together on one GPU.
// All cohorts share k, oversamples, maxp, tol and the operand form; every
// product of every cohort of one step runs in one persistent grouped launch,
// the per-cohort L x L steps run on the device, and the whole run (all
// epochs) is captured in one HIP graph and replayed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "batch_kernels.hpp"
#include "context.hpp"
#include "device.hpp"
#include "rsvd.hpp"

namespace pcaone {

namespace synth {
struct Spec;
}

struct BatchResult {
  uint32_t k = 0;
  std::vector<int> status;      // 0 or PCAONE_ERR_* per cohort
  std::vector<int> epochs;      // epochs run per cohort
  std::vector<int> nsv;         // singular values produced per cohort
  std::vector<double> S;        // ncoh x k
  std::vector<double> trace;    // ncoh x kBatchMaxEpochs (1 - mev per epoch)
};

struct BatchTuning {
  GemmShape g, h;
  uint64_t kchunk_g = 0, kchunk_h = 0;
  double g_ms = 0, h_ms = 0;  // timed grouped launch of the chosen configuration
  int candidates = 0;         // configurations timed
};

class Batch {
 public:
  explicit Batch(Context& ctx);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // ---- data: reserve, then upload (or generate), then finalize
  void reserve(const std::vector<uint32_t>& nsamples, const std::vector<uint64_t>& nsnps);
  void upload_rows(uint32_t c, const uint8_t* rows, uint64_t j0, uint64_t n);
  void generate(const std::vector<synth::Spec>& specs);
  void finalize();
  uint32_t size() const { return (uint32_t)N_.size(); }
  uint32_t nsamples(uint32_t c) const { return N_[c]; }
  uint64_t nsnps(uint32_t c) const { return M_[c]; }
  const std::vector<double>& allele_freq(uint32_t c) const { return F_[c]; }

  // ---- run: every cohort's randomized SVD; results stay on the device
  // until read with get_u / get_v.  use_graph: capture and replay.
  void run(const Param& p, bool want_v, bool use_graph, BatchResult& out);
  void get_u(uint32_t c, double* U) const;  // N_c x k, column-major
  void get_v(uint32_t c, double* V) const;  // M_c x k, column-major (run with want_v)
  const BatchTuning& tuning() const { return tune_; }
  double last_seconds = 0;

 private:
  Context& ctx_;
  hipStream_t stream_;
  // data
  std::vector<uint32_t> N_;
  std::vector<uint64_t> M_;
  std::vector<uint64_t> ks_, kt_, pitch_s_, pitch_t_, rows_s_, rows_t_, off_s_, off_t_;
  DevBuf<uint32_t> xs_, xt_;
  std::vector<std::vector<double>> F_;
  bool finalized_ = false;

  // run workspace (sized for form, L and want_v)
  struct Layout {
    int form = -1;
    int L = 0;
    bool want_v = false;
    uint64_t kchunk_g = 0, kchunk_h = 0;
    GemmShape g, h;
    bool operator==(const Layout& o) const {
      return form == o.form && L == o.L && want_v == o.want_v && kchunk_g == o.kchunk_g && kchunk_h == o.kchunk_h &&
             g.wm == o.g.wm && g.waves == o.g.waves && h.wm == o.h.wm && h.waves == o.h.waves;
    }
  } layout_;
  DevBuf<double> mats_, scf_, vecs_, partial_;
  DevBuf<unsigned char> pieces_, part_;
  DevBuf<BatchSmall> small_;
  DevBuf<BatchCohort> coh_;
  DevBuf<GemmCohort> gcoh_g_, gcoh_h_;
  DevBuf<GemmItem> items_g_, items_h_;
  DevBuf<RowItem> items_m_, items_n_, items_e_;
  DevBuf<int> active_;
  uint32_t n_items_g_ = 0, n_items_h_ = 0, n_items_m_ = 0, n_items_n_ = 0, n_items_e_ = 0;
  unsigned grid_g_ = 0, grid_h_ = 0;
  std::vector<BatchCohort> host_coh_;
  std::vector<uint64_t> part_cap_;  // bytes of each cohort's partial-product block
  uint64_t start_key_ = 0;          // (seed, L, rand, ploidy) of the uploaded Omega, sc, cF (+1)
  BatchTuning tune_;
  std::map<std::tuple<int, int>, BatchTuning> tuned_;  // by (form, L)
  // graph of the last run configuration
  hipGraphExec_t graph_ = nullptr;
  std::tuple<int, int, int, int, bool, double> graph_key_{-1, 0, 0, 0, false, 0.0};
  int last_k_ = 0;
  std::vector<int> stop_epoch_;

  void prepare(int form, int L, bool want_v, const GemmShape& g, const GemmShape& h, uint64_t kchunk_g,
               uint64_t kchunk_h);
  void build_gemm_items(bool g_product, const GemmShape& sh, uint64_t kchunk, std::vector<GemmItem>& items,
                        std::vector<uint64_t>& nch) const;
  void upload_start(const Param& p, int L);
  void enqueue_epochs(const BatchArgs& a, int p, bool want_v);
  void grouped(bool g_product);
  BatchTuning autotune(int form, int L);
  void release_graph();
};

}  // namespace pcaone
