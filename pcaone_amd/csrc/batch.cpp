// Batched randomized SVD driver (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "kernels.hpp"
#include "rng_host.hpp"
#include "synth_host.hpp"

namespace pcaone {

namespace {
constexpr double kVarTol = 1e-9;  // src/Data.hpp:7
constexpr int LP = kBatchLp;
const uint64_t kChunkCandG[] = {2048, 4096, 8192};
const uint64_t kChunkCandH[] = {4096, 8192, 16384};
const GemmShape kShapeCand[] = {{8, 4}, {4, 8}, {4, 4}, {2, 8}, {2, 4}, {1, 4}};

template <class T>
void upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
  d.resize(std::max<size_t>(h.size(), 1));
  if (!h.empty()) PCAONE_HIP(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
}  // namespace

Batch::Batch(Context& ctx) : ctx_(ctx), stream_(ctx.stream) {}

Batch::~Batch() { release_graph(); }

void Batch::release_graph() {
  if (graph_) (void)hipGraphExecDestroy(graph_);
  graph_ = nullptr;
}

// ------------------------------------------------------------------ data
void Batch::reserve(const std::vector<uint32_t>& nsamples, const std::vector<uint64_t>& nsnps) {
  require(!nsamples.empty() && nsamples.size() == nsnps.size(), kErrArgument, "batch: cohort lists differ in size");
  require(nsamples.size() < (1u << 20), kErrArgument, "batch: too many cohorts");
  const size_t n = nsamples.size();
  release_graph();
  layout_ = Layout{};
  tuned_.clear();
  finalized_ = false;
  N_ = nsamples;
  M_ = nsnps;
  ks_.assign(n, 0), kt_.assign(n, 0), pitch_s_.assign(n, 0), pitch_t_.assign(n, 0);
  rows_s_.assign(n, 0), rows_t_.assign(n, 0), off_s_.assign(n, 0), off_t_.assign(n, 0);
  F_.assign(n, {});
  uint64_t ws = 0, wt = 0;
  for (size_t c = 0; c < n; ++c) {
    require(N_[c] > 0 && M_[c] > 0, kErrArgument, "batch: empty cohort");
    require(M_[c] < (1ull << 32), kErrArgument, "batch: more than 2^32-1 SNPs in a cohort");
    ks_[c] = round_up(N_[c], kKStep);
    kt_[c] = round_up(M_[c], kKStep);
    pitch_s_[c] = ks_[c] / 16;
    pitch_t_[c] = kt_[c] / 16;
    rows_s_[c] = round_up(M_[c], 512);
    rows_t_[c] = round_up(N_[c], 512);
    off_s_[c] = ws;
    off_t_[c] = wt;
    ws += rows_s_[c] * pitch_s_[c];
    wt += rows_t_[c] * pitch_t_[c];
  }
  xs_.resize(ws);
  xt_.resize(wt);
  for (size_t c = 0; c < n; ++c) launch_fill_rows(xs_.get() + off_s_[c], pitch_s_[c], 0, rows_s_[c], stream_);
  PCAONE_HIP(hipStreamSynchronize(stream_));
}

void Batch::upload_rows(uint32_t c, const uint8_t* rows, uint64_t j0, uint64_t n) {
  require(c < size(), kErrArgument, "batch: cohort index out of range");
  require(j0 + n <= M_[c], kErrArgument, "batch: upload beyond the cohort's SNP count");
  if (!n) return;
  finalized_ = false;
  const uint64_t bps = bytes_per_snp(N_[c]);
  const uint64_t chunk = std::max<uint64_t>(1, (64ull << 20) / bps);
  DevBuf<uint8_t> raw(std::min(n, chunk) * bps);
  for (uint64_t r = 0; r < n; r += chunk) {
    const uint64_t m = std::min(chunk, n - r);
    PCAONE_HIP(hipMemcpyAsync(raw.get(), rows + r * bps, m * bps, hipMemcpyHostToDevice, stream_));
    launch_interleave(raw.get(), m, bps, N_[c], xs_.get() + off_s_[c] + (j0 + r) * pitch_s_[c], pitch_s_[c],
                      stream_);
    PCAONE_HIP(hipStreamSynchronize(stream_));
  }
}

void Batch::generate(const std::vector<synth::Spec>& specs) {
  require(specs.size() == size(), kErrArgument, "batch: one generator spec per cohort");
  for (uint32_t c = 0; c < size(); ++c) {
    synth::validate(specs[c]);
    require(specs[c].nsamples == N_[c] && specs[c].nsnps == M_[c], kErrArgument,
            "batch: generator spec dimensions differ from the reserved cohort");
    synth_rows(specs[c], xs_.get() + off_s_[c], pitch_s_[c], stream_);
  }
  finalized_ = false;
  PCAONE_HIP(hipStreamSynchronize(stream_));
}

void Batch::finalize() {
  require(size() > 0, kErrState, "batch: reserve first");
  release_graph();
  for (uint32_t c = 0; c < size(); ++c) {
    const uint64_t M = M_[c];
    DevBuf<uint32_t> cnt(4 * M);
    launch_code_counts(xs_.get() + off_s_[c], pitch_s_[c], M, N_[c], cnt.get(), stream_);
    std::vector<uint32_t> h(4 * M);
    PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), h.size() * 4, hipMemcpyDeviceToHost, stream_));
    PCAONE_HIP(hipStreamSynchronize(stream_));
    std::vector<double>& F = F_[c];
    F.assign(M, 0.0);
    for (uint64_t j = 0; j < M; ++j) {
      const uint64_t c00 = h[4 * j], c01 = h[4 * j + 1], c10 = h[4 * j + 2], c11 = h[4 * j + 3];
      require(c01 == 0, kErrArgument,
              "batch: cohort " + std::to_string(c) + " has missing genotypes (the batched form takes complete data)");
      const uint64_t n = c00 + c10 + c11;
      F[j] = n ? ((double)c00 + 0.5 * (double)c10) / (double)n : 0.0;  // src/FilePlink.cpp:37-55
    }
    launch_transpose_codes(xs_.get() + off_s_[c], pitch_s_[c], M, N_[c], xt_.get() + off_t_[c], pitch_t_[c],
                           rows_t_[c], stream_);
  }
  PCAONE_HIP(hipStreamSynchronize(stream_));
  layout_ = Layout{};
  finalized_ = true;
}

// ------------------------------------------------------------------ layout
void Batch::build_gemm_items(bool g_product, const GemmShape& sh, uint64_t kchunk, std::vector<GemmItem>& items,
                             std::vector<uint64_t>& nch) const {
  const uint64_t tile = (uint64_t)gemm_row_tile(sh);
  items.clear();
  nch.assign(size(), 0);
  for (uint32_t c = 0; c < size(); ++c) {
    const uint64_t rows = round_up(g_product ? M_[c] : N_[c], tile);
    const uint64_t K = g_product ? ks_[c] : kt_[c];
    const uint64_t n = (K + kchunk - 1) / kchunk;
    nch[c] = n;
    for (uint64_t r0 = 0; r0 < rows; r0 += tile)
      for (uint64_t z = 0; z < n; ++z)
        items.push_back(GemmItem{c, (uint32_t)z, r0, z * kchunk, std::min(K, (z + 1) * kchunk)});
  }
  // longest items first, so the persistent workgroups finish together
  std::stable_sort(items.begin(), items.end(),
                   [](const GemmItem& a, const GemmItem& b) { return a.k1 - a.k0 > b.k1 - b.k0; });
}

void Batch::prepare(int form, int L, bool want_v, const GemmShape& g, const GemmShape& h, uint64_t kchunk_g,
                    uint64_t kchunk_h) {
  Layout want;
  want.form = form, want.L = L, want.want_v = want_v, want.kchunk_g = kchunk_g, want.kchunk_h = kchunk_h;
  want.g = g, want.h = h;
  if (layout_ == want) return;
  release_graph();
  const Form f = (Form)form;
  require(gemm_shape_supported(f, g, LP) && gemm_shape_supported(f, h, LP), kErrArgument,
          "batch: unsupported GEMM shape");
  const uint32_t n = size();
  const size_t wsz = form_wbytes(f), P = form_pieces(f), acc = form_accbytes(f);
  const bool realloc = layout_.form != form || layout_.want_v != want_v;
  // ---- memory (allocated for the smallest candidate chunks, reused by every configuration)
  std::vector<uint64_t> o_mat(n * kNumBatchBufs), o_scf(n), o_wg(n), o_wh(n), o_part(n), part_cap(n);
  uint64_t nm = 0, nsc = 0, npc = 0, npart = 0;
  for (uint32_t c = 0; c < n; ++c) {
    for (int b = 0; b < kNumBatchBufs; ++b) {
      const bool mrows = b == kBufG || b == kBufV;
      const uint64_t rows = (b == kBufV && !want_v) ? 0 : (mrows ? M_[c] : N_[c]);
      o_mat[c * kNumBatchBufs + b] = nm;
      nm += rows * LP;
    }
    o_scf[c] = nsc;
    nsc += 2 * M_[c];
    o_wg[c] = npc;
    npc += P * LP * ks_[c] * wsz;
    o_wh[c] = npc;
    npc += P * LP * kt_[c] * wsz;
    npc = round_up(npc, 256);
    o_part[c] = npart;
    const uint64_t pg = (ks_[c] + kChunkCandG[0] - 1) / kChunkCandG[0] * rows_s_[c];
    const uint64_t ph = (kt_[c] + kChunkCandH[0] - 1) / kChunkCandH[0] * rows_t_[c];
    const uint64_t pgu = (ks_[c] + kchunk_g - 1) / kchunk_g * rows_s_[c];
    const uint64_t phu = (kt_[c] + kchunk_h - 1) / kchunk_h * rows_t_[c];
    part_cap[c] = std::max({pg, ph, pgu, phu}) * LP * acc;
    npart += round_up(part_cap[c], 256);
  }
  if (realloc || layout_.L == 0) {
    mats_.resize(nm);
    PCAONE_HIP(hipMemsetAsync(mats_.get(), 0, mats_.bytes(), stream_));
    pieces_.resize(npc);
    PCAONE_HIP(hipMemsetAsync(pieces_.get(), 0, pieces_.bytes(), stream_));
    small_.resize(n);
    PCAONE_HIP(hipMemsetAsync(small_.get(), 0, small_.bytes(), stream_));
    start_key_ = 0;  // Omega, sc and cF are re-sent
  }
  scf_.resize(nsc);
  vecs_.resize((uint64_t)n * 2 * LP);
  active_.resize(n);
  part_.resize(npart);
  part_cap_ = part_cap;
  // ---- item lists
  std::vector<RowItem> im, in, ie;
  host_coh_.assign(n, BatchCohort{});
  for (uint32_t c = 0; c < n; ++c) {
    BatchCohort& b = host_coh_[c];
    b.N = N_[c];
    b.M = M_[c];
    for (int k = 0; k < kNumBatchBufs; ++k) b.mat[k] = mats_.get() + o_mat[c * kNumBatchBufs + k];
    if (!want_v) b.mat[kBufV] = nullptr;
    b.wg = pieces_.get() + o_wg[c];
    b.wg_pitch = ks_[c];
    b.wh = pieces_.get() + o_wh[c];
    b.wh_pitch = kt_[c];
    b.sc = scf_.get() + o_scf[c];
    b.cf = scf_.get() + o_scf[c] + M_[c];
    b.colsum = vecs_.get() + (uint64_t)c * 2 * LP;
    b.bvec = b.colsum + LP;
    b.part = part_.get() + o_part[c];
    b.st = small_.get() + c;
    b.m_begin = (uint32_t)im.size();
    for (uint64_t r = 0; r < M_[c]; r += kBatchRowsPerItem)
      im.push_back(RowItem{c, (uint32_t)std::min<uint64_t>(kBatchRowsPerItem, M_[c] - r), r});
    b.m_end = (uint32_t)im.size();
    b.n_begin = (uint32_t)in.size();
    for (uint64_t r = 0; r < N_[c]; r += kBatchRowsPerItem)
      in.push_back(RowItem{c, (uint32_t)std::min<uint64_t>(kBatchRowsPerItem, N_[c] - r), r});
    b.n_end = (uint32_t)in.size();
    b.e_begin = (uint32_t)ie.size();
    for (uint64_t r = 0; r < M_[c]; r += kBatchEpiRows)
      ie.push_back(RowItem{c, (uint32_t)std::min<uint64_t>(kBatchEpiRows, M_[c] - r), r});
    b.e_end = (uint32_t)ie.size();
  }
  // ---- grouped GEMM items and tables
  std::vector<GemmItem> ig, ih;
  std::vector<uint64_t> nch_g, nch_h;
  build_gemm_items(true, g, kchunk_g, ig, nch_g);
  build_gemm_items(false, h, kchunk_h, ih, nch_h);
  require(ig.size() < (1ull << 32) && ih.size() < (1ull << 32), kErrArgument, "batch: too many GEMM tiles");
  std::vector<GemmCohort> tg(n), th(n);
  const uint64_t tile_g = gemm_row_tile(g), tile_h = gemm_row_tile(h);
  for (uint32_t c = 0; c < n; ++c) {
    BatchCohort& b = host_coh_[c];
    const uint64_t rg = round_up(M_[c], tile_g), rh = round_up(N_[c], tile_h);
    require(rg <= rows_s_[c] && rh <= rows_t_[c], kErrInternal, "batch: tile rows exceed the padded rows");
    require(nch_g[c] * rg * LP * acc <= part_cap[c] && nch_h[c] * rh * LP * acc <= part_cap[c], kErrInternal,
            "batch: partial products exceed their buffer");
    b.nch_g = nch_g[c];
    b.ychunk_g = rg * LP;
    b.nch_h = nch_h[c];
    b.ychunk_h = rh * LP;
    tg[c] = GemmCohort{xs_.get() + off_s_[c], pitch_s_[c], b.wg, ks_[c], (uint64_t)LP * ks_[c], b.part, b.ychunk_g};
    th[c] = GemmCohort{xt_.get() + off_t_[c], pitch_t_[c], b.wh, kt_[c], (uint64_t)LP * kt_[c], b.part, b.ychunk_h};
  }
  upload(coh_, host_coh_, stream_);
  upload(gcoh_g_, tg, stream_);
  upload(gcoh_h_, th, stream_);
  upload(items_g_, ig, stream_);
  upload(items_h_, ih, stream_);
  upload(items_m_, im, stream_);
  upload(items_n_, in, stream_);
  upload(items_e_, ie, stream_);
  n_items_g_ = (uint32_t)ig.size();
  n_items_h_ = (uint32_t)ih.size();
  n_items_m_ = (uint32_t)im.size();
  n_items_n_ = (uint32_t)in.size();
  n_items_e_ = (uint32_t)ie.size();
  // Gram partials are LP x LP per item; the epilogue and row sums need 2 LP per item
  partial_.resize(std::max<uint64_t>((uint64_t)std::max({n_items_m_, n_items_n_, 1u}) * LP * LP,
                                     (uint64_t)std::max({n_items_e_, n_items_n_, 1u}) * 2 * LP));
  int cus = 0;
  PCAONE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx_.device));
  grid_g_ = (unsigned)std::min<uint64_t>(n_items_g_, (uint64_t)cus * grouped_gemm_blocks_per_cu(f, g, LP));
  grid_h_ = (unsigned)std::min<uint64_t>(n_items_h_, (uint64_t)cus * grouped_gemm_blocks_per_cu(f, h, LP));
  PCAONE_HIP(hipStreamSynchronize(stream_));
  layout_ = want;
}

void Batch::upload_start(const Param& p, int L) {
  const uint64_t key = ((uint64_t)(uint32_t)p.seed << 32) ^ ((uint64_t)L << 16) ^ ((uint64_t)(p.rand != 0) << 8) ^
                       (uint64_t)(uint32_t)p.ploidy;
  if (start_key_ == key + 1) return;
  // sc and cF per SNP: the plain path always standardizes (src/Halko.cpp:283-289,
  // src/Data.cpp:351-362)
  for (uint32_t c = 0; c < size(); ++c) {
    std::vector<double> scf(2 * M_[c]);
    for (uint64_t j = 0; j < M_[c]; ++j) {
      const double fq = F_[c][j];
      const double sd = std::sqrt(fq * (1.0 - fq));
      scf[j] = sd > kVarTol ? std::sqrt((double)p.ploidy) / sd : 1.0;
      scf[M_[c] + j] = fq;
    }
    PCAONE_HIP(hipMemcpyAsync(const_cast<double*>(host_coh_[c].sc), scf.data(), scf.size() * 8,
                              hipMemcpyHostToDevice, stream_));
    PCAONE_HIP(hipStreamSynchronize(stream_));
  }
  // Omega of every cohort: the libstdc++ draw (src/Halko.cpp:15-23), N_c x L
  // column-major, stored row-major with LP columns in kBufOmega0.
  std::map<uint32_t, std::vector<double>> byn;
  for (uint32_t c = 0; c < size(); ++c) {
    auto& om = byn[N_[c]];
    if (om.empty()) {
      std::vector<double> cm((uint64_t)N_[c] * L);
      rng::omega(p.seed, p.rand != 0, N_[c], (uint64_t)L, cm.data());
      om.assign((uint64_t)N_[c] * LP, 0.0);
      for (uint64_t i = 0; i < N_[c]; ++i)
        for (int l = 0; l < L; ++l) om[i * LP + l] = cm[(uint64_t)l * N_[c] + i];
    }
    PCAONE_HIP(hipMemcpyAsync(host_coh_[c].mat[kBufOmega0], om.data(), om.size() * 8, hipMemcpyHostToDevice,
                              stream_));
  }
  PCAONE_HIP(hipStreamSynchronize(stream_));
  start_key_ = key + 1;
}

// ------------------------------------------------------------------ epochs
void Batch::grouped(bool g_product) {
  GroupedGemmArgs a{};
  a.items = g_product ? items_g_.get() : items_h_.get();
  a.nitems = g_product ? n_items_g_ : n_items_h_;
  a.cohorts = g_product ? gcoh_g_.get() : gcoh_h_.get();
  a.active = active_.get();
  launch_grouped_gemm((Form)layout_.form, g_product ? layout_.g : layout_.h, LP, a, g_product ? grid_g_ : grid_h_,
                      stream_);
}

void Batch::enqueue_epochs(const BatchArgs& a, int p, bool want_v) {
  hipStream_t s = stream_;
  double* part = partial_.get();
  launch_batch_small(a, kSmallInit, 0, s);
  for (int pi = 0; pi <= p; ++pi) {
    // computeGandH (src/Halko.cpp:99-127): Omega (pi == 0) or thin-Q(H) is ready
    if (pi == 0) {
      if (form_colscaled((Form)layout_.form))
        launch_batch_omega_colmax(a, items_n_.get(), n_items_n_, kBufOmega0, s);
      launch_batch_make_pieces(a, items_n_.get(), n_items_n_, kBufOmega0, false, s);
      launch_batch_colsum(a, items_n_.get(), n_items_n_, part, s);
    }
    launch_batch_active(a, active_.get(), s);
    grouped(true);
    launch_batch_epi_g(a, items_e_.get(), n_items_e_, part, s);
    grouped(false);
    launch_batch_epi_h(a, items_n_.get(), n_items_n_, s);
    // computeUSV (src/Halko.cpp:55-77): R of G, B B^T, Ucur, mev
    launch_batch_gram(a, items_m_.get(), n_items_m_, true, kBufG, -1, 0, 1, part, s);
    launch_batch_gram(a, items_n_.get(), n_items_n_, false, kBufH, -1, 1, 1, part, s);
    launch_batch_small(a, kSmallEpoch, pi, s);
    const int ucur = (pi & 1) ? kBufU1 : kBufU0, upre = (pi & 1) ? kBufU0 : kBufU1;
    launch_batch_apply(a, items_n_.get(), n_items_n_, kBufH, ucur, 0, 1, s);
    if (pi > 0) launch_batch_gram(a, items_n_.get(), n_items_n_, false, ucur, upre, 0, 1, part, s);
    launch_batch_small(a, kSmallStop, pi, s);
    // stopping cohorts: loadings V = Q U_B[:, :k] (src/Halko.cpp:84)
    if (want_v) launch_batch_apply(a, items_m_.get(), n_items_m_, kBufG, kBufV, 1, 2, s);
    launch_batch_small(a, kSmallFinish, pi, s);
    if (pi == p) break;
    // next Omega: thin-Q(H) with Householder signs, then flipOmg (src/Halko.cpp:121-123)
    launch_batch_small(a, kSmallOmega1, pi, s);
    launch_batch_apply(a, items_n_.get(), n_items_n_, kBufH, kBufTmpN, 0, 1, s);
    launch_batch_gram(a, items_n_.get(), n_items_n_, false, kBufTmpN, -1, 0, 1, part, s);
    launch_batch_small(a, kSmallOmega2, pi, s);
    launch_batch_apply(a, items_n_.get(), n_items_n_, kBufH, kBufOmegaC, 0, 1, s);
    launch_batch_flip(a, items_n_.get(), n_items_n_, part, s);
    launch_batch_small(a, kSmallFlip, pi, s);
    if (form_colscaled((Form)layout_.form))
      launch_batch_omega_colmax(a, items_n_.get(), n_items_n_, kBufOmegaC, s);
    launch_batch_make_pieces(a, items_n_.get(), n_items_n_, kBufOmegaC, true, s);
    launch_batch_colsum(a, items_n_.get(), n_items_n_, part, s);
  }
}

BatchTuning Batch::autotune(int form, int L) {
  // Times every supported (shape, chunk) of each grouped product once the
  // pieces are in place; all cohorts are running.
  const Form f = (Form)form;
  BatchTuning best;
  best.g_ms = best.h_ms = 1e30;
  hipEvent_t e0, e1;
  PCAONE_HIP(hipEventCreate(&e0));
  PCAONE_HIP(hipEventCreate(&e1));
  for (int prod = 0; prod < 2; ++prod) {
    const uint64_t* chunks = prod == 0 ? kChunkCandG : kChunkCandH;
    for (const GemmShape& sh : kShapeCand) {
      if (!gemm_shape_supported(f, sh, LP)) continue;
      for (int ci = 0; ci < 3; ++ci) {
        const GemmShape gs = prod == 0 ? sh : layout_.g, hs = prod == 0 ? layout_.h : sh;
        const uint64_t kg = prod == 0 ? chunks[ci] : layout_.kchunk_g, kh = prod == 0 ? layout_.kchunk_h : chunks[ci];
        prepare(form, L, layout_.want_v, gs, hs, kg, kh);
        BatchArgs a{coh_.get(), size(), L, 1, 0, 0.0, form};
        launch_batch_small(a, kSmallInit, 0, stream_);
        launch_batch_active(a, active_.get(), stream_);
        grouped(prod == 0);  // warm-up
        PCAONE_HIP(hipEventRecord(e0, stream_));
        grouped(prod == 0);
        PCAONE_HIP(hipEventRecord(e1, stream_));
        PCAONE_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PCAONE_HIP(hipEventElapsedTime(&ms, e0, e1));
        ++best.candidates;
        if (prod == 0 && ms < best.g_ms) best.g_ms = ms, best.g = sh, best.kchunk_g = chunks[ci];
        if (prod == 1 && ms < best.h_ms) best.h_ms = ms, best.h = sh, best.kchunk_h = chunks[ci];
      }
    }
    // fix the winner of this product before timing the other
    if (prod == 0) prepare(form, L, layout_.want_v, best.g, layout_.h, best.kchunk_g, layout_.kchunk_h);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

void Batch::run(const Param& p, bool want_v, bool use_graph, BatchResult& out) {
  require(finalized_, kErrState, "batch: finalize the cohorts first");
  require(p.svd == 1, kErrArgument, "batch: the batched form runs --svd 1");
  require(!p.emu, kErrArgument, "batch: the batched form has no --emu (cohorts are complete)");
  const int k = (int)p.k, L = (int)(p.k + p.oversamples);
  require(k >= 1 && L <= kBatchMaxL, kErrArgument, "batch: k + oversamples must be at most 32");
  require((int)p.maxp < kBatchMaxEpochs, kErrArgument, "batch: --maxp must be below 64");
  for (uint32_t c = 0; c < size(); ++c)
    require((uint64_t)L <= N_[c] && (uint64_t)L <= M_[c], kErrArgument,
            "batch: k + oversamples exceeds the dimensions of cohort " + std::to_string(c));
  const ProductForms pf = forms_of(p.precision);
  require(pf.g == pf.h, kErrArgument, "batch: mixed operand forms are not supported");
  require(pf.g != Form::I8, kErrArgument, "batch: the int8 form runs on the single-matrix path only");
  const int form = (int)pf.g;
  // configuration: explicit (tuning fields) or autotuned once per (form, L)
  const GemmTuning& t = p.tuning;
  const bool fixed = t.wm_g || t.waves_g || t.wm_h || t.waves_h || t.kchunk_g || t.kchunk_h;
  if (fixed) {
    const GemmShape dg = gemm_default_shape((Form)form, LP);
    tune_ = BatchTuning{};
    tune_.g = GemmShape{t.wm_g ? t.wm_g : dg.wm, t.waves_g ? t.waves_g : dg.waves};
    tune_.h = GemmShape{t.wm_h ? t.wm_h : dg.wm, t.waves_h ? t.waves_h : dg.waves};
    tune_.kchunk_g = t.kchunk_g ? round_up(t.kchunk_g, kKStep) : 4096;
    tune_.kchunk_h = t.kchunk_h ? round_up(t.kchunk_h, kKStep) : 8192;
  } else {
    auto it = tuned_.find({form, L});
    if (it == tuned_.end()) {
      const GemmShape dg = gemm_default_shape((Form)form, LP);
      prepare(form, L, want_v, dg, dg, 4096, 8192);
      it = tuned_.emplace(std::make_tuple(form, L), autotune(form, L)).first;
    }
    tune_ = it->second;
  }
  prepare(form, L, want_v, tune_.g, tune_.h, tune_.kchunk_g, tune_.kchunk_h);
  upload_start(p, L);
  const BatchArgs a{coh_.get(), size(), L, k, (int)p.maxp, p.tol, form};
  launch_batch_small(a, kSmallInit, 0, stream_);  // kernel attributes are set outside any capture
  PCAONE_HIP(hipStreamSynchronize(stream_));
  const auto t0 = std::chrono::steady_clock::now();
  if (use_graph) {
    const auto key = std::make_tuple(form, L, k, (int)p.maxp, want_v, p.tol);
    if (!graph_ || key != graph_key_) {
      release_graph();
      hipGraph_t g = nullptr;
      PCAONE_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      enqueue_epochs(a, (int)p.maxp, want_v);
      PCAONE_HIP(hipStreamEndCapture(stream_, &g));
      const hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        graph_ = nullptr;
        fail(kErrDevice, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      graph_key_ = key;
    }
    PCAONE_HIP(hipGraphLaunch(graph_, stream_));
  } else {
    enqueue_epochs(a, (int)p.maxp, want_v);
  }
  PCAONE_HIP(hipStreamSynchronize(stream_));
  last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // results
  std::vector<BatchSmall> st(size());
  PCAONE_HIP(hipMemcpyAsync(st.data(), small_.get(), st.size() * sizeof(BatchSmall), hipMemcpyDeviceToHost,
                            stream_));
  PCAONE_HIP(hipStreamSynchronize(stream_));
  out.k = (uint32_t)k;
  out.status.assign(size(), 0);
  out.epochs.assign(size(), 0);
  out.nsv.assign(size(), 0);
  out.S.assign((uint64_t)size() * k, 0.0);
  out.trace.assign((uint64_t)size() * kBatchMaxEpochs, 0.0);
  stop_epoch_.assign(size(), 0);
  for (uint32_t c = 0; c < size(); ++c) {
    out.status[c] = st[c].state < 0 ? -st[c].state : (st[c].state == 0 ? 0 : kErrInternal);
    out.epochs[c] = st[c].epochs;
    out.nsv[c] = st[c].nsv;
    stop_epoch_[c] = st[c].stop_epoch;
    for (int i = 0; i < k; ++i) out.S[(uint64_t)c * k + i] = st[c].sigma[i];
    std::memcpy(&out.trace[(uint64_t)c * kBatchMaxEpochs], st[c].trace, sizeof(st[c].trace));
  }
  last_k_ = k;
}

void Batch::get_u(uint32_t c, double* U) const {
  require(c < size() && c < stop_epoch_.size(), kErrState, "batch: no result for this cohort");
  const int buf = (stop_epoch_[c] & 1) ? kBufU1 : kBufU0;
  std::vector<double> h((uint64_t)N_[c] * LP);
  PCAONE_HIP(hipMemcpyAsync(h.data(), host_coh_[c].mat[buf], h.size() * 8, hipMemcpyDeviceToHost, stream_));
  PCAONE_HIP(hipStreamSynchronize(stream_));
  for (int j = 0; j < last_k_; ++j)
    for (uint64_t i = 0; i < N_[c]; ++i) U[(uint64_t)j * N_[c] + i] = h[i * LP + j];
}

void Batch::get_v(uint32_t c, double* V) const {
  require(c < size() && c < stop_epoch_.size(), kErrState, "batch: no result for this cohort");
  require(layout_.want_v && host_coh_[c].mat[kBufV], kErrState, "batch: run with loadings first");
  std::vector<double> h(M_[c] * LP);
  PCAONE_HIP(hipMemcpyAsync(h.data(), host_coh_[c].mat[kBufV], h.size() * 8, hipMemcpyDeviceToHost, stream_));
  PCAONE_HIP(hipStreamSynchronize(stream_));
  for (int j = 0; j < last_k_; ++j)
    for (uint64_t i = 0; i < M_[c]; ++i) V[(uint64_t)j * M_[c] + i] = h[i * LP + j];
}

}  // namespace pcaone
