// extern "C" entry points: argument checks and exception -> status translation.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/pcaone_amd.h"
#include "batch.hpp"
#include "common.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "linalg_host.hpp"
#include "noalloc.hpp"
#include "pgen_io.hpp"
#include "plink_io.hpp"
#include "rng_host.hpp"
#include "rsvd.hpp"
#include "synth_host.hpp"

namespace pcaone {
namespace {
thread_local std::string g_error;
}
void set_global_error(const std::string& msg) { g_error = msg; }
const char* global_error() { return g_error.c_str(); }
}  // namespace pcaone

using namespace pcaone;

struct pcaone_ctx {
  std::unique_ptr<Context> ctx;
  std::unique_ptr<Batch> batch;
  RunInfo last;
  BatchResult last_batch;
  std::string err;
};

namespace {

template <class F>
int guard_global(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_global_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_global_error("host out of memory");
    return kErrInternal;
  } catch (const std::exception& e) {
    set_global_error(e.what());
    return kErrInternal;
  }
}

template <class F>
int guard(pcaone_ctx* c, F&& f) {
  if (!c || !c->ctx) {
    set_global_error("null context");
    return kErrArgument;
  }
  try {
    PCAONE_HIP(hipSetDevice(c->ctx->device));
    f(*c->ctx);
    return kOk;
  } catch (const Error& e) {
    c->err = e.what();
    set_global_error(c->err);
    return e.status;
  } catch (const std::bad_alloc&) {
    c->err = "host out of memory";
    set_global_error(c->err);
    return kErrInternal;
  } catch (const std::exception& e) {
    c->err = e.what();
    set_global_error(c->err);
    return kErrInternal;
  }
}

synth::Spec to_spec(const pcaone_synth_spec* s) {
  require(s != nullptr, kErrArgument, "null synth spec");
  return synth::Spec{s->seed,   s->nsamples, s->npops,     s->nsnps,          s->freq_lo,       s->freq_hi,
                     s->fst_lo, s->fst_hi,   s->miss_rate, s->high_miss_frac, s->high_miss_rate};
}

int to_precision(int p) {
  require(p >= 0 && p < kNumPrecisions, kErrArgument, "unknown precision");
  return p;
}

Param to_param(const pcaone_params* p) {
  require(p != nullptr, kErrArgument, "null params");
  Param q;
  q.svd = p->svd;
  q.k = p->k;
  q.oversamples = p->oversamples;
  q.maxp = p->maxp;
  q.bands = p->bands;
  q.maxiter = p->maxiter;
  q.tol = p->tol;
  q.tolem = p->tolem;
  q.maf = p->maf;
  q.seed = p->seed;
  q.rand = p->rand;
  q.ploidy = p->ploidy;
  q.noshuffle = p->noshuffle != 0;
  q.emu = p->emu != 0;
  q.printv = p->printv != 0;
  q.verbose = p->verbose;
  q.precision = to_precision(p->precision);
  q.tuning.wm_g = p->wm_g;
  q.tuning.wm_h = p->wm_h;
  q.tuning.waves_g = p->waves_g;
  q.tuning.waves_h = p->waves_h;
  q.tuning.kchunk_g = p->kchunk_g;
  q.tuning.kchunk_h = p->kchunk_h;
  q.tuning.miss_gather = p->miss_gather != 0;
  q.memory = p->memory;
  q.buffer = p->buffer;
  q.ld = p->ld != 0;
  q.ld_stats = p->ld_stats;
  q.imaxiter = p->imaxiter;
  q.itol = p->itol;
  q.stream = p->stream != 0;
  q.hardcall = p->hardcall != 0;
  q.threads = p->threads;
  require(q.threads >= 1, kErrArgument, "threads must be >= 1");
  require(q.k >= 1, kErrArgument, "k must be >= 1");
  require(q.ploidy == 1 || q.ploidy == 2, kErrArgument, "ploidy must be 1 or 2");
  q.derive();
  return q;
}

void fill_info(pcaone_run_info* info, const RunInfo& r, const Context& c, int nsv) {
  if (!info) return;
  std::memset(info, 0, sizeof(*info));
  info->epochs = r.epochs;
  info->total_epochs = r.total_epochs;
  info->em_iters = r.em_iters;
  info->n_trace = (int)r.trace.size();
  info->n_sv = nsv;
  info->seconds = r.seconds;
  info->g_ms = c.ktimes.g_ms;
  info->h_ms = c.ktimes.h_ms;
  info->g_launches = c.ktimes.g_launches;
  info->h_launches = c.ktimes.h_launches;
  info->omega_ms = r.omega_ms;
  info->householder_fallbacks = c.householder_fallbacks;
  info->gram_second_passes = c.gram_second_passes;
  info->epoch_rollbacks = c.epoch_rollbacks;
  info->stream_loads = c.stream_loads;
  info->stream_load_ms = c.stream_load_ms;
}

}  // namespace

extern "C" {

const char* pcaone_last_error(const pcaone_ctx* ctx) {
  if (ctx && !ctx->err.empty()) return ctx->err.c_str();
  return global_error();
}

int pcaone_device_count(int* n) {
  return guard_global([&] {
    require(n != nullptr, kErrArgument, "null output");
    int c = 0;
    *n = hipGetDeviceCount(&c) == hipSuccess ? c : 0;
  });
}

int pcaone_open(int device, pcaone_ctx** out) {
  return guard_global([&] {
    require(out != nullptr, kErrArgument, "null output");
    *out = nullptr;
    auto c = std::make_unique<pcaone_ctx>();
    c->ctx = std::make_unique<Context>(device);
    *out = c.release();
  });
}

int pcaone_close(pcaone_ctx* ctx) {
  return guard_global([&] { delete ctx; });
}

void pcaone_params_default(pcaone_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->svd = 2;
  p->k = 10;
  p->oversamples = 10;
  p->maxp = 20;
  p->bands = 64;
  p->maxiter = 100;
  p->tol = 1e-4;
  p->tolem = 1e-5;
  p->seed = 112;
  p->rand = 1;
  p->ploidy = 2;
  p->precision = PCAONE_PREC_FP32;
  p->buffer = 2;
  p->threads = 12;
  p->imaxiter = 1000;
  p->itol = 1e-6;
}

int pcaone_bed_read(pcaone_ctx* c, const char* prefix, double maf, uint32_t* nsamples, uint64_t* nsnps) {
  return guard(c, [&](Context& x) {
    require(prefix != nullptr, kErrArgument, "null prefix");
    Param p;
    p.maf = maf;
    p.svd = 1;
    p.derive();
    read_plink(x, prefix, p);
    if (nsamples) *nsamples = x.N;
    if (nsnps) *nsnps = x.M;
  });
}

int pcaone_bed_dims(const char* prefix, uint32_t* nsamples, uint64_t* nsnps) {
  return guard_global([&] {
    require(prefix != nullptr, kErrArgument, "null prefix");
    const PlinkInfo pi = plink_dims(prefix);
    if (nsamples) *nsamples = (uint32_t)pi.nsamples;
    if (nsnps) *nsnps = pi.nsnps_file;
  });
}

int pcaone_pgen_dims(const char* prefix, uint32_t* nsamples, uint64_t* nsnps) {
  return guard_global([&] {
    require(prefix != nullptr, kErrArgument, "null prefix");
    const PlinkInfo pi = pgen_dims(prefix);
    if (nsamples) *nsamples = (uint32_t)pi.nsamples;
    if (nsnps) *nsnps = pi.nsnps_file;
  });
}

int pcaone_pgen_read(pcaone_ctx* c, const char* prefix, double maf, int hardcall, uint32_t* nsamples,
                     uint64_t* nsnps) {
  return guard(c, [&](Context& x) {
    require(prefix != nullptr, kErrArgument, "null prefix");
    Param p;
    p.maf = maf;
    p.svd = 1;
    p.hardcall = hardcall != 0;
    p.derive();
    read_pgen(x, prefix, p);
    if (nsamples) *nsamples = x.N;
    if (nsnps) *nsnps = x.M;
  });
}

int pcaone_pgen_decode(const char* prefix, int hardcall, uint64_t snp_begin, uint64_t snp_end, uint8_t* rows) {
  return guard_global([&] {
    require(prefix != nullptr && rows != nullptr, kErrArgument, "null argument");
    const uint64_t n = psam_samples(std::string(prefix) + ".psam");
    const PgenSource src(std::string(prefix) + ".pgen", n, hardcall != 0);
    require(snp_begin <= snp_end && snp_end <= src.nsnps(), kErrArgument, "pgen_decode: SNP range out of bounds");
    src.decode_rows(snp_begin, snp_end - snp_begin, nullptr, rows);
  });
}

int pcaone_stream_pgen(pcaone_ctx* c, const char* prefix, int hardcall, uint64_t block_snps, uint32_t* nsamples,
                       uint64_t* nsnps) {
  return guard(c, [&](Context& x) {
    require(prefix != nullptr, kErrArgument, "null prefix");
    Param p;
    p.hardcall = hardcall != 0;
    stream_pgen(x, prefix, p, block_snps);
    if (nsamples) *nsamples = x.N;
    if (nsnps) *nsnps = x.M;
  });
}

int pcaone_bed_reserve(pcaone_ctx* c, uint32_t nsamples, uint64_t nsnps) {
  return guard(c, [&](Context& x) { x.reserve(nsamples, nsnps); });
}

int pcaone_bed_upload(pcaone_ctx* c, const uint8_t* rows, uint64_t snp_begin, uint64_t n) {
  return guard(c, [&](Context& x) {
    require(rows != nullptr || n == 0, kErrArgument, "null rows");
    x.upload_rows(rows, snp_begin, n);
  });
}

int pcaone_bed_finalize(pcaone_ctx* c) {
  return guard(c, [&](Context& x) { x.finalize(); });
}

int pcaone_bed_generate(pcaone_ctx* c, const pcaone_synth_spec* spec) {
  return guard(c, [&](Context& x) { x.generate(to_spec(spec)); });
}

int pcaone_stream_bed(pcaone_ctx* c, const char* prefix, uint64_t block_snps, uint32_t* nsamples, uint64_t* nsnps) {
  return guard(c, [&](Context& x) {
    require(prefix != nullptr, kErrArgument, "null prefix");
    const PlinkInfo pi = plink_dims(prefix);
    x.stream_bed(std::string(prefix) + ".bed", (uint32_t)pi.nsamples, pi.nsnps_file, block_snps);
    if (nsamples) *nsamples = x.N;
    if (nsnps) *nsnps = x.M;
  });
}

int pcaone_stream_host(pcaone_ctx* c, const uint8_t* rows, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps) {
  return guard(c, [&](Context& x) { x.stream_host(rows, nsamples, nsnps, block_snps); });
}

int pcaone_bed_download(pcaone_ctx* c, uint8_t* rows) {
  return guard(c, [&](Context& x) {
    require(rows != nullptr, kErrArgument, "null output");
    x.download_rows(rows);
  });
}

int pcaone_dims(const pcaone_ctx* c, uint32_t* nsamples, uint64_t* nsnps, uint64_t* nmissing) {
  return guard(const_cast<pcaone_ctx*>(c), [&](Context& x) {
    if (nsamples) *nsamples = x.N;
    if (nsnps) *nsnps = x.M;
    if (nmissing) *nmissing = x.nmiss;
  });
}

int pcaone_allele_freq(const pcaone_ctx* c, double* F) {
  return guard(const_cast<pcaone_ctx*>(c), [&](Context& x) {
    require(F != nullptr, kErrArgument, "null output");
    require(!x.F.empty(), kErrState, "no data");
    std::memcpy(F, x.F.data(), x.F.size() * sizeof(double));
  });
}

int pcaone_standardize(pcaone_ctx* c, int ploidy) {
  return guard(c, [&](Context& x) { x.standardize(ploidy); });
}

int pcaone_permute_snps(pcaone_ctx* c, const int32_t* perm) {
  return guard(c, [&](Context& x) {
    require(perm != nullptr, kErrArgument, "null permutation");
    x.permute_columns(std::vector<int>(perm, perm + x.M));
  });
}

int pcaone_set_impute(pcaone_ctx* c, const double* U, const double* S, const double* V, uint32_t k) {
  return guard(c, [&](Context& x) {
    require(U && S && V, kErrArgument, "null U/S/V");
    x.set_impute(U, S, V, (int)k);
  });
}

int pcaone_configure(pcaone_ctx* c, int precision, uint32_t L) {
  return guard(c, [&](Context& x) { x.configure(to_precision(precision), (int)L); });
}

int pcaone_set_omega(pcaone_ctx* c, const double* omega) {
  return guard(c, [&](Context& x) {
    require(omega != nullptr, kErrArgument, "null omega");
    x.set_omega(omega);
  });
}

int pcaone_sketch(pcaone_ctx* c, uint64_t b, uint64_t e, int h_mode) {
  return guard(c, [&](Context& x) {
    require(h_mode >= 0 && h_mode <= 2, kErrArgument, "h_mode must be 0, 1 or 2");
    x.sketch(b, e, h_mode);
  });
}

int pcaone_update_omega(pcaone_ctx* c, int source) {
  return guard(c, [&](Context& x) {
    require(source == 0 || source == 1, kErrArgument, "source must be 0 or 1");
    x.omega_from(source);
  });
}

int pcaone_thin_q_omega(pcaone_ctx* c, int source) {
  return guard(c, [&](Context& x) {
    require(source == 0 || source == 1, kErrArgument, "source must be 0 or 1");
    x.omega_from(source, /*flip=*/false);
  });
}

int pcaone_zero_h(pcaone_ctx* c, int which) {
  return guard(c, [&](Context& x) {
    require(which == 1 || which == 2, kErrArgument, "which must be 1 or 2");
    x.zero_h(which);
  });
}

int pcaone_finish_epoch(pcaone_ctx* c, uint32_t k, int have_prev, double* sigma, double* diff) {
  return guard(c, [&](Context& x) {
    const EpochResult r = x.finish_epoch((int)k, have_prev != 0);
    if (sigma) std::memcpy(sigma, r.sigma.data(), r.sigma.size() * sizeof(double));
    if (diff) *diff = r.diff;
  });
}

int pcaone_save_prev(pcaone_ctx* c) {
  return guard(c, [&](Context& x) { x.save_prev(); });
}

int pcaone_get_u(pcaone_ctx* c, uint32_t k, double* U) {
  return guard(c, [&](Context& x) {
    require(U != nullptr && k >= 1 && (int)k <= x.L, kErrArgument, "bad U request");
    x.download_u((int)k, U);
  });
}

int pcaone_get_g(pcaone_ctx* c, uint64_t snp_begin, uint64_t snp_end, double* G) {
  return guard(c, [&](Context& x) {
    require(G != nullptr, kErrArgument, "null G");
    x.download_g(snp_begin, snp_end, G);
  });
}

int pcaone_get_h(pcaone_ctx* c, int which, double* H) {
  return guard(c, [&](Context& x) {
    require(H != nullptr, kErrArgument, "null H");
    x.download_h(which, H);
  });
}

int pcaone_get_v(pcaone_ctx* c, uint32_t k, double* V) {
  return guard(c, [&](Context& x) {
    require(V != nullptr && k >= 1 && (int)k <= x.L, kErrArgument, "bad V request");
    x.compute_v((int)k);
    x.download_v((int)k, V);
  });
}

int pcaone_run_halko(pcaone_ctx* c, const pcaone_params* p, double* U, double* S, double* E, double* V,
                     pcaone_run_info* info) {
  return guard(c, [&](Context& x) {
    Param q = to_param(p);
    if (V) q.printv = true;
    require(x.M > 0 && !x.F.empty(), kErrState, "no genotype data (pcaone_bed_read / upload first)");
    c->last = RunInfo{};
    x.ktimes = Context::KernelTimes{};
    const PcaResult r = run_pca(x, q, &c->last, U);
    const uint64_t M = x.M;
    if (S) std::memcpy(S, r.S.data(), r.S.size() * sizeof(double));
    if (E) std::memcpy(E, r.E.data(), r.E.size() * sizeof(double));
    if (V && !r.V.empty()) std::memcpy(V, r.V.data(), M * (uint64_t)r.k * sizeof(double));
    fill_info(info, c->last, x, (int)r.S.size());
  });
}

// ---------------------------------------------------------------- batched form
int pcaone_batch_reserve(pcaone_ctx* c, uint32_t ncohorts, const uint32_t* nsamples, const uint64_t* nsnps) {
  return guard(c, [&](Context& x) {
    require(ncohorts > 0 && nsamples && nsnps, kErrArgument, "batch: empty cohort list");
    if (!c->batch) c->batch = std::make_unique<Batch>(x);
    c->batch->reserve(std::vector<uint32_t>(nsamples, nsamples + ncohorts),
                      std::vector<uint64_t>(nsnps, nsnps + ncohorts));
  });
}

int pcaone_batch_upload(pcaone_ctx* c, uint32_t cohort, const uint8_t* rows, uint64_t snp_begin, uint64_t nsnps) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr, kErrState, "batch: reserve first");
    require(rows != nullptr || nsnps == 0, kErrArgument, "null rows");
    c->batch->upload_rows(cohort, rows, snp_begin, nsnps);
  });
}

int pcaone_batch_generate(pcaone_ctx* c, uint32_t ncohorts, const pcaone_synth_spec* specs) {
  return guard(c, [&](Context& x) {
    require(ncohorts > 0 && specs, kErrArgument, "batch: empty spec list");
    std::vector<synth::Spec> v;
    std::vector<uint32_t> n;
    std::vector<uint64_t> m;
    for (uint32_t i = 0; i < ncohorts; ++i) {
      v.push_back(to_spec(&specs[i]));
      n.push_back(specs[i].nsamples);
      m.push_back(specs[i].nsnps);
    }
    if (!c->batch) c->batch = std::make_unique<Batch>(x);
    c->batch->reserve(n, m);
    c->batch->generate(v);
    c->batch->finalize();
  });
}

int pcaone_batch_finalize(pcaone_ctx* c) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr, kErrState, "batch: reserve first");
    c->batch->finalize();
  });
}

int pcaone_batch_allele_freq(pcaone_ctx* c, uint32_t cohort, double* F) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr && cohort < c->batch->size() && F, kErrArgument, "batch: bad cohort");
    const auto& f = c->batch->allele_freq(cohort);
    require(!f.empty(), kErrState, "batch: finalize first");
    std::memcpy(F, f.data(), f.size() * sizeof(double));
  });
}

int pcaone_batch_run(pcaone_ctx* c, const pcaone_params* p, int use_graph, int want_v, double* S, double* E,
                     int* epochs, int* status, pcaone_batch_info* info) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr, kErrState, "batch: no cohorts");
    Param q = to_param(p);
    Batch& b = *c->batch;
    b.run(q, want_v != 0, use_graph != 0, c->last_batch);
    const BatchResult& r = c->last_batch;
    const uint32_t n = b.size(), k = r.k;
    for (uint32_t i = 0; i < n; ++i) {
      for (uint32_t j = 0; j < k; ++j) {
        const double sv = (int)j < r.nsv[i] ? r.S[(uint64_t)i * k + j] : 0.0;
        if (S) S[(uint64_t)i * k + j] = sv;
        if (E) E[(uint64_t)i * k + j] = sv * sv / (double)b.nsnps(i);  // src/Halko.cpp:339
      }
      if (epochs) epochs[i] = r.epochs[i];
      if (status) status[i] = r.status[i];
    }
    if (info) {
      std::memset(info, 0, sizeof(*info));
      const BatchTuning& t = b.tuning();
      info->seconds = b.last_seconds;
      info->wm_g = t.g.wm, info->waves_g = t.g.waves, info->wm_h = t.h.wm, info->waves_h = t.h.waves;
      info->kchunk_g = t.kchunk_g, info->kchunk_h = t.kchunk_h;
      info->tuned_g_ms = t.g_ms, info->tuned_h_ms = t.h_ms;
      info->tuned_candidates = t.candidates;
      int failed = 0;
      for (uint32_t i = 0; i < n; ++i) failed += r.status[i] != 0;
      info->failed = failed;
    }
  });
}

int pcaone_batch_get_u(pcaone_ctx* c, uint32_t cohort, double* U) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr && U, kErrArgument, "batch: bad request");
    c->batch->get_u(cohort, U);
  });
}

int pcaone_batch_get_v(pcaone_ctx* c, uint32_t cohort, double* V) {
  return guard(c, [&](Context&) {
    require(c->batch != nullptr && V, kErrArgument, "batch: bad request");
    c->batch->get_v(cohort, V);
  });
}

int pcaone_batch_trace(const pcaone_ctx* c, uint32_t cohort, double* out, int cap, int* n) {
  return guard(const_cast<pcaone_ctx*>(c), [&](Context&) {
    const BatchResult& r = c->last_batch;
    require(cohort < r.epochs.size(), kErrArgument, "batch: bad cohort");
    const int cnt = r.epochs[cohort] > 1 ? r.epochs[cohort] - 1 : 0;  // checked epochs 1 ..
    if (n) *n = cnt;
    if (out)
      for (int i = 0; i < cap && i < cnt; ++i) out[i] = r.trace[(uint64_t)cohort * kBatchMaxEpochs + 1 + i];
  });
}

int pcaone_run_trace(const pcaone_ctx* c, int em, double* out, int cap, int* n) {
  return guard(const_cast<pcaone_ctx*>(c), [&](Context&) {
    const auto& v = em ? c->last.em_trace : c->last.trace;
    if (n) *n = (int)v.size();
    if (out)
      for (int i = 0; i < cap && i < (int)v.size(); ++i) out[i] = v[i];
  });
}

int pcaone_run_bfile(pcaone_ctx* c, const char* bfile, const char* out, const pcaone_params* p,
                     pcaone_run_info* info) {
  return guard(c, [&](Context& x) {
    require(bfile && out, kErrArgument, "null path");
    const Param q = to_param(p);
    PlinkInfo pi;
    if (q.stream) {
      // the -m plan's blocks are read from the mapped .bed on every pass
      require(q.out_of_core, kErrArgument, "stream needs memory > 0 (the -m plan sets the blocks)");
      pi = plink_dims(bfile);
      const BlockPlan bp = plan_blocks(q, pi.nsamples, pi.nsnps_file);
      x.stream_bed(std::string(bfile) + ".bed", (uint32_t)pi.nsamples, pi.nsnps_file, bp.blocksize);
    } else {
      pi = read_plink(x, bfile, q);
    }
    c->last = RunInfo{};
    x.ktimes = Context::KernelTimes{};
    const PcaResult r = run_pca(x, q, &c->last);
    if (q.ld && q.svd != 3) write_residuals(x, r, q, out);  // src/Halko.cpp:338, src/Arnoldi.cpp:117
    write_outputs(x, r, q, bfile, out, pi);
    fill_info(info, c->last, x, (int)r.S.size());
  });
}

int pcaone_run_pgen(pcaone_ctx* c, const char* pfile, const char* out, const pcaone_params* p,
                    pcaone_run_info* info) {
  return guard(c, [&](Context& x) {
    require(pfile && out, kErrArgument, "null path");
    const Param q = to_param(p);
    PlinkInfo pi;
    if (q.stream) {
      // the -m plan's blocks are decoded from the mapped .pgen on every pass
      require(q.out_of_core, kErrArgument, "stream needs memory > 0 (the -m plan sets the blocks)");
      pi = pgen_dims(pfile);
      const BlockPlan bp = plan_blocks(q, pi.nsamples, pi.nsnps_file);
      pi = stream_pgen(x, pfile, q, bp.blocksize);
    } else {
      pi = read_pgen(x, pfile, q);
    }
    c->last = RunInfo{};
    x.ktimes = Context::KernelTimes{};
    const PcaResult r = run_pca(x, q, &c->last);
    if (q.ld && q.svd != 3) write_residuals(x, r, q, out);  // src/Halko.cpp:338, src/Arnoldi.cpp:117
    write_outputs(x, r, q, pfile, out, pi);
    fill_info(info, c->last, x, (int)r.S.size());
  });
}

int pcaone_residuals(pcaone_ctx* c, uint32_t k, const double* U, const double* S, const double* V, int as_read,
                     uint64_t snp_begin, uint64_t snp_end, float* out) {
  return guard(c, [&](Context& x) {
    require(out != nullptr, kErrArgument, "null output");
    x.residuals_prepare((int)k, U, S, V, as_read != 0);
    x.residuals_rows(snp_begin, snp_end, out);
  });
}

int pcaone_time_kernels(pcaone_ctx* c, int on) {
  return guard(c, [&](Context& x) { x.time_kernels = on != 0; });
}

int pcaone_synth_packed(const pcaone_synth_spec* spec, uint8_t* packed_out, uint32_t* attempts_out) {
  return guard_global([&] {
    require(packed_out != nullptr, kErrArgument, "null output buffer");
    synth::generate(to_spec(spec), packed_out, attempts_out);
  });
}

int pcaone_synth_write_plink(const pcaone_synth_spec* spec, const char* prefix) {
  return guard_global([&] {
    require(prefix != nullptr, kErrArgument, "null prefix");
    synth::write_plink(to_spec(spec), prefix);
  });
}

int pcaone_packed_gemm(pcaone_ctx* c, int precision, const uint8_t* rows, uint64_t nrows, uint32_t ncodes,
                       const double* W, uint32_t L, double* Y) {
  return guard(c, [&](Context& x) {
    require(rows && W && Y && nrows > 0 && ncodes > 0 && L >= 1 && L <= 128, kErrArgument, "bad gemm arguments");
    require(precision >= 0 && precision < kNumPrecisions && forms_of(precision).g == forms_of(precision).h,
            kErrArgument, "packed_gemm: precision must be a single-form precision (0..3, 5, 7 or 8)");
    const Form f = forms_of(precision).g;
    require(f != Form::I8 || round_up(ncodes, kKStep) <= kI8ChunkCap, kErrArgument,
            "packed_gemm: the int8 form takes at most 65536 codes per call");
    const int lp = (int)round_up(L, 16);
    const GemmShape shape = gemm_default_shape(f, lp);
    const uint64_t tile = gemm_row_tile(shape);
    const uint64_t rows_pad = round_up(nrows, tile);
    const uint64_t K = round_up(ncodes, kKStep), pitch = K / 16, bps = bytes_per_snp(ncodes);
    DevBuf<uint8_t> raw(nrows * bps);
    DevBuf<uint32_t> X(rows_pad * pitch);
    launch_fill_rows(X.get(), pitch, 0, rows_pad, x.stream);
    PCAONE_HIP(hipMemcpyAsync(raw.get(), rows, nrows * bps, hipMemcpyHostToDevice, x.stream));
    launch_interleave(raw.get(), nrows, bps, ncodes, X.get(), pitch, x.stream);
    std::vector<double> w(K * lp, 0.0);
    for (uint64_t k = 0; k < ncodes; ++k)
      for (uint32_t l = 0; l < L; ++l) w[k * lp + l] = W[(uint64_t)l * ncodes + k];
    DevBuf<double> dw(K * lp);
    PCAONE_HIP(hipMemcpyAsync(dw.get(), w.data(), w.size() * 8, hipMemcpyHostToDevice, x.stream));
    const int P = form_pieces(f);
    DevBuf<unsigned char> wt((uint64_t)P * lp * K * form_wbytes(f));
    DevBuf<double> i8s(2 * 128);
    DevBuf<unsigned long long> i8max(128);
    if (form_colscaled(f))
      launch_col_scales(f, dw.get(), ncodes, lp, nullptr, i8max.get(), i8s.get(), i8s.get() + 128, x.stream);
    launch_make_pieces(f, dw.get(), ncodes, K, lp, nullptr, nullptr, wt.get(), K, (uint64_t)lp * K, nullptr,
                       nullptr, x.stream, i8s.get());
    DevBuf<unsigned char> out(rows_pad * lp * form_accbytes(f));
    GemmArgs a{};
    a.X = X.get();
    a.xpitch = pitch;
    a.rows = rows_pad;
    a.Wt = wt.get();
    a.wpitch = K;
    a.wpiece = (uint64_t)lp * K;
    a.Y = out.get();
    a.ychunk = rows_pad * lp;
    a.k_begin = 0;
    a.k_end = K;
    a.kchunk = K;
    a.lp = lp;
    launch_packed_gemm(f, shape, a, x.stream);
    std::vector<unsigned char> h(out.bytes());
    std::vector<double> mul(128, 1.0);
    PCAONE_HIP(hipMemcpyAsync(h.data(), out.get(), h.size(), hipMemcpyDeviceToHost, x.stream));
    if (form_colscaled(f))
      PCAONE_HIP(hipMemcpyAsync(mul.data(), i8s.get() + 128, 128 * 8, hipMemcpyDeviceToHost, x.stream));
    PCAONE_HIP(hipStreamSynchronize(x.stream));
    for (uint64_t r = 0; r < nrows; ++r)
      for (uint32_t l = 0; l < L; ++l) {
        const uint64_t idx = r * lp + l;
        double v;
        if (f == Form::FP64)
          v = reinterpret_cast<const double*>(h.data())[idx];
        else if (f == Form::I8)  // int32 sum of (v/2) q, times 2 s_l
          v = (double)reinterpret_cast<const int32_t*>(h.data())[idx] * mul[l];
        else if (f == Form::F8)  // FP32 sum of v (w / s_l), times s_l
          v = (double)reinterpret_cast<const float*>(h.data())[idx] * mul[l];
        else
          v = reinterpret_cast<const float*>(h.data())[idx];
        Y[(uint64_t)l * nrows + r] = v;
      }
  });
}

int pcaone_packed_gemm_noalloc(pcaone_ctx* c, int precision, const uint8_t* rows, uint64_t nrows, uint32_t ncodes,
                               const double* W, uint32_t L, double* Y) {
  return guard(c, [&](Context& x) {
    require(rows && W && Y && nrows > 0 && ncodes > 0 && L >= 1 && L <= 128, kErrArgument, "bad gemm arguments");
    require(precision >= 0 && precision < kNumPrecisions && forms_of(precision).g == forms_of(precision).h,
            kErrArgument, "packed_gemm_noalloc: precision must be a single-form precision (0..3, 5, 7 or 8)");
    const Form f = forms_of(precision).g;
    require(f != Form::I8 || round_up(ncodes, kKStep) <= kI8ChunkCap, kErrArgument,
            "packed_gemm_noalloc: the int8 form takes at most 65536 codes per call");
    packed_gemm_noalloc(x, f, rows, nrows, ncodes, W, L, Y);
  });
}

int pcaone_copy_probe(pcaone_ctx* c, uint64_t bytes) {
  return guard(c, [&](Context& x) {
    const uint64_t n = bytes / 16 * 16;
    DevBuf<unsigned char> a(n), b(n);
    PCAONE_HIP(hipMemsetAsync(a.get(), 1, n, x.stream));
    launch_copy_probe(a.get(), b.get(), n, x.stream);
    PCAONE_HIP(hipStreamSynchronize(x.stream));
  });
}

int pcaone_block_plan(const pcaone_params* p, uint32_t nsamples, uint64_t nsnps, uint64_t* blocksize,
                      uint64_t* nblocks, uint64_t* band_factor, uint64_t* start, uint64_t cap) {
  return guard_global([&] {
    const Param q = to_param(p);
    const BlockPlan b = plan_blocks(q, nsamples, nsnps);
    if (blocksize) *blocksize = b.blocksize;
    if (nblocks) *nblocks = b.nblocks;
    if (band_factor) *band_factor = b.band_factor;
    if (start)
      for (uint64_t i = 0; i < b.nblocks && i < cap; ++i) start[i] = b.start[i];
  });
}

int pcaone_permute_plink_order(uint64_t nsamples, uint64_t nsnps, uint32_t buffer_gb, uint32_t bands,
                               int32_t* order) {
  return guard_global([&] {
    require(order != nullptr, kErrArgument, "null output");
    const std::vector<int> o = permute_plink_order(nsamples, nsnps, buffer_gb, bands);
    std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
  });
}

int pcaone_pgen_perm_order(uint64_t nsnps, uint32_t nbatches, uint64_t blocksize, uint32_t nthreads, int seed,
                           int32_t* order) {
  return guard_global([&] {
    require(order != nullptr || nsnps == 0, kErrArgument, "null output");
    const std::vector<int> o = pgen_perm_order(nsnps, nbatches, blocksize, nthreads, seed);
    if (!o.empty()) std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
  });
}

int pcaone_host_omega(int seed, int gaussian, uint64_t nrows, uint64_t ncols, double* out) {
  return guard_global([&] {
    require(out != nullptr, kErrArgument, "null output");
    rng::omega(seed, gaussian != 0, nrows, ncols, out);
  });
}

int pcaone_host_sym_eig(int n, const double* A, double* w, double* V) {
  return guard_global([&] {
    require(n >= 1 && A && w && V, kErrArgument, "bad arguments");
    la::Mat a(n);
    std::memcpy(a.a.data(), A, (size_t)n * n * sizeof(double));
    std::vector<double> ww;
    la::Mat vv;
    la::sym_eig_desc(a, ww, vv);
    std::memcpy(w, ww.data(), n * sizeof(double));
    std::memcpy(V, vv.a.data(), (size_t)n * n * sizeof(double));
  });
}

int pcaone_host_householder_signs(int n, const double* Y, double* s) {
  return guard_global([&] {
    require(n >= 1 && Y && s, kErrArgument, "bad arguments");
    la::Mat y(n);
    std::memcpy(y.a.data(), Y, (size_t)n * n * sizeof(double));
    const auto v = la::householder_signs(y);
    std::memcpy(s, v.data(), n * sizeof(double));
  });
}

int pcaone_host_chol_upper(int n, const double* A, double* R) {
  return guard_global([&] {
    require(n >= 1 && A && R, kErrArgument, "bad arguments");
    la::Mat a(n);
    std::memcpy(a.a.data(), A, (size_t)n * n * sizeof(double));
    const la::Mat r = la::chol_upper(a);
    std::memcpy(R, r.a.data(), (size_t)n * n * sizeof(double));
  });
}

int pcaone_host_householder_q(uint64_t nrows, int ncols, double* X) {
  return guard_global([&] {
    require(ncols >= 1 && nrows >= (uint64_t)ncols && X, kErrArgument, "bad arguments");
    la::householder_q(X, nrows, ncols, ncols);
  });
}

}  // extern "C"
