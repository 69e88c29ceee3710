// Randomized-SVD drivers on the device context (mirror of src/Halko.cpp).
#include "rsvd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <thread>

#include "common.hpp"
#include "exact.hpp"
#include "pgen_io.hpp"
#include "rng_host.hpp"

namespace pcaone {

void Param::derive() {
  ncv = 20 > (2 * k + 1) ? 20 : (2 * k + 1);        // src/Cmd.cpp:215
  oversamples = oversamples > k ? oversamples : k;  // src/Cmd.cpp:216
  filterSNP = maf > 0;                              // src/Cmd.cpp:220
  if (filterSNP) require(maf > 0 && maf < 0.5, kErrArgument, "--maf has to be between (0, 0.5)");
  missme = emu;                                     // src/Cmd.cpp:228-232
  if (!missme) maxiter = 0;
  require(bands >= 4 && bands % 2 == 0, kErrArgument,
          "the -w/--batches must be a power of 2 and the minimun is 4.");  // src/Cmd.cpp:235-236
  perm = (svd == 2 && !noshuffle);                 // src/Cmd.cpp:238
  require(svd >= 0 && svd <= 3, kErrArgument, "--svd must be 0 (IRAM), 1, 2 or 3 (exact PCA)");
  out_of_core = memory > 0 && svd != 3;            // src/Cmd.cpp:218
  require(!(filterSNP && out_of_core), kErrArgument,
          "does not support --maf filters for out-of-core mode yet! ");  // src/Cmd.cpp:221-224
}

BlockPlan plan_blocks(const Param& p, uint64_t nsamples, uint64_t nsnps) {
  // src/Data.cpp:42-84.  The reference holds the counts in 32-bit unsigned
  // integers (src/Data.hpp:43-45) and evaluates its RAM model
  // (3 n l + 2 m l + 5 m doubles) in that type.
  require(nsamples > 0 && nsamples < (1ull << 32) && nsnps > 0 && nsnps < (1ull << 32), kErrArgument,
          "out-of-core: sample and SNP counts must fit 32 bits");
  const uint32_t n = (uint32_t)nsamples, m = (uint32_t)nsnps;
  const uint32_t l = p.k + p.oversamples;
  double ram = (double)(3u * n * l + 2u * m * l + 5u * m) / 134217728;
  if (p.memory > 1.1 * ram) ram = 0;  // otherwise the reference warns and adds the minimum
  const double bs = std::ceil((double)((ram + p.memory) * 134217728 - 3u * n * l - 2u * m * l - 5u * m) / n);
  require(bs >= 1, kErrArgument, "out-of-core: -m/--memory leaves no room for a block of SNPs");
  require(bs < (double)m, kErrArgument, "only one block exists. please remove -m option");  // src/Data.cpp:63
  BlockPlan b;
  uint32_t blocksize = (uint32_t)bs;
  uint32_t nblocks = (uint32_t)std::ceil((double)m / blocksize);
  uint32_t factor = 1;
  if (p.svd == 2) {  // src/Data.cpp:64-76: blocks become the winSVD windows
    if (nblocks < p.bands) {
      blocksize = (uint32_t)std::ceil((double)m / p.bands);
    } else {
      factor = (uint32_t)std::ceil((double)nblocks / p.bands);
      blocksize = (uint32_t)std::ceil((double)m / (p.bands * factor));
    }
    nblocks = (uint32_t)std::ceil((double)m / blocksize);
  }
  b.blocksize = blocksize;
  b.nblocks = nblocks;
  b.band_factor = factor;
  b.start.resize(nblocks);
  b.stop.resize(nblocks);
  for (uint32_t i = 0; i < nblocks; ++i) {  // src/Data.cpp:79-83
    b.start[i] = (uint64_t)i * blocksize;
    b.stop[i] = std::min<uint64_t>(b.start[i] + blocksize - 1, (uint64_t)m - 1);
  }
  return b;
}

std::vector<int> permute_plink_order(uint64_t nsamples, uint64_t nsnps, uint32_t gb, uint32_t nbands) {
  // src/FilePlink.cpp:303-408: the file is read in buffers of a multiple of
  // nbands SNPs; SNP j nbands + b of buffer i goes to band b's run of the
  // output, after the runs of the earlier buffers.  Only the index map is
  // needed here (the device gathers its rows once).
  require(nsnps > 0 && nsnps < (1ull << 31) && nbands > 0, kErrArgument, "permute_plink: bad sizes");
  const uint32_t M = (uint32_t)nsnps;
  const uint32_t bytes_per_snp = (uint32_t)((nsamples + 3) >> 2);
  uint32_t buf_snps = (uint32_t)std::floor((double)1073741824 * gb / bytes_per_snp);
  if (buf_snps > M) buf_snps = M;
  uint32_t bufsize = (uint32_t)std::floor((double)buf_snps / nbands);
  require(bufsize >= 1, kErrArgument, "permute_plink: the --buffer holds fewer SNPs than -w bands");
  buf_snps = bufsize * nbands;
  const uint32_t nbuf = (M + buf_snps - 1) / buf_snps;
  uint32_t modr2 = M % buf_snps;
  std::vector<uint64_t> bandidx(nbands);
  const uint32_t modr = M % nbands, bandsize = (M + nbands - 1) / nbands;
  for (uint32_t i = 0; i < nbands; ++i)
    bandidx[i] = (modr == 0 || i < modr) ? (uint64_t)i * bandsize
                                         : (uint64_t)modr * bandsize + (uint64_t)(bandsize - 1) * (i - modr);
  std::vector<int> order(M, -1);
  std::vector<char> seen(M, 0);
  const uint64_t bufidx = bufsize;
  auto place = [&](uint64_t ia, uint64_t ib) {
    require(ia < M && ib < M && order[ib] < 0 && !seen[ia], kErrInternal, "permute_plink: not a permutation");
    order[ib] = (int)ia;
    seen[ia] = 1;
  };
  for (uint64_t i = 0; i < nbuf; ++i) {
    if (i == nbuf - 1 && modr2 != 0) {  // the last, shorter buffer
      const uint64_t last = M - (uint64_t)(nbuf - 1) * buf_snps;
      bufsize = (uint32_t)((last + nbands - 1) / nbands);
      modr2 = (uint32_t)(last % nbands);
    }
    for (uint64_t b = 0; b < nbands; ++b) {
      uint64_t j = 0;
      for (; j + 1 < bufsize; ++j) place(i * buf_snps + j * nbands + b, i * bufidx + bandidx[b] + j);
      if (i != nbuf - 1 || b < modr2 || modr2 == 0) place(i * buf_snps + j * nbands + b, i * bufidx + bandidx[b] + j);
    }
  }
  for (uint32_t i = 0; i < M; ++i) require(order[i] >= 0, kErrInternal, "permute_plink: not a permutation");
  return order;
}

RsvdOpData::RsvdOpData(Context& ctx, const Param& p, int k, int os) : ctx_(ctx), params_(p), nk_(k), os_(os) {
  ctx_.configure(p.precision, k + os, p.tuning);
}

// src/Halko.cpp:15-23 with src/RSVD.hpp:25-59: libstdc++ default_random_engine
// seeded with --seed; one draw per element in column-major order, drawn
// again at epoch 0 of every computeUSV as the reference does.  The draw runs
// in parallel straight into the context's pinned staging memory.
void RsvdOpData::initOmg() {
  const auto t0 = std::chrono::steady_clock::now();
  rng::omega(params_.seed, params_.rand != 0, cols(), (uint64_t)size(), ctx_.omega_staging());
  if (info) info->omega_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ctx_.set_omega_staged();
}

void RsvdOpData::fit_with_pi() {
  // src/Data.cpp:318-348: missing entries <- clamp(U S V^T, -F, 1-F).  The
  // EMU branch only runs with --emu; the device recomputes the imputed values
  // inside every product from (U S, V).
  if (!params_.emu) return;
  require(!S.empty(), kErrState, "fit_with_pi before the first decomposition");
  if (device_uv) {
    ctx_.set_impute_device(S.data(), (int)S.size());
    return;
  }
  require(!U.empty() && !V.empty(), kErrState, "fit_with_pi before the first decomposition");
  ctx_.set_impute(U.data(), S.data(), V.data(), (int)S.size());
}

void RsvdOpData::computeUSV(int p, double tol) {
  // src/Halko.cpp:46-97.  Where the operator allows it, epoch pi + 1's device
  // work is enqueued between the two halves of epoch pi's host steps, so the
  // L x L work on the host overlaps the next products; the results are the
  // same, and a stop at epoch pi puts epoch pi's G~ and H back.
  const int nk = ranks();
  bool ahead = false;  // epoch pi's products were enqueued during epoch pi - 1
  for (int pi = 0; pi <= p; ++pi) {
    if (!ahead) computeGandH(pi);
    ctx_.finish_begin();
    ahead = pi < p && enqueue_next(pi + 1);
    const EpochResult r = ctx_.finish_end(nk, pi > 0);
    if (info) info->epochs = pi + 1, info->total_epochs += 1;
    if (pi > 0) {
      const double diff = r.diff;
      if (info) info->trace.push_back(diff);
      if (params_.verbose && !params_.missme) {
        std::printf("running of epoch = %d, diff = %f\n", pi, diff);
        std::fflush(stdout);
      }
      if (diff < tol || pi == p) {
        if (params_.svd == 2 && std::pow(2, pi) < params_.bands) {
          p = (int)std::log2(params_.bands);
        } else {
          if (ahead) ctx_.unhold();  // the enqueued next epoch is not used
          ahead = false;
          S.assign(r.sigma.begin(), r.sigma.begin() + nk);
          ctx_.compute_v(nk);
          if (need_v_host && !device_uv) {
            V.resize(rows() * (uint64_t)nk);
            ctx_.download_v(nk, V.data());
          }
          break;
        }
      } else {
        ctx_.save_prev();
      }
    } else {
      ctx_.save_prev();
    }
  }
  if (device_uv) return;  // U stays in the device's ucur
  if (u_out && !params_.missme) {
    ctx_.download_u(nk, u_out);  // straight into the caller's N x k buffer
  } else {
    U.resize(cols() * (uint64_t)nk);
    ctx_.download_u(nk, U.data());
  }
}

void NormalRsvdOpData::computeGandH(int pi) {
  // src/Halko.cpp:99-127 (in-core) and :129-152 (out-of-core).  The block
  // reads re-apply (update, standardize) on every pass from U, S, V, which
  // stay fixed during computeUSV, so the device state set at pi == 0 is the
  // matrix every pass reads.  H = sum over blocks of A_b^T G_b is one product
  // over all blocks.  The out-of-core branch has no flipOmg.
  if (pi == 0) initOmg();
  if (pi == 0) {
    if (update) fit_with_pi();
    if (standardize && params_.scale == -9) ctx_.standardize(params_.ploidy);
  }
  if (pi > 0) ctx_.omega_from(0, /*flip=*/plan == nullptr);
  products();
}

void NormalRsvdOpData::products() {
  // G = A Omega and H = A^T G over all SNPs: one product on resident
  // genotypes; on streamed ones block by block, H summed over the blocks
  // (src/Halko.cpp:137-152)
  if (!ctx_.streamed()) {
    ctx_.sketch(0, rows(), 0);
    return;
  }
  ctx_.zero_h(0);
  for (uint64_t b = 0; b < plan->nblocks; ++b) ctx_.sketch(plan->start[b], plan->stop[b] + 1, 3);
}

bool NormalRsvdOpData::enqueue_next(int pi) {
  // computeGandH for pi > 0 with this epoch's G~ and H held aside between the
  // thin-Q (which reads H) and the products (which overwrite G~ and H)
  if (pi == 0 || params_.missme || device_uv || !ctx_.overlap_ok()) return false;
  ctx_.omega_from(0, /*flip=*/plan == nullptr);
  ctx_.hold();
  products();
  return true;
}

namespace {

// Consecutive windows that use one Omega and add into one of H1 / H2 form a
// single product over their SNP range: G_b = A_b Omega for every window b of
// the range and H_x += sum_b A_b^T G_b are the same numbers whether the
// windows go one by one (src/Halko.cpp:195-203) or as one range, up to the
// order of the summation.  Windows are merged only on resident genotypes.
class Segments {
 public:
  Segments(Context& ctx, bool merge) : ctx_(ctx), merge_(merge) {}
  void add(uint64_t j0, uint64_t j1, int hmode) {
    if (merge_ && j1_ > j0_ && hmode == hmode_ && j0 == j1_) {
      j1_ = j1;
      return;
    }
    flush();
    j0_ = j0, j1_ = j1, hmode_ = hmode;
    if (!merge_) flush();
  }
  void flush() {
    if (j1_ > j0_) ctx_.sketch(j0_, j1_, hmode_);
    j0_ = j1_ = 0;
  }

 private:
  Context& ctx_;
  bool merge_;
  uint64_t j0_ = 0, j1_ = 0;
  int hmode_ = 0;
};

}  // namespace

void FancyRsvdOpData::computeGandH(int pi) {
  // src/Halko.cpp:155-224 (in-core) and :225-268 (out-of-core).  Windows
  // starting past the last SNP (possible when M < bands^2) are treated as
  // empty.
  const uint64_t bands = params_.bands;
  if (pi == 0) initOmg();
  if (std::pow(2, pi) >= (double)bands) {
    ctx_.zero_h(1);
    ctx_.zero_h(2);
  }
  if (pi == 0) {
    if (update) fit_with_pi();
    if (standardize && params_.scale == -9) ctx_.standardize(params_.ploidy);
    if (plan) {
      bandsize_ = plan->band_factor;
    } else {
      bandsize_ = 1;
      blocksize_ = (uint64_t)std::ceil((double)rows() / bands);
      if (params_.perm && !(update && params_.emu)) {
        std::vector<int> p(rows());
        std::iota(p.begin(), p.end(), 0);
        auto rng = std::default_random_engine{};
        std::shuffle(p.data(), p.data() + p.size(), rng);  // src/RSVD.hpp:61-70
        ctx_.permute_columns(p);
      }
    }
  }
  // The epoch's Omega updates run with their safety test on the device; if
  // one fails it (H near rank deficient), the epoch is run again from its
  // checkpoint with the host test, which takes the Householder fallback.
  const bool defer = ctx_.deferred_ok();
  const uint64_t bandsize0 = bandsize_, updates0 = omega_updates;
  if (defer) {
    ctx_.epoch_checkpoint();
    ctx_.deferred_thinq = true;
  }
  auto run = [&] { plan ? windows_blocks(pi) : windows(pi); };
  try {
    run();
  } catch (...) {
    ctx_.deferred_thinq = false;
    throw;
  }
  ctx_.deferred_thinq = false;
  if (defer && ctx_.epoch_thinq_failed()) {
    ctx_.epoch_rollback();
    bandsize_ = bandsize0;
    omega_updates = updates0;
    run();
  }
}

void FancyRsvdOpData::windows(int pi) {
  const uint64_t bands = params_.bands;
  bandsize_ = std::min<uint64_t>(bandsize_ * 2, bands);
  const uint64_t M = rows();
  Segments seg(ctx_, !ctx_.streamed());
  for (uint64_t b = 0, i = 1; b < bands; ++b, ++i) {
    const uint64_t start = b * blocksize_;
    const uint64_t stop = (b + 1) * blocksize_ >= M ? M - 1 : (b + 1) * blocksize_ - 1;
    if (start <= stop) seg.add(start, stop + 1, i <= bandsize_ / 2 ? 1 : 2);
    const bool adjacent = (pi > 0 && (double)(b + 1) == std::pow(2, pi - 1) && std::pow(2, pi) < (double)bands);
    if ((b + 1) < bandsize_ && !adjacent) continue;
    if (!((i == bandsize_) || (i == bandsize_ / 2) || adjacent)) continue;
    seg.flush();
    ctx_.omega_from(1);
    ++omega_updates;
    if (i == bandsize_) {
      ctx_.zero_h(1);
      i = 0;
    } else {
      ctx_.zero_h(2);
    }
  }
  seg.flush();
}

void FancyRsvdOpData::windows_blocks(int pi) {
  // src/Halko.cpp:225-268: the windows are the plan's blocks and the bands
  // grow from band_factor blocks.
  const BlockPlan& bp = *plan;
  const uint64_t bands = params_.bands;
  bandsize_ = std::min<uint64_t>(bandsize_ * 2, bp.nblocks);
  Segments seg(ctx_, !ctx_.streamed());  // streamed blocks are loaded one at a time
  for (uint64_t b = 0, i = 1; b < bp.nblocks; ++b, ++i) {
    seg.add(bp.start[b], bp.stop[b] + 1, i <= bandsize_ / 2 ? 1 : 2);
    const bool adjacent = (pi > 0 && (double)(b + 1) == std::pow(2, pi - 1) * (double)bp.band_factor &&
                           std::pow(2, pi) < (double)bands);
    if ((b + 1) < bandsize_ && !adjacent) continue;
    if (!((i == bandsize_) || (i == bandsize_ / 2) || adjacent)) continue;
    seg.flush();
    ctx_.omega_from(1);
    ++omega_updates;
    if (i == bandsize_) {
      ctx_.zero_h(1);
      i = 0;
    } else {
      ctx_.zero_h(2);
    }
  }
  seg.flush();
}

PcaResult run_pca(Context& ctx, const Param& params, RunInfo* info, double* U_out) {
  if (params.svd == 1 || params.svd == 2) return run_pca_with_halko(ctx, params, info, U_out);
  PcaResult r = params.svd == 3 ? run_pca_full(ctx, params, info) : run_pca_arnoldi(ctx, params, info);
  if (U_out && !r.U.empty()) {
    std::memcpy(U_out, r.U.data(), r.U.size() * sizeof(double));
    r.U.clear();
  }
  return r;
}

PcaResult run_pca_with_halko(Context& ctx, const Param& params, RunInfo* info, double* U_out) {
  // src/Halko.cpp:271-345 (genetic input, in-core, no LD output)
  ctx.reset_columns();  // restores the as-read state between runs; not part of the timed work
  BlockPlan plan;
  require(!ctx.streamed() || params.out_of_core, kErrArgument,
          "streamed genotypes run the out-of-core schedule: set memory (-m)");
  if (params.out_of_core) {
    // Data::prepare's plan (src/Data.cpp:42-84); with the shuffle, the file
    // is first rewritten in permute_plink order (src/Main.cpp:120-151) and
    // read from there, so allele frequencies follow their SNPs
    plan = plan_blocks(params, ctx.N, ctx.M);
    require(!ctx.streamed() || plan.blocksize <= ctx.stream_block_snps(), kErrArgument,
            "stream: the -m plan's blocks of " + std::to_string(plan.blocksize) +
                " SNPs exceed the streamed block capacity of " + std::to_string(ctx.stream_block_snps()));
    if (params.perm) {
      // a .pgen is not rewritten: its reader threads take a logical order
      // (src/Main.cpp:169-172, max_threads of src/Main.cpp:53-54)
      const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
      ctx.permute_columns(ctx.pgen_input ? pgen_perm_order(ctx.M, params.bands, plan.blocksize,
                                                           std::min(params.threads, hw), params.seed)
                                         : permute_plink_order(ctx.N, ctx.M, params.buffer, params.bands),
                          /*with_freq=*/true);
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<RsvdOpData> rsvd;
  if (params.svd == 2) {
    auto f = std::make_unique<FancyRsvdOpData>(ctx, params, (int)params.k, (int)params.oversamples);
    if (params.out_of_core) f->plan = &plan;
    rsvd = std::move(f);
  } else {
    auto n = std::make_unique<NormalRsvdOpData>(ctx, params, (int)params.k, (int)params.oversamples);
    if (params.out_of_core) n->plan = &plan;
    rsvd = std::move(n);
  }
  rsvd->info = info;
  rsvd->u_out = U_out;
  rsvd->need_v_host = params.printv || params.missme || params.ld;
  const int k = (int)params.k;
  const uint64_t N = ctx.N, M = ctx.M;
  if (!params.missme) {
    rsvd->setFlags(false, !params.ld);  // src/Halko.cpp:283-285: --ld keeps the matrix unscaled
    rsvd->computeUSV((int)params.maxp, params.tol);
  } else {
    if (ctx.nmiss == 0 && params.verbose) std::printf("warning: there is no missing values\n");
    // U and V stay on the device through the EM iterations (flip_UV, mev
    // and the imputation inputs computed there); read back once at the end
    rsvd->device_uv = true;
    rsvd->setFlags(false, false);
    rsvd->computeUSV((int)params.maxp, params.tol);
    ctx.flip_uv_device(k);
    for (uint32_t i = 0; i < params.maxiter; ++i) {
      rsvd->setFlags(true, false);
      ctx.save_vpre();
      rsvd->computeUSV((int)params.maxp, params.tol);
      ctx.flip_uv_device(k);
      const double diff = 1.0 - ctx.mev_vpre(k);
      if (info) info->em_trace.push_back(diff), info->em_iters = (int)i + 1;
      if (params.verbose) {
        std::printf("individual allele frequencies estimated iter = %u, diff = %f\n", i + 1, diff);
        std::fflush(stdout);
      }
      if (diff < params.tolem) break;
    }
    if (params.emu) {
      rsvd->setFlags(true, true);
      rsvd->computeUSV((int)params.maxp, params.tol);
      ctx.flip_uv_device(k);
    }
    if (U_out) {
      ctx.download_u(k, U_out);
    } else {
      rsvd->U.resize(N * (uint64_t)k);
      ctx.download_u(k, rsvd->U.data());
    }
    rsvd->V.resize(M * (uint64_t)k);
    ctx.download_v(k, rsvd->V.data());
  }
  PcaResult res;
  res.k = k;
  res.U = std::move(rsvd->U);  // empty when U went straight to U_out
  if (U_out && !res.U.empty()) std::memcpy(U_out, res.U.data(), res.U.size() * sizeof(double));
  res.S = rsvd->S;
  res.E.resize(res.S.size());
  for (size_t i = 0; i < res.S.size(); ++i) res.E[i] = res.S[i] * res.S[i] / (double)M;  // src/Halko.cpp:339
  if (rsvd->need_v_host && !rsvd->V.empty()) {
    // loadings in the original SNP order (src/Data.cpp:226-233)
    res.V.assign(rsvd->V.size(), 0.0);
    for (int c = 0; c < k; ++c)
      for (uint64_t i = 0; i < M; ++i) res.V[(uint64_t)c * M + ctx.perm[i]] = rsvd->V[(uint64_t)c * M + i];
  }
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace pcaone
