// Device context implementation (see context.hpp for the HBM layout).
#include "context.hpp"

#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstring>

#include "kernels.hpp"
#include "synth_host.hpp"

namespace pcaone {

namespace {
constexpr double kVarTol = 1e-9;   // src/Data.hpp:7
constexpr uint64_t kChunkCapH = 16384;  // largest split-K chunk of the H product (codes)
constexpr uint64_t kL2Bytes = 2ull << 20;  // working set per chunk of gathered rows (half an XCD's L2)
constexpr uint64_t kMissPartialElems = 32ull << 20;  // FP64 partials of chunked corrections (256 MB)

uint64_t bytes_of(Form f) { return form_wbytes(f); }
}  // namespace

Context::Context(int dev) : device(dev) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) fail(kErrNoDevice, "no HIP device available");
  require(dev >= 0 && dev < n, kErrArgument, "device index out of range");
  PCAONE_HIP(hipSetDevice(dev));
  int ncu = 0;
  PCAONE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  cus_ = ncu > 0 ? (uint64_t)ncu : 0;
  PCAONE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

Context::~Context() {
  stream_close();
  if (stream) {
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
    if (gram_ev_) (void)hipEventDestroy(gram_ev_);
    if (copy_stream_) {
      (void)hipStreamSynchronize(copy_stream_);
      (void)hipStreamDestroy(copy_stream_);
    }
    (void)hipStreamDestroy(stream);
  }
}

// ------------------------------------------------------------------ data
void Context::reserve(uint32_t nsamples, uint64_t nsnps) {
  require(nsamples > 0 && nsnps > 0, kErrArgument, "empty genotype matrix");
  require(nsnps < (1ull << 32), kErrArgument, "more than 2^32-1 SNPs");
  stream_close();  // resident rows replace streamed ones
  N = nsamples;
  M = nsnps;
  bps_ = bytes_per_snp(N);
  ks_ = round_up(N, kKStep);
  kt_ = round_up(M, kKStep);
  pitch_s_ = ks_ / 16;
  pitch_t_ = kt_ / 16;
  rows_s_ = round_up(M, kRowPad) + kRowPad;
  rows_state_ = rows_s_;
  rows_t_ = round_up(N, kRowPad);
  xs_.resize(rows_s_ * pitch_s_);
  launch_fill_rows(xs_.get(), pitch_s_, 0, rows_s_, stream);
  nmiss = 0;
  F.clear();
  perm.clear();
  L = lp = 0;
}

void Context::upload_rows(const uint8_t* rows, uint64_t j0, uint64_t n) {
  require(N > 0, kErrState, "reserve() first");
  require(j0 + n <= M, kErrArgument, "upload beyond the reserved SNP count");
  if (!n) return;
  // Two staging slots (pinned host + raw device rows): the host copies chunk
  // c+1 into its pinned slot while the DMA of chunk c runs on the copy
  // stream, and the interleave of chunk c runs beside the DMA of chunk c+1
  // (the host waits for a slot's DMA, the copy stream for its interleave).
  if (!copy_stream_) PCAONE_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
  const uint64_t chunk = std::max<uint64_t>(1, (64ull << 20) / std::max<uint64_t>(bps_, 1));
  const uint64_t slot = std::min(n, chunk) * bps_;
  raw_.resize(2 * slot);
  pinned_.resize(2 * slot);
  Event dma[2], done[2];
  for (uint64_t c = 0, r = 0; r < n; ++c, r += chunk) {
    const int s = (int)(c & 1);
    const uint64_t m = std::min(chunk, n - r), bytes = m * bps_;
    if (c >= 2) PCAONE_HIP(hipEventSynchronize(dma[s].get()));
    unsigned char* host = pinned_.get() + s * slot;
    const uint8_t* src = rows + r * bps_;
    const int64_t pieces = 16;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < pieces; ++q) {
      const uint64_t b0 = bytes * q / pieces, b1 = bytes * (q + 1) / pieces;
      std::memcpy(host + b0, src + b0, b1 - b0);
    }
    if (c >= 2) PCAONE_HIP(hipStreamWaitEvent(copy_stream_, done[s].get(), 0));
    PCAONE_HIP(hipMemcpyAsync(raw_.get() + s * slot, host, bytes, hipMemcpyHostToDevice, copy_stream_));
    PCAONE_HIP(hipEventRecord(dma[s].get(), copy_stream_));
    PCAONE_HIP(hipStreamWaitEvent(stream, dma[s].get(), 0));
    launch_interleave(raw_.get() + s * slot, m, bps_, N, xs_.get() + (j0 + r) * pitch_s_, pitch_s_, stream);
    PCAONE_HIP(hipEventRecord(done[s].get(), stream));
  }
  PCAONE_HIP(hipStreamSynchronize(stream));  // every interleave waited for its DMA
}

void synth_rows(const synth::Spec& spec, uint32_t* xs, uint64_t pitch, hipStream_t stream) {
  // Per-SNP population frequencies (Beta draws) on the host, genotypes on the
  // device with the shared integer generator; SNPs that come out
  // monomorphic are redrawn with the next attempt, as the host writer does.
  const uint64_t M = spec.nsnps;
  const uint32_t N = spec.nsamples;
  const uint32_t P = spec.npops;
  const auto fst = synth::pop_fst(spec);
  const auto flags = synth::sample_flags(spec);
  const synth::MissSpec ms = synth::miss_spec(spec);
  std::vector<uint64_t> thr(M * P);
  std::vector<uint32_t> att(M, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t j = 0; j < (int64_t)M; ++j) synth::snp_thresholds(spec, fst, (uint64_t)j, 0, &thr[(uint64_t)j * P]);
  DevBuf<uint64_t> dthr(M * P);
  DevBuf<uint32_t> datt(M), drows(M), cnt(4 * M);
  DevBuf<uint8_t> dflags(N);
  PCAONE_HIP(hipMemcpyAsync(dthr.get(), thr.data(), thr.size() * 8, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipMemcpyAsync(datt.get(), att.data(), M * 4, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipMemcpyAsync(dflags.get(), flags.data(), N, hipMemcpyHostToDevice, stream));
  launch_synth(spec.seed, N, P, dthr.get(), datt.get(), dflags.get(), ms, nullptr, M, xs, pitch, stream);
  std::vector<uint32_t> h(4 * M);
  for (uint32_t round = 0;; ++round) {
    launch_code_counts(xs, pitch, M, N, cnt.get(), stream);
    PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), h.size() * 4, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    std::vector<uint32_t> redo;
    for (uint64_t j = 0; j < M; ++j) {
      const uint32_t c00 = h[4 * j], c10 = h[4 * j + 2], c11 = h[4 * j + 3];
      if (!(c10 > 0 || (c00 > 0 && c11 > 0))) redo.push_back((uint32_t)j);
    }
    if (redo.empty()) break;
    require(round < 63, kErrNumeric, "synth: a SNP stayed monomorphic after 64 attempts");
    for (uint32_t j : redo) {
      ++att[j];
      synth::snp_thresholds(spec, fst, j, att[j], &thr[(uint64_t)j * P]);
    }
    PCAONE_HIP(hipMemcpyAsync(dthr.get(), thr.data(), thr.size() * 8, hipMemcpyHostToDevice, stream));
    PCAONE_HIP(hipMemcpyAsync(datt.get(), att.data(), M * 4, hipMemcpyHostToDevice, stream));
    PCAONE_HIP(hipMemcpyAsync(drows.get(), redo.data(), redo.size() * 4, hipMemcpyHostToDevice, stream));
    launch_synth(spec.seed, N, P, dthr.get(), datt.get(), dflags.get(), ms, drows.get(), redo.size(), xs, pitch,
                 stream);
  }
}

void Context::generate(const synth::Spec& spec) {
  synth::validate(spec);
  reserve(spec.nsamples, spec.nsnps);
  synth_rows(spec, xs_.get(), pitch_s_, stream);
  finalize();
}

void Context::download_rows(uint8_t* out) {
  require(!F.empty(), kErrState, "no data");
  require(!streamed(), kErrArgument, "download: the genotypes are streamed from the host");
  const uint64_t chunk = std::max<uint64_t>(1, (256ull << 20) / bps_);
  raw_.resize(std::min(chunk, M) * bps_);
  for (uint64_t j0 = 0; j0 < M; j0 += chunk) {
    const uint64_t m = std::min(chunk, M - j0);
    launch_deinterleave(xs_.get() + j0 * pitch_s_, pitch_s_, m, N, raw_.get(), bps_, stream);
    PCAONE_HIP(hipMemcpyAsync(out + j0 * bps_, raw_.get(), m * bps_, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
  }
}

void Context::finalize() {
  require(N > 0 && M > 0, kErrState, "no genotypes uploaded");
  DevBuf<uint32_t> cnt(4 * M);
  launch_code_counts(xs_.get(), pitch_s_, M, N, cnt.get(), stream);
  std::vector<uint32_t> h(4 * M);
  PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), h.size() * 4, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  F.assign(M, 0.0);
  nmiss = 0;
  for (uint64_t j = 0; j < M; ++j) {
    const uint64_t c00 = h[4 * j], c01 = h[4 * j + 1], c10 = h[4 * j + 2], c11 = h[4 * j + 3];
    const uint64_t c = c00 + c10 + c11;
    // the reference sums BED2GENO values (1 and 0.5, exact) then divides once
    F[j] = c ? ((double)c00 + 0.5 * (double)c10) / (double)c : 0.0;
    nmiss += c01;
  }
  perm.resize(M);
  for (uint64_t j = 0; j < M; ++j) perm[j] = (int)j;
  perm_identity_ = true;
  freq_follows_ = false;
  dF_.resize(M);
  PCAONE_HIP(hipMemcpyAsync(dF_.get(), F.data(), M * 8, hipMemcpyHostToDevice, stream));
  pristine_columns();
  clear_impute();
  build_derived();
}

void Context::pristine_columns() {
  // padded (zeros) so window tiles may read past M
  cF_.resize(rows_state_);
  sc_.resize(rows_state_);
  launch_pristine_cols(dF_.get(), M, rows_state_, cF_.get(), sc_.get(), stream);
  weights_dirty_ = true;
  cols_pristine_ = true;
}

void Context::build_derived() {
  require(!streamed(), kErrInternal, "build_derived on streamed genotypes");
  xt_.resize(rows_t_ * pitch_t_);
  launch_transpose_codes(xs_.get(), pitch_s_, M, N, xt_.get(), pitch_t_, rows_t_, stream);
  build_lists(M, M);
}

void Context::build_lists(uint64_t rows, uint64_t ncodes_t) {
  // CSR lists of the code-01 entries of the loaded layouts: by SNP row of xs
  // (sample indices) and by sample row of xt (SNP indices local to xt)
  weights_dirty_ = true;
  lrows_ = rows;
  lnmiss_ = 0;
  if (nmiss == 0) {
    rp_s_.resize(rows + 1);
    rp_t_.resize((uint64_t)N + 1);
    PCAONE_HIP(hipMemsetAsync(rp_s_.get(), 0, (rows + 1) * 8, stream));
    PCAONE_HIP(hipMemsetAsync(rp_t_.get(), 0, ((uint64_t)N + 1) * 8, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    return;
  }
  auto lists = [&](const uint32_t* x, uint64_t pitch, uint64_t nrows, uint64_t ncodes, DevBuf<uint64_t>& rp,
                   DevBuf<uint32_t>& ix) {
    DevBuf<uint32_t> cnt(4 * nrows);
    launch_code_counts(x, pitch, nrows, ncodes, cnt.get(), stream);
    std::vector<uint32_t> h(4 * nrows);
    PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), h.size() * 4, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    std::vector<uint64_t> p(nrows + 1, 0);
    for (uint64_t r = 0; r < nrows; ++r) p[r + 1] = p[r] + h[4 * r + 1];
    rp.resize(nrows + 1);
    PCAONE_HIP(hipMemcpyAsync(rp.get(), p.data(), (nrows + 1) * 8, hipMemcpyHostToDevice, stream));
    ix.resize(std::max<uint64_t>(p[nrows], 1));
    launch_missing_lists(x, pitch, nrows, ncodes, rp.get(), ix.get(), stream);
    PCAONE_HIP(hipStreamSynchronize(stream));
    return p[nrows];
  };
  const uint64_t ns = lists(xs_.get(), pitch_s_, rows, N, rp_s_, ix_s_);
  const uint64_t nt = lists(xt_.get(), pitch_t_, N, ncodes_t, rp_t_, ix_t_);
  require(ns == nt && (streamed() || ns == nmiss), kErrInternal, "missing-entry lists disagree with the code counts");
  lnmiss_ = ns;
  mt_s_.built = mt_t_.built = false;
}

void Context::build_tiles(MissTiles& t, const DevBuf<uint64_t>& rp, const DevBuf<uint32_t>& ix, uint64_t nrows,
                          uint64_t kext, uint32_t span, bool balance) {
  // (1) entries per (row block, granule) on the device; (2) the tiles of each
  // block on the host: runs of granules from the first one holding entries,
  // as long as they hold at most kMissEMax entries and span at most `span`
  // indexes; (3) row offsets and local indexes on the device.
  require(kext <= 0xffffffffull && span >= (uint32_t)kMissKG, kErrInternal, "build_tiles: extent");
  const uint64_t nb = (nrows + kMissRB - 1) / kMissRB;
  const uint64_t ngran = (kext + kMissKG - 1) / kMissKG;
  t.rperm.release();
  t.rpos.release();
  mwork_.clear();
  if (balance && nrows > 1) {
    // the lane groups of a workgroup own fixed rows: rows of like entry
    // counts share a block, so no group waits for one with a much longer row
    std::vector<uint64_t> hp(nrows + 1);
    PCAONE_HIP(hipMemcpyAsync(hp.data(), rp.get(), hp.size() * 8, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    std::vector<uint32_t> order(nrows);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return hp[x + 1] - hp[x] > hp[y + 1] - hp[y]; });
    std::vector<uint32_t> perm(nb * kMissRB, 0xffffffffu), pos(nrows);
    for (uint64_t p = 0; p < nrows; ++p) {
      perm[p] = order[p];
      pos[order[p]] = (uint32_t)p;
    }
    t.rperm.resize(perm.size());
    t.rpos.resize(nrows);
    PCAONE_HIP(hipMemcpyAsync(t.rperm.get(), perm.data(), perm.size() * 4, hipMemcpyHostToDevice, stream));
    PCAONE_HIP(hipMemcpyAsync(t.rpos.get(), pos.data(), nrows * 4, hipMemcpyHostToDevice, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
  }
  DevBuf<uint32_t> cnt(std::max<uint64_t>(nb * ngran, 1));
  PCAONE_HIP(hipMemsetAsync(cnt.get(), 0, cnt.bytes(), stream));
  launch_miss_hist(rp.get(), ix.get(), nrows, t.rpos.size() ? t.rpos.get() : nullptr, (uint32_t)ngran, cnt.get(),
                   stream);
  std::vector<uint32_t> h(nb * ngran);
  if (!h.empty()) PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), h.size() * 4, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  t.h_tiles.clear();
  t.h_tb.assign(nb + 1, 0);
  uint64_t off = 0;
  const uint64_t gspan = span / kMissKG;
  for (uint64_t b = 0; b < nb; ++b) {
    const uint32_t* c = h.data() + b * ngran;
    uint64_t g = 0;
    while (true) {
      while (g < ngran && c[g] == 0) ++g;
      if (g == ngran) break;
      const uint64_t g0 = g;
      uint64_t n = 0;
      while (g < ngran && g - g0 < gspan && n + c[g] <= (uint64_t)kMissEMax) n += c[g++];
      MissTile tl{};
      tl.e0 = off;
      tl.k0 = (uint32_t)(g0 * kMissKG);
      tl.k1 = (uint32_t)std::min<uint64_t>(g * kMissKG, kext);
      tl.ne = (uint32_t)n;
      t.h_tiles.push_back(tl);
      off += miss_stored((uint32_t)n);
    }
    t.h_tb[b + 1] = (uint32_t)t.h_tiles.size();
  }
  const uint64_t ntiles = t.h_tiles.size();
  t.tiles.resize(std::max<uint64_t>(ntiles, 1));
  t.tb.resize(nb + 1);
  t.trow.resize(std::max<uint64_t>(ntiles, 1) * kMissRowStride);
  t.eix.resize(std::max<uint64_t>(off, 8));
  t.ew.resize(std::max<uint64_t>(off, 8));
  // the padding entries stay zero (the fill and the weights write [0, ne))
  PCAONE_HIP(hipMemsetAsync(t.eix.get(), 0, t.eix.bytes(), stream));
  PCAONE_HIP(hipMemsetAsync(t.ew.get(), 0, t.ew.bytes(), stream));
  if (ntiles)
    PCAONE_HIP(hipMemcpyAsync(t.tiles.get(), t.h_tiles.data(), ntiles * sizeof(MissTile), hipMemcpyHostToDevice,
                              stream));
  PCAONE_HIP(hipMemcpyAsync(t.tb.get(), t.h_tb.data(), (nb + 1) * 4, hipMemcpyHostToDevice, stream));
  DevBuf<int> st(1);
  PCAONE_HIP(hipMemsetAsync(st.get(), 0, sizeof(int), stream));
  launch_miss_fill(rp.get(), ix.get(), nrows, t.rperm.size() ? t.rperm.get() : nullptr, t.tiles.get(), t.tb.get(),
                   nb, t.trow.get(), t.eix.get(), st.get(), stream);
  int bad = 0;
  PCAONE_HIP(hipMemcpyAsync(&bad, st.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));  // the host copies and cnt go out of scope
  require(!bad, kErrInternal, "missing-entry tiles disagree with their planned entry counts");
  t.nrows = nrows;
  t.kext = kext;
  t.nblocks = nb;
  t.span = span;
  t.built = true;
}

const Context::MissWorkSet& Context::miss_work(const MissTiles& t, uint64_t r_lo, uint64_t r_hi, uint64_t k_lo,
                                               uint64_t k_hi) {
  const std::array<uint64_t, 6> key{(uint64_t)(&t == &mt_s_ ? 0 : 1), r_lo, r_hi, k_lo, k_hi, (uint64_t)L};
  auto it = mwork_.find(key);
  if (it != mwork_.end()) return it->second;
  // tiles of each block in the index range
  const uint64_t b0 = r_lo / kMissRB, b1 = (r_hi - 1) / kMissRB;
  std::vector<uint32_t> ta(b1 - b0 + 1), tz(b1 - b0 + 1);
  uint64_t total = 0;
  for (uint64_t b = b0; b <= b1; ++b) {
    const MissTile* lo = t.h_tiles.data() + t.h_tb[b];
    const MissTile* hi = t.h_tiles.data() + t.h_tb[b + 1];
    const MissTile* x0 = std::partition_point(lo, hi, [&](const MissTile& m) { return m.k1 <= k_lo; });
    const MissTile* x1 = std::partition_point(x0, hi, [&](const MissTile& m) { return m.k0 < k_hi; });
    ta[b - b0] = (uint32_t)(x0 - t.h_tiles.data());
    tz[b - b0] = (uint32_t)(x1 - t.h_tiles.data());
    total += x1 - x0;
  }
  // shares of about `per` tiles: about sixteen rounds of workgroups (one
  // per CU at a time), at least 8 tiles a share, the partial rows within
  // their buffer
  const uint64_t cus = std::max<uint64_t>(cus_, 1);
  uint64_t per = std::max<uint64_t>(8, (total + 16 * cus - 1) / (16 * cus));
  const uint64_t cap = mpart_.size() / std::max<uint64_t>((r_hi - r_lo) * (uint64_t)L, 1);
  std::vector<uint32_t> ns(ta.size());
  uint32_t slots = 1;
  while (true) {
    slots = 1;
    for (size_t i = 0; i < ta.size(); ++i) {
      ns[i] = (uint32_t)std::max<uint64_t>(1, (tz[i] - ta[i] + per - 1) / per);
      slots = std::max(slots, ns[i]);
    }
    if (slots == 1 || slots <= cap) break;
    per *= 2;
  }
  std::vector<MissWork> work;
  for (size_t i = 0; i < ta.size(); ++i) {
    const uint32_t n = tz[i] - ta[i];
    for (uint32_t z = 0; z < ns[i]; ++z)
      work.push_back(MissWork{(uint32_t)(b0 + i), ta[i] + (uint32_t)((uint64_t)n * z / ns[i]),
                              ta[i] + (uint32_t)((uint64_t)n * (z + 1) / ns[i]), z});
  }
  // the longest shares first, so that the short ones fill the last round
  std::stable_sort(work.begin(), work.end(),
                   [](const MissWork& x, const MissWork& y) { return x.t1 - x.t0 > y.t1 - y.t0; });
  MissWorkSet& ws = mwork_[key];
  ws.work.resize(work.size());
  ws.bslots.resize(ns.size());
  PCAONE_HIP(hipMemcpyAsync(ws.work.get(), work.data(), work.size() * sizeof(MissWork), hipMemcpyHostToDevice,
                            stream));
  PCAONE_HIP(hipMemcpyAsync(ws.bslots.get(), ns.data(), ns.size() * 4, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  ws.nwork = (uint32_t)work.size();
  ws.slots = slots;
  return ws;
}

void Context::filter_snps(const std::vector<uint32_t>& keep) {
  require(!keep.empty(), kErrArgument, "no SNPs left after filtering!");
  require(!streamed(), kErrArgument, "--maf filters do not apply to streamed genotypes (the reference's -m)");
  for (uint32_t j : keep) require(j < M, kErrArgument, "filter index out of range");
  DevBuf<uint32_t> idx(keep.size());
  PCAONE_HIP(hipMemcpyAsync(idx.get(), keep.data(), keep.size() * 4, hipMemcpyHostToDevice, stream));
  const uint64_t newM = keep.size();
  const uint64_t new_rows = round_up(newM, kRowPad) + kRowPad;
  tmpx_.resize(new_rows * pitch_s_);
  launch_fill_rows(tmpx_.get(), pitch_s_, 0, new_rows, stream);
  launch_gather_rows(xs_.get(), pitch_s_, idx.get(), newM, tmpx_.get(), stream);
  PCAONE_HIP(hipStreamSynchronize(stream));
  std::swap(xs_, tmpx_);
  tmpx_.release();
  std::vector<double> Fk(newM);
  for (uint64_t i = 0; i < newM; ++i) Fk[i] = F[keep[i]];
  M = newM;
  kt_ = round_up(M, kKStep);
  pitch_t_ = kt_ / 16;
  rows_s_ = new_rows;
  rows_state_ = new_rows;
  finalize();
  F = Fk;  // identical to the recount, kept explicit
}

void Context::permute_columns(const std::vector<int>& p, bool with_freq) {
  require(p.size() == M, kErrArgument, "permutation size mismatch");
  require(with_freq || !freq_follows_, kErrState, "column order mixes two permutation kinds");
  std::vector<uint32_t> idx(M);
  for (uint64_t i = 0; i < M; ++i) {
    require(p[i] >= 0 && (uint64_t)p[i] < M, kErrArgument, "permutation index out of range");
    idx[i] = (uint32_t)p[i];
  }
  DevBuf<uint32_t> d(M);
  PCAONE_HIP(hipMemcpyAsync(d.get(), idx.data(), M * 4, hipMemcpyHostToDevice, stream));
  if (!streamed()) {
    tmpx_.resize(rows_s_ * pitch_s_);
    launch_fill_rows(tmpx_.get(), pitch_s_, M, rows_s_, stream);
    launch_gather_rows(xs_.get(), pitch_s_, d.get(), M, tmpx_.get(), stream);
  }
  // the column state follows its columns (padding untouched)
  DevBuf<double> t(2 * M);
  launch_gather_f64(cF_.get(), d.get(), M, t.get(), stream);
  launch_gather_f64(sc_.get(), d.get(), M, t.get() + M, stream);
  PCAONE_HIP(hipMemcpyAsync(cF_.get(), t.get(), M * 8, hipMemcpyDeviceToDevice, stream));
  PCAONE_HIP(hipMemcpyAsync(sc_.get(), t.get() + M, M * 8, hipMemcpyDeviceToDevice, stream));
  if (with_freq) {
    launch_gather_f64(dF_.get(), d.get(), M, t.get(), stream);
    PCAONE_HIP(hipMemcpyAsync(dF_.get(), t.get(), M * 8, hipMemcpyDeviceToDevice, stream));
  }
  PCAONE_HIP(hipStreamSynchronize(stream));
  if (!streamed()) std::swap(xs_, tmpx_);  // the old layout stays allocated for the next permutation (every winSVD run has two)
  std::vector<int> np(M);
  perm_identity_ = true;
  for (uint64_t i = 0; i < M; ++i) {
    np[i] = perm[idx[i]];
    perm_identity_ = perm_identity_ && np[i] == (int)i;
  }
  perm = np;
  if (with_freq) freq_follows_ = !perm_identity_;
  weights_dirty_ = true;
  cols_pristine_ = false;
  if (emu_k_ > 0) {  // imputation loadings follow their columns
    const uint64_t kp = emu_kp_;
    std::vector<double> v(M * kp), w(M * kp);
    PCAONE_HIP(hipMemcpyAsync(v.data(), vimp_.get(), v.size() * 8, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    for (uint64_t i = 0; i < M; ++i) std::memcpy(&w[i * kp], &v[(uint64_t)idx[i] * kp], kp * 8);
    PCAONE_HIP(hipMemcpyAsync(vimp_.get(), w.data(), w.size() * 8, hipMemcpyHostToDevice, stream));
  }
  if (streamed()) {
    blk1_ = 0;  // the next product loads its block in the new order
    PCAONE_HIP(hipStreamSynchronize(stream));
  } else {
    build_derived();
  }
}

void Context::standardize(int ploidy) {
  require(!F.empty() && sc_.size() >= M, kErrState, "no data");
  // F as read, indexed by the current column: the reference indexes F by the
  // column position after any permutation (src/Data.cpp:356)
  launch_standardize(dF_.get(), M, std::sqrt((double)ploidy), kVarTol, sc_.get(), stream);
  weights_dirty_ = true;
  cols_pristine_ = false;
}

void Context::set_impute(const double* U, const double* S, const double* V, int kk) {
  require(kk > 0, kErrArgument, "impute rank must be positive");
  // rows padded to whole 128-byte lines, with room for three per-SNP values
  // after the k loadings of V (the device reads the rows as 16-byte pairs)
  const int kp = (kk + 3 + 15) / 16 * 16;
  std::vector<double> us((uint64_t)N * kp, 0.0), v(M * kp, 0.0);
  for (uint64_t i = 0; i < N; ++i)
    for (int t = 0; t < kk; ++t) us[i * kp + t] = U[(uint64_t)t * N + i] * S[t];
  for (uint64_t j = 0; j < M; ++j)
    for (int t = 0; t < kk; ++t) v[j * kp + t] = V[(uint64_t)t * M + j];
  us_.resize(us.size());
  vimp_.resize(v.size());
  PCAONE_HIP(hipMemcpyAsync(us_.get(), us.data(), us.size() * 8, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipMemcpyAsync(vimp_.get(), v.data(), v.size() * 8, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  emu_k_ = kk;
  emu_kp_ = kp;
  weights_dirty_ = true;
}

void Context::set_impute_device(const double* S, int kk) {
  require(kk > 0 && kk <= L, kErrArgument, "impute rank must be in [1, L]");
  const int kp = (kk + 3 + 15) / 16 * 16;  // as set_impute
  double* dS = small_.get() + 128 * 128 + 256;
  PCAONE_HIP(hipMemcpyAsync(dS, S, kk * sizeof(double), hipMemcpyHostToDevice, stream));
  us_.resize((uint64_t)N * kp);
  vimp_.resize(M * (uint64_t)kp);
  launch_impute_rows(ucur_.get(), N, lp, kk, kp, dS, us_.get(), stream);
  launch_impute_rows(vbuf_.get(), M, lp, kk, kp, nullptr, vimp_.get(), stream);
  PCAONE_HIP(hipStreamSynchronize(stream));  // dS came from caller memory
  emu_k_ = kk;
  emu_kp_ = kp;
  weights_dirty_ = true;
}

void Context::flip_uv_device(int k) {
  require(k >= 1 && k <= L, kErrArgument, "k out of range");
  launch_flip_cols_by_max(vbuf_.get(), M, ucur_.get(), N, lp, k, red_.get(), small_.get() + 128 * 128 + 128, stream);
}

void Context::save_vpre() {
  vpre_.resize(vbuf_.size());
  PCAONE_HIP(hipMemcpyAsync(vpre_.get(), vbuf_.get(), vbuf_.bytes(), hipMemcpyDeviceToDevice, stream));
}

double Context::mev_vpre(int k) {
  // mev(V, Vpre) = (1/k) sum_i || V^T vpre_i ||   (src/Utils.cpp:194-200)
  const la::Mat G = gram_l(vbuf_.get(), M, vpre_.get());
  double m = 0.0;
  for (int i = 0; i < k; ++i) {
    double s2 = 0.0;
    for (int a = 0; a < k; ++a) s2 += G(a, i) * G(a, i);
    m += std::sqrt(s2);
  }
  return m / k;
}

void Context::clear_impute() {
  if (emu_k_) weights_dirty_ = true;
  emu_k_ = 0;
}

void Context::reset_columns() {
  require(!F.empty(), kErrState, "no data");
  if (cols_pristine_ && !emu_k_) return;  // nothing changed since the last reset
  clear_impute();
  if (!perm_identity_) {
    std::vector<int> inv(M);
    for (uint64_t i = 0; i < M; ++i) inv[perm[i]] = (int)i;
    permute_columns(inv, freq_follows_);
  }
  pristine_columns();
}

// ------------------------------------------------------------------ rsvd
void Context::configure(int prec, int Lr, const GemmTuning& t) {
  require(N > 0 && M > 0 && !F.empty(), kErrState, "no genotype data (finalize first)");
  require(Lr >= 1 && Lr <= 128, kErrArgument, "k + oversamples must be in [1, 128]");
  require((uint64_t)Lr <= N && (uint64_t)Lr <= M, kErrArgument, "k + oversamples exceeds the matrix dimensions");
  require(prec >= 0 && prec < kNumPrecisions, kErrArgument, "unknown precision");
  precision = prec;
  form_g = forms_of(prec).g;
  form_h = forms_of(prec).h;
  L = Lr;
  lp = (int)round_up((uint64_t)Lr, 16);
  tuning = t;
  const uint64_t nrows_n = std::max<uint64_t>(rows_t_, ks_);
  omega_.resize(nrows_n * lp);
  omega_r_.resize(nrows_n * lp);
  omega_c_.resize(nrows_n * lp);
  colsum_.resize(lp);
  bvec_.resize(lp);
  if (form_colscaled(form_g) || form_colscaled(form_h)) {
    i8s_.resize(4 * 128);
    i8max_.resize(128);
  }
  wg_pitch_ = ks_;
  wg_.resize((uint64_t)form_pieces(form_g) * lp * wg_pitch_ * bytes_of(form_g));
  PCAONE_HIP(hipMemsetAsync(wg_.get(), 0, wg_.bytes(), stream));
  wh_pitch_ = kt_ + kKStep;
  wh_.resize((uint64_t)form_pieces(form_h) * lp * wh_pitch_ * bytes_of(form_h));
  gt_.resize(rows_state_ * lp);
  tmp_m_.resize(rows_state_ * lp);
  vbuf_.resize(rows_state_ * lp);
  H_.resize(nrows_n * lp);
  H1_.resize(nrows_n * lp);
  H2_.resize(nrows_n * lp);
  tmp_n_.resize(nrows_n * lp);
  ucur_.resize(nrows_n * lp);
  upre_.resize(nrows_n * lp);
  for (auto* b : {&H_, &H1_, &H2_, &gt_})
    PCAONE_HIP(hipMemsetAsync(b->get(), 0, b->bytes(), stream));
  small_.resize(2 * 128 * 128);
  red_cap_ = std::max<uint64_t>({gram_partial_elems(std::max<uint64_t>(M, N), lp), 2048 * 256,
                                 epi_g_fused_partial_elems(rows_s_, lp)});
  red_.resize(red_cap_ + 2 * 128 * 128);
  // partial buffers: the largest G window is M rows, the H product has rows_t_ rows
  const GemmShape dg = gemm_default_shape(form_g, lp), dh = gemm_default_shape(form_h, lp);
  shape_g = GemmShape{tuning.wm_g ? tuning.wm_g : dg.wm, tuning.waves_g ? tuning.waves_g : dg.waves};
  shape_h = GemmShape{tuning.wm_h ? tuning.wm_h : dh.wm, tuning.waves_h ? tuning.waves_h : dh.waves};
  require(gemm_shape_supported(form_g, shape_g, lp) && gemm_shape_supported(form_h, shape_h, lp), kErrArgument,
          "unsupported GEMM workgroup shape for this precision");
  miss_f32_ = precision != 0;  // FP64 form: FP64 sources
  cp_ = (int)round_up((uint64_t)L, miss_f32_ ? 4 : 2);
  // compact rows start on 128-byte lines, so a row read touches the fewest lines
  cpitch_ = (int)(round_up((uint64_t)cp_ * (miss_f32_ ? 4 : 8), 128) / (miss_f32_ ? 4 : 8));
  if (nmiss) {
    const uint64_t esz = miss_f32_ ? 4 : 8;
    csg_.resize(nrows_n * cpitch_ * esz);
    csh_.resize(rows_s_ * cpitch_ * esz);
    mpart_.resize(kMissPartialElems);
  }
  // the tiled corrections take FP32 sources of at most kMissMaxCp columns;
  // tiles are planned for the span that this L's source rows allow
  const bool tiled = nmiss && miss_f32_ && cp_ <= kMissMaxCp && !tuning.miss_gather;
  if (tiled != miss_tiled_) weights_dirty_ = true;
  miss_tiled_ = tiled;
  if (tiled && mt_s_.span != miss_tile_span(cp_)) {
    mt_s_.built = mt_t_.built = false;
    weights_dirty_ = true;
  }
  fixed_g_ = tuning.wm_g || tuning.waves_g;
  fixed_h_ = tuning.wm_h || tuning.waves_h;
  tuning.kchunk_g = round_up(tuning.kchunk_g, kKStep);  // 0 stays 0 (planned per launch)
  tuning.kchunk_h = round_up(tuning.kchunk_h, kKStep);
  const uint64_t pg =
      plan_partial_rows(rows_s_, ks_, tuning.kchunk_g, form_kcap(form_g, ks_)) * lp * form_accbytes(form_g);
  const uint64_t ph = plan_partial_rows(rows_t_, wh_pitch_, tuning.kchunk_h, form_kcap(form_h, kChunkCapH)) * lp *
                      form_accbytes(form_h);
  part_.resize(std::max(pg, ph));
  h_version_ = 0;
  gram_version_ = ~0ull;
  h_safe_version_ = ~0ull;
  gram_dev_version_ = ~0ull;
  if (thinq_on_device(lp)) {
    ldev_.resize(3 * (uint64_t)lp * lp + lp);
    lstatus_.resize(1);
    PCAONE_HIP(hipMemsetAsync(lstatus_.get(), 0, sizeof(int), stream));
  }
  ktimes = KernelTimes{};
  householder_fallbacks = gram_second_passes = epoch_rollbacks = stream_loads = 0;
  stream_load_ms = 0.0;
}

void Context::set_t(const la::Mat& T) {
  std::vector<double> h((size_t)lp * lp, 0.0);
  for (int i = 0; i < T.n; ++i)
    for (int j = 0; j < T.n; ++j) h[(size_t)i * lp + j] = T(i, j);
  PCAONE_HIP(hipMemcpyAsync(small_.get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, stream));
}

la::Mat Context::gram_l(const double* X, uint64_t rows, const double* Y, double* dev_out) {
  double* out = dev_out ? dev_out : red_.get() + red_cap_;
  launch_gram(X, nullptr, Y, rows, lp, red_.get(), red_cap_, out, stream);
  std::vector<double> h((size_t)lp * lp);
  PCAONE_HIP(hipMemcpyAsync(h.data(), out, h.size() * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  la::Mat G(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) G(i, j) = h[(size_t)i * lp + j];
  return G;
}

void Context::stage_in(const double* host, uint64_t count, double* dev) {
  // host -> pinned (parallel copy) -> device
  pinned_.resize(count * 8);
  unsigned char* st = pinned_.get();
  const int64_t pieces = 16;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < pieces; ++q) {
    const uint64_t b0 = count * q / pieces, b1 = count * (q + 1) / pieces;
    std::memcpy(st + b0 * 8, host + b0, (b1 - b0) * 8);
  }
  PCAONE_HIP(hipMemcpyAsync(dev, st, count * 8, hipMemcpyHostToDevice, stream));
}

void Context::stage_out(const double* dev, uint64_t count, double* host) {
  pinned_.resize(count * 8);
  unsigned char* st = pinned_.get();
  PCAONE_HIP(hipMemcpyAsync(st, dev, count * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  const int64_t pieces = 16;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < pieces; ++q) {
    const uint64_t b0 = count * q / pieces, b1 = count * (q + 1) / pieces;
    std::memcpy(host + b0, st + b0 * 8, (b1 - b0) * 8);
  }
}

void Context::set_omega(const double* om) {
  require(L > 0, kErrState, "configure() first");
  // column-major N x L as given, laid out row-major (ks_ x lp) on the device
  stage_in(om, (uint64_t)N * L, tmp_n_.get());
  omega_pieces();
}

double* Context::omega_staging() {
  require(L > 0, kErrState, "configure() first");
  omega_pinned_.resize((uint64_t)N * L * 8);
  return reinterpret_cast<double*>(omega_pinned_.get());
}

void Context::set_omega_staged() {
  require(L > 0 && omega_pinned_.get(), kErrState, "set_omega_staged: nothing staged");
  PCAONE_HIP(hipMemcpyAsync(tmp_n_.get(), omega_pinned_.get(), (uint64_t)N * L * 8, hipMemcpyHostToDevice, stream));
  omega_pieces();
}

void Context::omega_pieces() {
  launch_colmajor_to_rows(tmp_n_.get(), N, L, lp, ks_, omega_c_.get(), stream);
  if (form_colscaled(form_g))
    launch_col_scales(form_g, omega_c_.get(), N, lp, nullptr, i8max_.get(), i8_scale(0), i8_mul(0), stream);
  launch_make_pieces(form_g, omega_c_.get(), N, ks_, lp, nullptr, nullptr, wg_.get(), wg_pitch_,
                     (uint64_t)lp * wg_pitch_, omega_r_.get(), omega_.get(), stream, i8_scale(0),
                     nullptr, nullptr, red_.get(), red_cap_, colsum_.get());
  PCAONE_HIP(hipStreamSynchronize(stream));
}

void Context::run_gemm(bool g_product, const GemmShape& shape, uint64_t nchunks, const GemmArgs& a) {
  require(nchunks * a.rows * (uint64_t)lp * form_accbytes(g_product ? form_g : form_h) <= part_.bytes(),
          kErrInternal, "partial products exceed their buffer");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (time_kernels) {
    PCAONE_HIP(hipEventCreate(&e0));
    PCAONE_HIP(hipEventCreate(&e1));
    PCAONE_HIP(hipEventRecord(e0, stream));
  }
  launch_packed_gemm(g_product ? form_g : form_h, shape, a, stream);
  if (time_kernels) {
    PCAONE_HIP(hipEventRecord(e1, stream));
    PCAONE_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PCAONE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (g_product ? ktimes.g_ms : ktimes.h_ms) += ms;
    (g_product ? ktimes.g_launches : ktimes.h_launches) += 1;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
}

void Context::sketch(uint64_t j0, uint64_t j1, int hmode) {
  require(L > 0, kErrState, "configure() first");
  require(j0 < j1 && j1 <= M, kErrArgument, "sketch: bad SNP range");
  if (streamed()) load_block(j0, j1);
  const uint64_t nwin = j1 - j0;
  const uint64_t jl = j0 - xoff();  // the window's first row in the resident layouts
  const uint64_t wb = bytes_of(form_h);
  const int P = form_pieces(form_h);

  // (1) G[j0:j1] = A Omega : SNP-major rows, K = samples
  const GemmPlan pg = plan_gemm(form_g, shape_g, fixed_g_, lp, nwin, ks_, tuning.kchunk_g, form_kcap(form_g, ks_),
                                 cus_);
  GemmArgs g{};
  g.X = xs_.get() + jl * pitch_s_;
  g.xpitch = pitch_s_;
  g.rows = pg.rows;
  g.Wt = wg_.get();
  g.wpitch = wg_pitch_;
  g.wpiece = (uint64_t)lp * wg_pitch_;
  g.Y = part_.get();
  g.ychunk = g.rows * lp;
  g.k_begin = 0;
  g.k_end = ks_;
  g.kchunk = pg.kchunk;
  g.lp = lp;
  require(jl + g.rows <= rows_s_, kErrInternal, "G window exceeds padded rows");
  run_gemm(true, pg.shape, pg.nchunks, g);
  const uint64_t nch_g = pg.nchunks;
  double* gwin = gt_.get() + j0 * lp;
  // The window's operand pieces of diag(sc) G~ sit at K offset jl % kKStep of
  // wh_ (the window's K offset in xt); the K margins around them read as zero.
  const uint64_t off = jl % kKStep;
  const uint64_t klen = round_up(off + nwin, kKStep);
  launch_zero_ranges(wh_.get(), wh_pitch_ * wb, (uint64_t)P * lp, 0, off * wb, (off + nwin) * wb,
                     (klen - off - nwin) * wb, stream);
  if (!lnmiss_ && !form_colscaled(form_g) && !form_colscaled(form_h)) {
    launch_epi_g_fused(form_h, part_.get(), nch_g, g.ychunk, nwin, lp, sc_.get() + j0, cF_.get() + j0,
                       colsum_.get(), gwin, wh_.get() + off * wb, wh_pitch_, (uint64_t)lp * wh_pitch_, red_.get(),
                       bvec_.get(), stream);
  } else {
    launch_epi_g(form_g, part_.get(), nch_g, g.ychunk, nwin, nwin, lp, sc_.get() + j0, cF_.get() + j0,
                 colsum_.get(), gwin, stream, i8_mul(0));
  }
  if (lnmiss_ && miss_tiled_) {
    update_missing_weights();
    launch_compact_rows(omega_r_.get(), N, lp, L, cpitch_, true, csg_.get(), stream);
    MissTiledSpmmArgs m{};
    m.t = mt_s_.list();
    m.ew = mt_s_.ew.get();
    m.r_lo = jl;
    m.r_hi = jl + nwin;
    m.k_lo = 0;
    m.k_hi = N;
    m.src = reinterpret_cast<const float*>(csg_.get());
    m.src_k0 = 0;
    m.cp = cp_;
    m.cpitch = cpitch_;
    m.ncols = L;
    const MissWorkSet& ws = miss_work(mt_s_, m.r_lo, m.r_hi, m.k_lo, m.k_hi);
    m.work = ws.work.get();
    m.nwork = ws.nwork;
    m.slots = ws.slots;
    m.bslots = ws.bslots.get();
    m.rpos = mt_s_.rpos.size() ? mt_s_.rpos.get() : nullptr;
    m.partial = mpart_.get();
    m.partial_cap = mpart_.size();
    m.out = gwin;
    m.lp = lp;
    launch_missing_spmm_tiled(m, stream);
  } else if (lnmiss_) {
    update_missing_weights();
    launch_compact_rows(omega_r_.get(), N, lp, L, cpitch_, miss_f32_, csg_.get(), stream);
    MissSpmmArgs m{};
    m.rp = rp_s_.get() + jl;
    m.ix = ix_s_.get();
    m.w = w_s_.get();
    m.nrows = nwin;
    m.k_lo = 0;
    m.k_hi = N;
    m.src = csg_.get();
    m.f32 = miss_f32_;
    m.src_k0 = 0;
    m.cp = cp_;
    m.cpitch = cpitch_;
    m.ncols = L;
    m.kspan = miss_span(nwin, N);
    m.partial = mpart_.get();
    m.partial_cap = mpart_.size();
    m.out = gwin;
    m.lp = lp;
    launch_missing_spmm(m, stream);
  }
  if (lnmiss_ || form_colscaled(form_g) || form_colscaled(form_h)) {
    if (form_colscaled(form_h))
      launch_col_scales(form_h, gwin, nwin, lp, sc_.get() + j0, i8max_.get(), i8_scale(1), i8_mul(1), stream);
    launch_make_pieces(form_h, gwin, nwin, nwin, lp, sc_.get() + j0, nullptr, wh_.get() + off * wb, wh_pitch_,
                       (uint64_t)lp * wh_pitch_, gwin, nullptr, stream, i8_scale(1), sc_.get() + j0, cF_.get() + j0,
                       red_.get(), red_cap_, bvec_.get());
  }

  h_half(j0, j1, hmode);
}

void Context::h_half(uint64_t j0, uint64_t j1, int hmode) {
  // H (or H1 / H2) += A[j0:j1]^T G~[j0:j1], with the operand pieces of
  // diag(sc) G~ for these SNPs in wh_ at K offset j0 % kKStep and the b
  // terms in bvec_.
  const uint64_t nwin = j1 - j0;
  const uint64_t jl = j0 - xoff();
  const uint64_t off = jl % kKStep;
  const uint64_t klen = round_up(off + nwin, kKStep);
  double* gwin = gt_.get() + j0 * lp;
  // sample-major rows, K = the window's SNPs
  const GemmPlan ph = plan_gemm(form_h, shape_h, fixed_h_, lp, N, klen, tuning.kchunk_h,
                                 form_kcap(form_h, kChunkCapH), cus_);
  GemmArgs h{};
  h.X = xt_.get() + (jl - off) / 16;
  h.xpitch = pitch_t_;
  h.rows = ph.rows;
  h.Wt = wh_.get();
  h.wpitch = wh_pitch_;
  h.wpiece = (uint64_t)lp * wh_pitch_;
  h.Y = part_.get();
  h.ychunk = h.rows * lp;
  h.k_begin = 0;
  h.k_end = klen;
  h.kchunk = ph.kchunk;
  h.lp = lp;
  require(h.rows <= rows_t_, kErrInternal, "H rows exceed padding");
  run_gemm(false, ph.shape, ph.nchunks, h);
  const uint64_t nch_h = ph.nchunks;
  // hmode 0: H = ..., 1: H1 += ..., 2: H2 += ..., 3: H += ... (blocks of one pass)
  double* target = (hmode == 0 || hmode == 3) ? H_.get() : (hmode == 1 ? H1_.get() : H2_.get());
  launch_epi_h(form_h, part_.get(), nch_h, h.ychunk, N, rows_t_, lp, bvec_.get(), target, hmode == 0 ? 0 : 1, stream,
               i8_mul(1));
  if (lnmiss_ && miss_tiled_) {
    launch_compact_rows(gwin, nwin, lp, L, cpitch_, true, csh_.get(), stream);
    MissTiledSpmmArgs m{};
    m.t = mt_t_.list();
    m.ew = mt_t_.ew.get();
    m.r_lo = 0;
    m.r_hi = N;
    m.k_lo = jl;
    m.k_hi = jl + nwin;
    m.src = reinterpret_cast<const float*>(csh_.get());
    m.src_k0 = jl;
    m.cp = cp_;
    m.cpitch = cpitch_;
    m.ncols = L;
    const MissWorkSet& ws = miss_work(mt_t_, m.r_lo, m.r_hi, m.k_lo, m.k_hi);
    m.work = ws.work.get();
    m.nwork = ws.nwork;
    m.slots = ws.slots;
    m.bslots = ws.bslots.get();
    m.rpos = mt_t_.rpos.size() ? mt_t_.rpos.get() : nullptr;
    m.partial = mpart_.get();
    m.partial_cap = mpart_.size();
    m.out = target;
    m.lp = lp;
    launch_missing_spmm_tiled(m, stream);
  } else if (lnmiss_) {
    launch_compact_rows(gwin, nwin, lp, L, cpitch_, miss_f32_, csh_.get(), stream);
    MissSpmmArgs m{};
    m.rp = rp_t_.get();
    m.ix = ix_t_.get();
    m.w = w_t_.get();
    m.nrows = N;
    m.k_lo = jl;
    m.k_hi = jl + nwin;
    m.src = csh_.get();
    m.f32 = miss_f32_;
    m.src_k0 = jl;
    m.cp = cp_;
    m.cpitch = cpitch_;
    m.ncols = L;
    m.kspan = miss_span(N, nwin);
    m.partial = mpart_.get();
    m.partial_cap = mpart_.size();
    m.out = target;
    m.lp = lp;
    launch_missing_spmm(m, stream);
  }
  if (hmode == 0 || hmode == 3) ++h_version_;
}

void Context::h_from_g(int hmode) {
  // H = A^T G~ over all SNPs for a G~ the caller wrote into gt_ (rows 0..M)
  require(L > 0, kErrState, "configure() first");
  require(!form_colscaled(form_h), kErrArgument, "h_from_g: not in the int8 or fp8 form");
  require(!streamed(), kErrArgument, "h_from_g: not with streamed genotypes");
  const uint64_t wb = bytes_of(form_h);
  const int P = form_pieces(form_h);
  const uint64_t klen = round_up(M, kKStep);
  launch_zero_ranges(wh_.get(), wh_pitch_ * wb, (uint64_t)P * lp, 0, 0, M * wb, (klen - M) * wb, stream);
  launch_make_pieces(form_h, gt_.get(), M, M, lp, sc_.get(), nullptr, wh_.get(), wh_pitch_, (uint64_t)lp * wh_pitch_,
                     gt_.get(), nullptr, stream, nullptr, sc_.get(), cF_.get(), red_.get(), red_cap_, bvec_.get());
  if (lnmiss_) update_missing_weights();
  h_half(0, M, hmode);
}

uint64_t Context::miss_span(uint64_t nrows, uint64_t kext) const {
  // chunks of source rows that fit in L2, as many as the partial buffer holds
  const uint64_t esz = miss_f32_ ? 4 : 8;
  uint64_t sp = std::max<uint64_t>(1024, kL2Bytes / ((uint64_t)cpitch_ * esz));
  const uint64_t zmax = std::max<uint64_t>(1, mpart_.size() / std::max<uint64_t>(nrows * (uint64_t)L, 1));
  if ((kext + sp - 1) / sp > zmax) sp = (kext + zmax - 1) / zmax;
  return std::max(sp, (kext + 65534) / 65535);
}

void Context::update_missing_weights() {
  if (!lnmiss_ || !weights_dirty_) return;
  const uint64_t xo = xoff();
  if (miss_tiled_) {
    const uint32_t span = miss_tile_span(cp_);
    if (!mt_s_.built || mt_s_.span != span) build_tiles(mt_s_, rp_s_, ix_s_, lrows_, N, span, false);
    if (!mt_t_.built || mt_t_.span != span) build_tiles(mt_t_, rp_t_, ix_t_, N, lrows_, span, true);
    if (emu_k_) launch_pack_col_state(vimp_.get(), M, emu_kp_, emu_k_, sc_.get(), cF_.get(), dF_.get(), stream);
    const auto weights = [&](MissTiles& t, bool by_snp) {
      MissTiledWeightArgs a{};
      a.t = t.list();
      a.ew = t.ew.get();
      a.nrows = t.nrows;
      a.nblocks = t.nblocks;
      a.by_snp = by_snp;
      a.sc = sc_.get() + xo;
      a.cf = cF_.get() + xo;
      a.Fclamp = emu_k_ ? dF_.get() + xo : nullptr;
      a.US = us_.get();
      a.V = emu_k_ ? vimp_.get() + xo * emu_kp_ : nullptr;
      a.kk = emu_k_;
      a.kp = emu_kp_;
      if (!t.nrows) return;
      const MissWorkSet& ws = miss_work(t, 0, t.nrows, 0, t.kext);
      a.work = ws.work.get();
      a.nwork = ws.nwork;
      launch_missing_weights_tiled(a, stream);
    };
    weights(mt_s_, true);
    weights(mt_t_, false);
    weights_dirty_ = false;
    weights_tiled_ = true;
    return;
  }
  weights_tiled_ = false;
  w_s_.resize(lnmiss_);
  w_t_.resize(lnmiss_);
  // the lists index SNPs from the first loaded row: the per-SNP state too
  MissWeightArgs a{};
  a.sc = sc_.get() + xo;
  a.cf = cF_.get() + xo;
  a.Fclamp = emu_k_ ? dF_.get() + xo : nullptr;
  a.US = us_.get();
  a.V = emu_k_ ? vimp_.get() + xo * emu_kp_ : nullptr;
  a.kk = emu_k_;
  a.kp = emu_kp_;
  if (emu_k_) launch_pack_col_state(vimp_.get(), M, emu_kp_, emu_k_, sc_.get(), cF_.get(), dF_.get(), stream);
  // index chunks keep the gathered rows of US (by SNP) or V (by sample) in L2
  const auto span = [&](uint64_t kext) {
    if (!emu_k_) return std::max<uint64_t>(kext, 1);
    uint64_t sp = std::max<uint64_t>(1024, kL2Bytes / ((uint64_t)emu_kp_ * 8));
    return std::max(sp, (kext + 65534) / 65535);
  };
  a.rp = rp_s_.get();
  a.ix = ix_s_.get();
  a.nrows = lrows_;
  a.by_snp = true;
  a.w = w_s_.get();
  a.kext = N;
  a.kspan = span(N);
  launch_missing_weights(a, stream);
  a.rp = rp_t_.get();
  a.ix = ix_t_.get();
  a.nrows = N;
  a.by_snp = false;
  a.w = w_t_.get();
  a.kext = lrows_;
  a.kspan = span(lrows_);
  launch_missing_weights(a, stream);
  weights_dirty_ = false;
}

void Context::zero_h(int which) {
  DevBuf<double>& b = which == 0 ? H_ : (which == 1 ? H1_ : H2_);
  PCAONE_HIP(hipMemsetAsync(b.get(), 0, b.bytes(), stream));
  if (which == 0) ++h_version_;
}

void Context::omega_from(int src, bool flip) {
  require(L > 0, kErrState, "configure() first");
  if (src == 1) {
    launch_add(H1_.get(), H2_.get(), H_.get(), H_.size(), stream);
    ++h_version_;
  }
  const double* X = H_.get();
  const bool deferred = deferred_thinq && deferred_ok();
  if (!deferred && !h_thinq_safe()) {
    // H is (nearly) rank deficient: CholeskyQR2 could break down or lose
    // orthogonality.  The thin Q comes from a Householder QR instead, the
    // reference's own factorization (src/Halko.cpp:121-122), which gives L
    // orthonormal columns whatever the rank.
    omega_householder(flip);
    return;
  }
  if (thinq_on_device(lp)) {
    // every step stream-ordered on the device: no host round trip
    double* gh = ldev_.get();
    double* g2 = gh + (uint64_t)lp * lp;
    double* r1 = g2 + (uint64_t)lp * lp;
    double* dsign = r1 + (uint64_t)lp * lp;
    if (gram_dev_version_ != h_version_) {
      launch_gram(X, nullptr, nullptr, N, lp, red_.get(), red_cap_, gh, stream);
      gram_dev_version_ = h_version_;
    }
    launch_thinq_first(gh, L, lp, r1, small_.get(), lstatus_.get(), stream, deferred ? kThinqMinPivot : 0.0);
    launch_apply(X, small_.get(), N, lp, tmp_n_.get(), stream);
    launch_gram(tmp_n_.get(), nullptr, nullptr, N, lp, red_.get(), red_cap_, g2, stream);
    launch_thinq_second(g2, r1, X, L, lp, small_.get(), lstatus_.get(), stream);
    launch_apply(X, small_.get(), N, lp, omega_c_.get(), stream);
    if (flip) {  // flipOmg (src/RSVD.hpp:80-89) against Omg2
      double* norms = red_.get() + red_cap_;
      launch_flip_norms(omega_.get(), omega_c_.get(), N, lp, red_.get(), norms, stream);
      launch_flip_signs(norms, L, lp, dsign, stream);
    } else {
      dsign = nullptr;
    }
    if (form_colscaled(form_g))
      launch_col_scales(form_g, omega_c_.get(), N, lp, nullptr, i8max_.get(), i8_scale(0), i8_mul(0), stream);
    launch_make_pieces(form_g, omega_c_.get(), N, ks_, lp, nullptr, dsign, wg_.get(), wg_pitch_,
                       (uint64_t)lp * wg_pitch_, omega_r_.get(), omega_.get(), stream, i8_scale(0),
                       nullptr, nullptr, red_.get(), red_cap_, colsum_.get());
    return;
  }
  const la::Mat& G1 = gram_h_;  // h_thinq_safe() left the Gram of this H there
  const la::Mat R1 = la::chol_upper(G1);
  set_t(la::inv_upper(R1));
  launch_apply(X, small_.get(), N, lp, tmp_n_.get(), stream);
  const la::Mat R2 = la::chol_upper(gram_l(tmp_n_.get(), N));
  const la::Mat Rinv = la::inv_upper(la::matmul(R2, R1));
  // Householder signs from the top L x L block of Q = X R^-1
  std::vector<double> top((size_t)L * lp);
  PCAONE_HIP(hipMemcpyAsync(top.data(), X, top.size() * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  la::Mat Xt(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) Xt(i, j) = top[(size_t)i * lp + j];
  const std::vector<double> s = la::householder_signs(la::matmul(Xt, Rinv));
  la::Mat T = Rinv;
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) T(i, j) *= s[j];
  set_t(T);
  launch_apply(X, small_.get(), N, lp, omega_c_.get(), stream);
  double* dsign = nullptr;
  if (flip) {  // flipOmg (src/RSVD.hpp:80-89) against Omg2
    double* norms = red_.get() + red_cap_;
    launch_flip_norms(omega_.get(), omega_c_.get(), N, lp, red_.get(), norms, stream);
    std::vector<double> ab(2 * (size_t)lp);
    PCAONE_HIP(hipMemcpyAsync(ab.data(), norms, ab.size() * 8, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    std::vector<double> sign(lp, 1.0);
    for (int l = 0; l < L; ++l)
      if (ab[l] > 2.0 * ab[lp + l]) sign[l] = -1.0;
    dsign = small_.get() + 128 * 128;
    PCAONE_HIP(hipMemcpyAsync(dsign, sign.data(), lp * 8, hipMemcpyHostToDevice, stream));
  }
  if (form_colscaled(form_g))
    launch_col_scales(form_g, omega_c_.get(), N, lp, nullptr, i8max_.get(), i8_scale(0), i8_mul(0), stream);
  launch_make_pieces(form_g, omega_c_.get(), N, ks_, lp, nullptr, dsign, wg_.get(), wg_pitch_,
                     (uint64_t)lp * wg_pitch_, omega_r_.get(), omega_.get(), stream, i8_scale(0),
                     nullptr, nullptr, red_.get(), red_cap_, colsum_.get());
  PCAONE_HIP(hipStreamSynchronize(stream));
}

bool Context::deferred_ok() const { return L > 0 && thinq_on_device(lp) && !streamed(); }

namespace {
template <class T>
void copy_dev(DevBuf<T>& dst, const DevBuf<T>& src, hipStream_t s) {
  if (dst.size() < src.size()) dst.resize(src.size());
  if (src.size()) PCAONE_HIP(hipMemcpyAsync(dst.get(), src.get(), src.bytes(), hipMemcpyDeviceToDevice, s));
}
template <class T>
void restore_dev(DevBuf<T>& dst, const DevBuf<T>& src, hipStream_t s) {
  if (dst.size()) PCAONE_HIP(hipMemcpyAsync(dst.get(), src.get(), dst.bytes(), hipMemcpyDeviceToDevice, s));
}
}  // namespace

void Context::epoch_checkpoint() {
  copy_dev(ck_H1_, H1_, stream);
  copy_dev(ck_H2_, H2_, stream);
  copy_dev(ck_omega_, omega_, stream);
  copy_dev(ck_omega_r_, omega_r_, stream);
  copy_dev(ck_colsum_, colsum_, stream);
  copy_dev(ck_i8s_, i8s_, stream);
  copy_dev(ck_wg_, wg_, stream);
  if (lstatus_.get()) PCAONE_HIP(hipMemsetAsync(lstatus_.get(), 0, sizeof(int), stream));
}

bool Context::epoch_thinq_failed() {
  int st = 0;
  if (lstatus_.get()) PCAONE_HIP(hipMemcpyAsync(&st, lstatus_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  return st != 0;
}

void Context::epoch_rollback() {
  restore_dev(H1_, ck_H1_, stream);
  restore_dev(H2_, ck_H2_, stream);
  restore_dev(omega_, ck_omega_, stream);
  restore_dev(omega_r_, ck_omega_r_, stream);
  restore_dev(colsum_, ck_colsum_, stream);
  restore_dev(i8s_, ck_i8s_, stream);
  restore_dev(wg_, ck_wg_, stream);
  if (lstatus_.get()) PCAONE_HIP(hipMemsetAsync(lstatus_.get(), 0, sizeof(int), stream));
  ++h_version_;  // every cached Gram of H is stale
  ++epoch_rollbacks;
}

la::Mat Context::from_pinned_gram(int which) const {
  const double* src = reinterpret_cast<const double*>(gram_pinned_.get()) + (uint64_t)which * lp * lp;
  la::Mat G(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) G(i, j) = src[(size_t)i * lp + j];
  return G;
}

bool Context::h_thinq_safe() {
  if (h_safe_version_ == h_version_) return h_safe_;
  if (gram_version_ != h_version_) {
    if (gram_ev_ && gram_h_sent_ && begin_h_version_ == h_version_) {
      // finish_begin already sent this H's Gram to pinned memory
      PCAONE_HIP(hipEventSynchronize(gram_ev_));
      gram_h_ = from_pinned_gram(1);
    } else if (thinq_on_device(lp)) {
      // the device thin-Q reuses this Gram
      gram_h_ = gram_l(H_.get(), N, nullptr, ldev_.get());
      gram_dev_version_ = h_version_;
    } else {
      gram_h_ = gram_l(H_.get(), N);
    }
    gram_version_ = h_version_;
  }
  la::Mat R;
  double ratio = 0.0;
  h_safe_ = la::chol_try(gram_h_, R, ratio) && ratio > kThinqMinPivot;
  h_safe_version_ = h_version_;
  return h_safe_;
}

void Context::omega_householder(bool flip) {
  ++householder_fallbacks;
  std::vector<double> h((uint64_t)N * lp);
  PCAONE_HIP(hipMemcpyAsync(h.data(), H_.get(), h.size() * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  la::householder_q(h.data(), N, L, lp);
  for (uint64_t i = 0; i < N; ++i)
    for (int j = L; j < lp; ++j) h[i * lp + j] = 0.0;
  PCAONE_HIP(hipMemcpyAsync(omega_c_.get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, stream));
  double* dsign = nullptr;
  if (flip) {  // flipOmg (src/RSVD.hpp:80-89) against Omg2
    double* norms = red_.get() + red_cap_;
    launch_flip_norms(omega_.get(), omega_c_.get(), N, lp, red_.get(), norms, stream);
    dsign = small_.get() + 128 * 128;
    launch_flip_signs(norms, L, lp, dsign, stream);
  }
  if (form_colscaled(form_g))
    launch_col_scales(form_g, omega_c_.get(), N, lp, nullptr, i8max_.get(), i8_scale(0), i8_mul(0), stream);
  launch_make_pieces(form_g, omega_c_.get(), N, ks_, lp, nullptr, dsign, wg_.get(), wg_pitch_,
                     (uint64_t)lp * wg_pitch_, omega_r_.get(), omega_.get(), stream, i8_scale(0),
                     nullptr, nullptr, red_.get(), red_cap_, colsum_.get());
  PCAONE_HIP(hipStreamSynchronize(stream));
}

void Context::orthonormal_factor(const la::Mat& G1, const double* g) {
  // W such that Q = G~ W has orthonormal columns spanning range(G~).
  // - Well conditioned (every Cholesky pivot above kOnePassMinPivot of the
  //   largest diagonal entry): W = R1^-1, one CholeskyQR pass; Q is
  //   orthonormal to about u / min pivot ratio <= 1e-10.
  // - Ill conditioned: a real second pass over G~ on the device, R2 from the
  //   Gram of G~ W1, W = W1 R2^-1 (CholeskyQR2).
  // - (Nearly) rank deficient (breakdown, or a pivot below kThinqMinPivot):
  //   W1 = P_r Lambda_r^-1/2 from the eigen-decomposition of the Gram, keeping
  //   the r directions above kRankTol of the largest, then the second pass.
  //   The other L - r columns of W are zero, like the zero rows that the
  //   reference's rank-revealing solve (fullPivHouseholderQr, src/Halko.cpp:68)
  //   gives B for pivots below its threshold.
  la::Mat R1;
  double ratio = 0.0;
  const bool ok = la::chol_try(G1, R1, ratio);
  if (ok && ratio > kOnePassMinPivot) {
    W_ = la::inv_upper(R1);
    return;
  }
  ++gram_second_passes;
  la::Mat W1(L);
  int r = L;
  if (ok && ratio > kThinqMinPivot) {
    W1 = la::inv_upper(R1);
  } else {
    std::vector<double> lam;
    la::Mat P;
    la::sym_eig_desc(G1, lam, P);
    r = 0;
    while (r < L && lam[r] > kRankTol * lam[0] && lam[r] > 0.0) ++r;
    require(r > 0, kErrNumeric, "G is zero: no direction to orthonormalize");
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < r; ++j) W1(i, j) = P(i, j) / std::sqrt(lam[j]);
  }
  // second pass: G2 = (G~ W1)^T (G~ W1), R2 = chol of its leading r x r block
  set_t(W1);
  launch_apply(g, small_.get(), M, lp, tmp_m_.get(), stream);
  const la::Mat G2 = gram_l(tmp_m_.get(), M);
  la::Mat G2r(r);
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < r; ++j) G2r(i, j) = G2(i, j);
  const la::Mat R2i = la::inv_upper(la::chol_upper(G2r));
  W_ = la::Mat(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < r; ++j) {
      double v = 0.0;
      for (int q = 0; q <= j; ++q) v += W1(i, q) * R2i(q, j);
      W_(i, j) = v;
    }
}

EpochResult Context::finish_epoch(int k, bool have_prev) {
  finish_begin();
  return finish_end(k, have_prev);
}

void Context::finish_begin() {
  require(L > 0, kErrState, "configure() first");
  const uint64_t l2 = (uint64_t)lp * lp;
  gram_pinned_.resize(2 * l2 * 8 + 64);
  double* hp = reinterpret_cast<double*>(gram_pinned_.get());
  if (!gram_ev_) PCAONE_HIP(hipEventCreateWithFlags(&gram_ev_, hipEventDisableTiming));
  held_ = false;
  // Gram of G~ (first QR of G, src/Halko.cpp:55-60)
  double* gout = red_.get() + red_cap_;
  launch_gram(gt_.get(), nullptr, nullptr, M, lp, red_.get(), red_cap_, gout, stream);
  PCAONE_HIP(hipMemcpyAsync(hp, gout, l2 * 8, hipMemcpyDeviceToHost, stream));
  // Gram of H (B B^T); the device copy serves the next epoch's thin-Q of H
  gram_h_sent_ = gram_version_ != h_version_;
  begin_h_version_ = h_version_;
  if (gram_h_sent_) {
    const bool dev = thinq_on_device(lp);
    double* hout = dev ? ldev_.get() : red_.get() + red_cap_ + l2;
    launch_gram(H_.get(), nullptr, nullptr, N, lp, red_.get(), red_cap_, hout, stream);
    PCAONE_HIP(hipMemcpyAsync(hp + l2, hout, l2 * 8, hipMemcpyDeviceToHost, stream));
    if (dev) gram_dev_version_ = h_version_;
  }
  // a breakdown in the device thin-Q of the last epoch
  int* st = reinterpret_cast<int*>(hp + 2 * l2);
  *st = 0;
  if (lstatus_.get()) PCAONE_HIP(hipMemcpyAsync(st, lstatus_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipEventRecord(gram_ev_, stream));
}

bool Context::overlap_ok() const { return L > 0 && thinq_on_device(lp) && !time_kernels; }

void Context::hold() {
  if (gt_alt_.size() < gt_.size()) {
    gt_alt_.resize(gt_.size());
    PCAONE_HIP(hipMemsetAsync(gt_alt_.get(), 0, gt_alt_.bytes(), stream));
  }
  if (H_alt_.size() < H_.size()) {
    H_alt_.resize(H_.size());
    PCAONE_HIP(hipMemsetAsync(H_alt_.get(), 0, H_alt_.bytes(), stream));
  }
  std::swap(gt_, gt_alt_);
  std::swap(H_, H_alt_);
  held_ = true;
}

void Context::unhold() {
  if (!held_) return;
  std::swap(gt_, gt_alt_);
  std::swap(H_, H_alt_);
  held_ = false;
  h_version_ = begin_h_version_;  // H is the held epoch's again
  h_safe_version_ = ~0ull;
  // the discarded epoch's thin-Q may have flagged a breakdown: not this run's
  if (lstatus_.get()) PCAONE_HIP(hipMemsetAsync(lstatus_.get(), 0, sizeof(int), stream));
}

EpochResult Context::finish_end(int k, bool have_prev) {
  require(k >= 1 && k <= L, kErrArgument, "k out of range");
  PCAONE_HIP(hipEventSynchronize(gram_ev_));
  const uint64_t l2 = (uint64_t)lp * lp;
  const double* hp = reinterpret_cast<const double*>(gram_pinned_.get());
  if (*reinterpret_cast<const int*>(hp + 2 * l2)) {  // a device thin-Q that passed h_thinq_safe() broke down
    PCAONE_HIP(hipMemsetAsync(lstatus_.get(), 0, sizeof(int), stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    unhold();
    fail(kErrNumeric, "Cholesky breakdown (Gram matrix not positive definite)");
  }
  // QR of G twice (src/Halko.cpp:55-65): Q = G~ W with W = R^-1.  Only the
  // L x L factor is formed: W from the Cholesky factor of the Gram of G~.
  const la::Mat G1 = from_pinned_gram(0);
  orthonormal_factor(G1, held_ ? gt_alt_.get() : gt_.get());
  const la::Mat& Rinv = W_;
  // B B^T = R^-T (H^T H) R^-1; B = R^-T H^T (src/Halko.cpp:68)
  la::Mat HH;
  if (gram_h_sent_) {
    HH = from_pinned_gram(1);
    gram_h_ = HH;
    gram_version_ = begin_h_version_;
  } else {
    HH = gram_h_;
  }
  const la::Mat C = la::matmul(la::transpose(Rinv), la::matmul(HH, Rinv));
  std::vector<double> w;
  la::sym_eig_desc(C, w, WB_);
  sigma_.resize(L);
  for (int i = 0; i < L; ++i) sigma_[i] = std::sqrt(std::max(w[i], 0.0));
  // Ucur = right singular vectors of B = H R^-1 W_B Sigma^-1 (first k)
  la::Mat T(L);
  const la::Mat RW = la::matmul(Rinv, WB_);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < k; ++j) T(i, j) = sigma_[j] > 0 ? RW(i, j) / sigma_[j] : 0.0;
  set_t(T);
  launch_apply(held_ ? H_alt_.get() : H_.get(), small_.get(), N, lp, ucur_.get(), stream);
  EpochResult r;
  r.sigma = sigma_;
  if (have_prev) {
    // mev(Ucur, Upre) = (1/k) sum_i || Ucur^T upre_i ||   (src/Utils.cpp:194-200)
    double* out = red_.get() + red_cap_;
    launch_gram(ucur_.get(), nullptr, upre_.get(), N, lp, red_.get(), red_cap_, out, stream);
    std::vector<double> h((size_t)lp * lp);
    PCAONE_HIP(hipMemcpyAsync(h.data(), out, h.size() * 8, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    double m = 0.0;
    for (int i = 0; i < k; ++i) {
      double s2 = 0.0;
      for (int a = 0; a < k; ++a) s2 += h[(size_t)a * lp + i] * h[(size_t)a * lp + i];
      m += std::sqrt(s2);
    }
    r.diff = 1.0 - m / k;
  }
  return r;
}

void Context::save_prev() {
  PCAONE_HIP(hipMemcpyAsync(upre_.get(), ucur_.get(), (uint64_t)N * lp * 8, hipMemcpyDeviceToDevice, stream));
}

void Context::download_u(int k, double* U) {
  launch_rows_to_colmajor(ucur_.get(), N, lp, k, tmp_n_.get(), stream);
  stage_out(tmp_n_.get(), (uint64_t)N * k, U);
}

void Context::download_rowmajor(const double* src, uint64_t rows, double* out) {
  std::vector<double> h(rows * (uint64_t)lp);
  PCAONE_HIP(hipMemcpyAsync(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  for (int j = 0; j < L; ++j)
    for (uint64_t i = 0; i < rows; ++i) out[(uint64_t)j * rows + i] = h[i * lp + j];
}

void Context::download_g(uint64_t j0, uint64_t j1, double* G) {
  require(L > 0, kErrState, "configure() first");
  require(j0 < j1 && j1 <= M, kErrArgument, "get_g: bad SNP range");
  download_rowmajor(gt_.get() + j0 * lp, j1 - j0, G);
}

void Context::download_h(int which, double* H) {
  require(L > 0, kErrState, "configure() first");
  require(which >= 0 && which <= 2, kErrArgument, "get_h: which must be 0 (H), 1 (H1) or 2 (H2)");
  download_rowmajor(which == 0 ? H_.get() : (which == 1 ? H1_.get() : H2_.get()), N, H);
}

void Context::residuals_prepare(int k, const double* U, const double* S, const double* V, bool as_read) {
  require(N > 0 && M > 0, kErrState, "no data");
  require(!streamed(), kErrArgument, "residuals need the genotypes resident (not streamed)");
  require(k >= 0 && k <= kResidMaxK, kErrArgument, "residuals: rank must be in [0, " + std::to_string(kResidMaxK) + "]");
  require(k == 0 || (U && S && V), kErrArgument, "residuals: null U, S or V");
  std::vector<double> ut((uint64_t)std::max(k, 1) * N, 0.0), sv(M * (uint64_t)std::max(k, 1), 0.0);
  for (int t = 0; t < k; ++t)
    for (uint64_t i = 0; i < N; ++i) ut[(uint64_t)t * N + i] = U[(uint64_t)t * N + i];
  for (uint64_t j = 0; j < M; ++j)
    for (int t = 0; t < k; ++t) sv[j * k + t] = S[t] * V[(uint64_t)t * M + j];
  res_ut_.resize(ut.size());
  res_sv_.resize(sv.size());
  PCAONE_HIP(hipMemcpyAsync(res_ut_.get(), ut.data(), ut.size() * 8, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipMemcpyAsync(res_sv_.get(), sv.data(), sv.size() * 8, hipMemcpyHostToDevice, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  res_k_ = k;
  res_as_read_ = as_read;
}

void Context::residuals_rows(uint64_t j0, uint64_t j1, float* out) {
  require(res_sv_.get(), kErrState, "residuals_prepare first");
  require(j0 < j1 && j1 <= M, kErrArgument, "residuals: bad SNP range");
  const uint64_t rows = j1 - j0;
  res_out_.resize(rows * N);
  res_part_.resize(residual_partial_elems(N, rows));
  ResidualArgs a{};
  a.xs = xs_.get();
  a.pitch = pitch_s_;
  a.n = N;
  a.j0 = j0;
  a.j1 = j1;
  a.sc = res_as_read_ ? nullptr : sc_.get();
  a.cf = res_as_read_ ? dF_.get() : cF_.get();
  a.k = res_k_;
  a.ut = res_ut_.get();
  a.sv = res_sv_.get();
  if (!res_as_read_ && emu_k_ > 0) {
    a.kimp = emu_k_;
    a.kpimp = emu_kp_;
    a.us = us_.get();
    a.vimp = vimp_.get();
    a.fclamp = dF_.get();
  }
  a.part = res_part_.get();
  a.out = res_out_.get();
  launch_residuals(a, stream);
  PCAONE_HIP(hipMemcpyAsync(out, res_out_.get(), rows * N * sizeof(float), hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
}

void Context::compute_v(int k) {
  // V = Q U_B[:, :k] = G~ R^-1 W_B[:, :k]   (src/Halko.cpp:84)
  const la::Mat RW = la::matmul(W_, WB_);
  la::Mat T(L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < k; ++j) T(i, j) = RW(i, j);
  set_t(T);
  launch_apply(gt_.get(), small_.get(), M, lp, vbuf_.get(), stream);
  PCAONE_HIP(hipStreamSynchronize(stream));
}

void Context::download_v(int k, double* V) {
  std::vector<double> h(M * (uint64_t)lp);
  PCAONE_HIP(hipMemcpyAsync(h.data(), vbuf_.get(), h.size() * 8, hipMemcpyDeviceToHost, stream));
  PCAONE_HIP(hipStreamSynchronize(stream));
  for (uint64_t i = 0; i < M; ++i)
    for (int j = 0; j < k; ++j) V[(uint64_t)j * M + i] = h[i * lp + j];
}

}  // namespace pcaone
