// Streamed genotypes: the packed matrix stays on the host and each product
// loads the SNP block it covers (the out-of-core reads of the reference,
// src/FilePlink.cpp:122-298, for matrices beyond HBM).
//
// Device memory then holds one block in both packed layouts (block x N / 4
// bytes each), its missing-entry lists, the raw rows of the block being
// uploaded, and the per-SNP state (M doubles per vector), instead of the
// whole matrix.  A block load is a parallel host gather of its rows (in the
// current column order, so a permuted order reads the file's rows in that
// order, as the reference reads its rewritten file) into pinned memory, the
// DMA and the interleave, chunk by chunk with the gather of one chunk beside
// the DMA of the one before, then the transpose and the missing lists.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstring>

#include "common.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "pgen_io.hpp"

namespace pcaone {

void Context::stream_close() {
  if (map_) munmap(map_, map_len_);
  map_ = nullptr;
  map_len_ = 0;
  src_rows_ = nullptr;
  pgen_.reset();
  pgen_input = false;
  sblock_ = 0;
  blk0_ = blk1_ = 0;
}

void Context::stream_setup(uint32_t nsamples, uint64_t nsnps, uint64_t block_snps) {
  require(nsamples > 0 && nsnps > 0, kErrArgument, "empty genotype matrix");
  require(nsnps < (1ull << 32), kErrArgument, "more than 2^32-1 SNPs");
  require(block_snps > 0, kErrArgument, "stream: block size must be positive");
  N = nsamples;
  M = nsnps;
  sblock_ = std::min<uint64_t>(block_snps, nsnps);
  bps_ = bytes_per_snp(N);
  ks_ = round_up(N, kKStep);
  kt_ = round_up(sblock_, kKStep);  // K extent of one block in the sample-major layout
  pitch_s_ = ks_ / 16;
  pitch_t_ = kt_ / 16;
  rows_s_ = round_up(sblock_, kRowPad) + kRowPad;  // one block of SNP rows
  rows_state_ = round_up(M, kRowPad) + kRowPad;     // the per-SNP state covers all SNPs
  rows_t_ = round_up(N, kRowPad);
  xs_.resize(rows_s_ * pitch_s_);
  xt_.resize(rows_t_ * pitch_t_);
  blk0_ = blk1_ = 0;
  L = lp = 0;
  stream_loads = 0;
  stream_load_ms = 0.0;
  perm.resize(M);
  for (uint64_t j = 0; j < M; ++j) perm[j] = (int)j;
  perm_identity_ = true;
  freq_follows_ = false;
  // one pass over all blocks counts the codes: F (src/FilePlink.cpp:37-55)
  // and the missing entries
  std::vector<uint32_t> h(4 * sblock_);
  DevBuf<uint32_t> cnt(4 * sblock_);
  F.assign(M, 0.0);
  nmiss = 0;
  for (uint64_t j0 = 0; j0 < M; j0 += sblock_) {
    const uint64_t n = std::min(sblock_, M - j0);
    block_rows_to_device(j0, n);
    launch_code_counts(xs_.get(), pitch_s_, n, N, cnt.get(), stream);
    PCAONE_HIP(hipMemcpyAsync(h.data(), cnt.get(), 4 * n * 4, hipMemcpyDeviceToHost, stream));
    PCAONE_HIP(hipStreamSynchronize(stream));
    for (uint64_t r = 0; r < n; ++r) {
      const uint64_t c00 = h[4 * r], c01 = h[4 * r + 1], c10 = h[4 * r + 2], c11 = h[4 * r + 3];
      const uint64_t c = c00 + c10 + c11;
      F[j0 + r] = c ? ((double)c00 + 0.5 * (double)c10) / (double)c : 0.0;
      nmiss += c01;
    }
  }
  dF_.resize(M);
  PCAONE_HIP(hipMemcpyAsync(dF_.get(), F.data(), M * 8, hipMemcpyHostToDevice, stream));
  pristine_columns();
  clear_impute();
  blk0_ = blk1_ = 0;
  lnmiss_ = lrows_ = 0;
  PCAONE_HIP(hipStreamSynchronize(stream));
}

void Context::stream_host(const uint8_t* rows, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps) {
  require(rows != nullptr, kErrArgument, "stream: null rows");
  stream_close();
  src_rows_ = rows;
  try {
    stream_setup(nsamples, nsnps, block_snps);
  } catch (...) {
    stream_close();
    throw;
  }
}

void Context::stream_bed(const std::string& bed, uint32_t nsamples, uint64_t nsnps, uint64_t block_snps) {
  stream_close();
  const int fd = ::open(bed.c_str(), O_RDONLY);
  require(fd >= 0, kErrIO, "can not open " + bed);
  struct stat st {};
  const bool ok = fstat(fd, &st) == 0;
  const uint64_t need = 3 + (uint64_t)nsnps * bytes_per_snp(nsamples);
  if (!ok || (uint64_t)st.st_size < need) {
    ::close(fd);
    fail(kErrIO, "truncated .bed file: " + bed);
  }
  void* p = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_SHARED, fd, 0);
  ::close(fd);
  require(p != MAP_FAILED, kErrIO, "can not map " + bed);
  const uint8_t* b = static_cast<const uint8_t*>(p);
  if (!(b[0] == 0x6c && b[1] == 0x1b && b[2] == 0x01)) {  // src/FilePlink.hpp:24-27
    munmap(p, (size_t)st.st_size);
    fail(kErrIO, "Invalid .bed file! Expected first two bytes to be the PLINK magic numbers and SNP-major mode.");
  }
  madvise(p, (size_t)st.st_size, MADV_SEQUENTIAL);
  map_ = p;
  map_len_ = (size_t)st.st_size;
  src_rows_ = b + 3;
  try {
    stream_setup(nsamples, nsnps, block_snps);
  } catch (...) {
    stream_close();
    throw;
  }
}

void Context::stream_pgen(std::shared_ptr<const PgenSource> src, uint64_t block_snps) {
  require(src != nullptr, kErrArgument, "stream: null .pgen source");
  stream_close();
  pgen_ = std::move(src);
  try {
    stream_setup((uint32_t)pgen_->nsamples(), pgen_->nsnps(), block_snps);
    pgen_input = true;
  } catch (...) {
    stream_close();
    throw;
  }
}

void Context::block_rows_to_device(uint64_t j0, uint64_t n) {
  // rows j0..j0+n of the current column order (file rows perm[j]) -> the
  // block's SNP-major layout xs_.  Two staging slots (pinned host + raw
  // device rows) of about 32 MB: the host gathers chunk c+1 into its pinned
  // slot while the DMA of chunk c runs on the copy stream, and the
  // interleave of chunk c on the context stream runs beside the DMA of
  // chunk c+1.  A slot is refilled once its DMA has read the pinned rows
  // (host wait) and its interleave has read the raw rows (copy stream wait).
  const auto t0 = std::chrono::steady_clock::now();
  if (!copy_stream_) PCAONE_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
  launch_fill_rows(xs_.get(), pitch_s_, 0, rows_s_, stream);
  const uint64_t chunk = std::max<uint64_t>(1, (32ull << 20) / std::max<uint64_t>(bps_, 1));
  const uint64_t slot = std::max<uint64_t>(std::min(n, chunk), 1) * bps_;
  pinned_.resize(2 * slot);
  raw_.resize(2 * slot);
  const bool ident = perm_identity_;
  // the raw slots are free here: each load ends with the interleaves done
  Event dma[2], done[2];
  for (uint64_t c = 0, r = 0; r < n; ++c, r += chunk) {
    const int s = (int)(c & 1);
    const uint64_t m = std::min(chunk, n - r), bytes = m * bps_;
    if (c >= 2) PCAONE_HIP(hipEventSynchronize(dma[s].get()));  // chunk c-2's DMA has read pinned slot s
    unsigned char* host = pinned_.get() + s * slot;
    if (pgen_) {
      pgen_->decode_rows(j0 + r, m, ident ? nullptr : perm.data(), host);
    } else {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)m; ++i) {
        const uint64_t j = j0 + r + (uint64_t)i;
        const uint64_t src = ident ? j : (uint64_t)perm[j];
        std::memcpy(host + (uint64_t)i * bps_, src_rows_ + src * bps_, bps_);
      }
    }
    if (c >= 2) PCAONE_HIP(hipStreamWaitEvent(copy_stream_, done[s].get(), 0));  // chunk c-2's interleave
    PCAONE_HIP(hipMemcpyAsync(raw_.get() + s * slot, host, bytes, hipMemcpyHostToDevice, copy_stream_));
    PCAONE_HIP(hipEventRecord(dma[s].get(), copy_stream_));
    PCAONE_HIP(hipStreamWaitEvent(stream, dma[s].get(), 0));
    launch_interleave(raw_.get() + s * slot, m, bps_, N, xs_.get() + r * pitch_s_, pitch_s_, stream);
    PCAONE_HIP(hipEventRecord(done[s].get(), stream));
  }
  // every interleave waited for its DMA: after this the pinned and raw
  // slots are free for the next load
  PCAONE_HIP(hipStreamSynchronize(stream));
  ++stream_loads;
  stream_load_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Context::load_block(uint64_t j0, uint64_t j1) {
  if (blk1_ != 0 && j0 >= blk0_ && j1 <= blk1_) return;  // already resident
  require(j1 - j0 <= sblock_, kErrArgument,
          "stream: a product covers " + std::to_string(j1 - j0) + " SNPs, more than the " + std::to_string(sblock_) +
              "-SNP device block");
  const uint64_t n = j1 - j0;
  block_rows_to_device(j0, n);
  launch_transpose_codes(xs_.get(), pitch_s_, n, N, xt_.get(), pitch_t_, rows_t_, stream);
  blk0_ = j0;
  blk1_ = j1;
  build_lists(n, n);  // synchronizes
}

}  // namespace pcaone
