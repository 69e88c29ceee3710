// PLINK 2 .pgen hardcall reader (src/FilePgen.cpp:27-287, src/FilePgen.hpp:14-48).
//
// The reference reads a .pgen through pgenlib (PgenReader::ReadHardcalls,
// allele index 1: ALT counts 0/1/2 and -3 for missing) one variant per call.
// Here the file is memory-mapped, its variant index is read once, and rows of
// variants are decoded in parallel straight into .bed rows:
//   ALT count 2 -> code 00, 1 -> 10, 0 -> 11, missing -> 01,
// the code whose BED2GENO value (src/Common.hpp:49-59) is count / 2, the value
// pgen2dosage gives (src/FilePgen.cpp:16-20).
//
// File layout read here (the PGEN format as pgenlib reads it,
// external/pgenlib/include/pgenlib_read.cc):
//   bytes 0-1 6c 1b; byte 2 the storage mode:
//     0x01        a PLINK 1 .bed payload (SNP-major) after the 3 bytes;
//     0x02-0x04   fixed width: uint32 M, uint32 N, a control byte, optional
//                 non-reference flags, then M records of ceil(N/4) bytes
//                 (+ 2N dosage bytes for 0x03, + 4N for 0x04);
//     0x10, 0x11  variable width (0x11: extensions follow the index);
//     0x20, 0x21  variable width with the index in <file>.pgi.
//   Variable width: uint32 M, uint32 N, control byte c; one uint64 file offset
//   per block of 65,536 variants (the first gives the first record); per
//   block the record types (4 bits each, low nibble first, when c & 15 < 4;
//   8 bits when 4..7; a packed code table for 8..14), the record lengths
//   (1 + (c & 3) bytes each), allele counts when (c >> 4) & 3 (rejected, as
//   the reference's loader rejects them), non-reference flags when c >> 6 == 3.
//   Main-track record types (type & 7):
//     0 a 2-bit array; 1 two genotypes as a bit array (a code byte lo * 4 +
//     (hi - lo), then N bits) plus a difference list; 2 a difference list
//     against the last earlier variant whose type is not 2 or 3, 3 the same
//     and then 0 and 2 swapped; 4 / 6 / 7 genotype 0 / 2 / 3 everywhere plus a
//     difference list; 5 all genotype 0, no bytes.
//   Difference list: varint length L (at most N / 8); for every group of 64
//   entries its first sample index (1 + floor(log2 N) / 8 bytes); one byte
//   per group but the last (skip sizes, not needed here); the L 2-bit
//   genotypes; within each group the index deltas as varints.
//   Bits 3 (multiallelic hardcalls) of a record type are rejected; bits 4-7
//   (phase, dosage) mark tracks after the main one, which hardcalls skip.
#include "pgen_io.hpp"

#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>

#include "common.hpp"

namespace pcaone {

namespace {

constexpr uint64_t kVblock = 65536;

// A map of 2-bit values applied to the 4 entries of a byte.
struct ByteLut {
  uint8_t t[256];
  explicit ByteLut(const uint8_t (&code)[4]) {
    for (int b = 0; b < 256; ++b) {
      uint8_t o = 0;
      for (int e = 0; e < 4; ++e) o |= (uint8_t)(code[(b >> (2 * e)) & 3] << (2 * e));
      t[b] = o;
    }
  }
};
constexpr uint8_t kAltToBed[4] = {3, 2, 0, 1};  // ALT count 0, 1, 2, missing -> .bed code
constexpr uint8_t kSwap[4] = {2, 1, 0, 3};       // record type 3: ALT counts 0 and 2 swapped
const ByteLut kToBed(kAltToBed);
const ByteLut kSwap02(kSwap);

struct Mapped {
  void* p = nullptr;
  size_t len = 0;
};

Mapped map_file(const std::string& path) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  require(fd >= 0, kErrIO, "Failed to open " + path);
  struct stat st {};
  if (fstat(fd, &st) != 0) {
    ::close(fd);
    fail(kErrIO, "can not stat " + path);
  }
  Mapped m;
  m.len = (size_t)st.st_size;
  if (m.len > 0) {
    m.p = mmap(nullptr, m.len, PROT_READ, MAP_SHARED, fd, 0);
    if (m.p == MAP_FAILED) {
      ::close(fd);
      fail(kErrIO, "can not map " + path);
    }
  }
  ::close(fd);
  return m;
}

uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

uint64_t rdn(const uint8_t* p, int nbytes) {
  uint64_t v = 0;
  for (int b = 0; b < nbytes; ++b) v |= (uint64_t)p[b] << (8 * b);
  return v;
}

[[noreturn]] void malformed(uint64_t j) {
  fail(kErrIO, "malformed .pgen record (variant " + std::to_string(j + 1) + ")");
}

// varint of at most 31 bits (7 bits per byte, low first)
uint32_t varint(const uint8_t*& p, const uint8_t* end, uint64_t j) {
  uint32_t v = 0;
  for (int shift = 0; shift < 35; shift += 7) {
    if (p >= end) malformed(j);
    const uint32_t b = *p++;
    v |= (b & 127u) << shift;
    if (b < 128) {
      if (shift == 28 && b > 7) malformed(j);
      return v;
    }
  }
  malformed(j);
}

inline void set2(uint8_t* g, uint64_t i, uint32_t v) {
  uint8_t& b = g[i >> 2];
  const int s = (int)(i & 3) * 2;
  b = (uint8_t)((b & ~(3u << s)) | (v << s));
}

}  // namespace

struct PgenSource::Scratch {
  std::vector<uint8_t> ld;  // decoded genotypes of the cached LD base
  uint64_t ld_idx = ~0ull;
};

uint64_t psam_samples(const std::string& psam) {
  std::ifstream f(psam);
  require(f.is_open(), kErrIO, "Cannot open psam file: " + psam);
  uint64_t n = 0;
  std::string line;
  while (std::getline(f, line))
    if (!line.empty() && line[0] != '#') ++n;
  return n;
}

PgenSource::PgenSource(const std::string& pgen, uint64_t nsamples, bool hardcall) {
  require(nsamples > 0 && nsamples < (1ull << 31), kErrIO, "bad sample count for " + pgen);
  Mapped m = map_file(pgen);
  map_ = m.p;
  map_len_ = m.len;
  base_ = static_cast<const uint8_t*>(map_);
  try {
    const uint64_t fsize = map_len_;
    require(fsize >= 4, kErrIO, "Error: " + pgen + " is too small to be a valid .pgen file.");
    require(base_[0] == 0x6c && base_[1] == 0x1b, kErrIO,
            "Error: " + pgen + " is not a .pgen file (first two bytes don't match the magic number).");
    mode_ = base_[2];
    n_ = nsamples;
    const uint64_t bps = bytes_per_snp(n_);
    if (mode_ == 0x00) fail(kErrIO, "Error: pgenlib does not directly support sample-major PLINK 1 .bed files.");
    if (mode_ == 0x01) {
      // PLINK 1 payload: constant records of ceil(N/4) bytes after 3 bytes
      require((fsize - 3) % bps == 0, kErrIO, "Error: Unexpected PLINK 1 .bed file size.");
      m_ = (fsize - 3) / bps;
      plink1_ = true;
      const_width_ = bps;
      const_offset_ = 3;
      return;
    }
    require(mode_ <= 0x04 || mode_ == 0x10 || mode_ == 0x11 || mode_ == 0x20 || mode_ == 0x21, kErrIO,
            "Error: Third byte of " + pgen + " does not correspond to a storage mode supported by this reader.");
    // the index: in the file itself, or in <file>.pgi
    Mapped pgi;
    const uint8_t* h = base_;
    uint64_t hlen = fsize;
    if (mode_ >= 0x20) {
      pgi = map_file(pgen + ".pgi");
      h = static_cast<const uint8_t*>(pgi.p);
      hlen = pgi.len;
      if (!(hlen >= 12 && h[0] == 0x6c && h[1] == 0x1b && (h[2] & 0xfe) == 0x30)) {
        if (pgi.p) munmap(pgi.p, pgi.len);
        fail(kErrIO, "Error: " + pgen + ".pgi is not a .pgen.pgi file.");
      }
    }
    struct Unmap {
      Mapped& m;
      ~Unmap() {
        if (m.p) munmap(m.p, m.len);
      }
    } unmap{pgi};
    require(hlen >= 12, kErrIO, "Error: " + pgen + " is too small to be a valid .pgen file.");
    m_ = rd32(h + 3);
    const uint64_t file_n = rd32(h + 7);
    const uint32_t ctrl = h[11];
    require(m_ > 0 && m_ < (1ull << 31), kErrIO, "Error: Invalid variant count in " + pgen);
    require(file_n == n_, kErrIO,
            "Error: the .psam lists " + std::to_string(n_) + " samples, but " + pgen + " contains " +
                std::to_string(file_n));
    const uint32_t nonref = ctrl >> 6;
    if (mode_ <= 0x04) {
      require((ctrl & 63) == 0, kErrIO, "Error: fixed-width .pgen with a variable-width control byte: " + pgen);
      const_offset_ = 12 + (nonref == 3 ? (m_ + 7) / 8 : 0);
      const_width_ = bps + (mode_ == 3 ? 2 * n_ : mode_ == 4 ? 4 * n_ : 0);
      const_vrtype_ = mode_ == 3 ? 0x40 : mode_ == 4 ? 0xc0 : 0;
      dosage_ = mode_ >= 3;
      require(const_offset_ + m_ * const_width_ == fsize, kErrIO, "Error: Unexpected .pgen file size: " + pgen);
    } else {
      require(((ctrl >> 4) & 3) == 0, kErrIO,
              "Storing of allele count information is not supported (only bi-allelic variants should be present).");
      const uint32_t st = ctrl & 15;
      const uint64_t nvb = (m_ + kVblock - 1) / kVblock;
      uint64_t at = 12;
      require(at + 8 * nvb <= hlen, kErrIO, "Error: truncated .pgen header: " + pgen);
      uint64_t pos = rdn(h + at, 8);
      at += 8 * nvb;
      vrtype_.resize(m_ + 1, 0);
      fpos_.resize(m_ + 1);
      // packed single-sample encodings (st 8..14): entry -> (type, length)
      uint8_t tab_t[16] = {5, 0};
      uint64_t tab_l[16] = {0, bps};
      int entry_bits = 2;
      if (st >= 8) {
        require(st != 13 && st != 15, kErrIO, "Error: Twelfth byte of " + pgen + " is not a supported format.");
        for (int e = 2; e < 16; ++e) tab_t[e] = 0xff, tab_l[e] = 0;
        if (st == 8) entry_bits = 1;
        if (st == 9) tab_t[2] = tab_t[3] = 8, tab_l[2] = bps + 2, tab_l[3] = bps + 3;
        if (st == 10) tab_t[2] = tab_t[3] = 16, tab_l[2] = bps + 1, tab_l[3] = bps + 2;
        if (st == 11) {
          entry_bits = 4;
          for (int e = 2; e < 8; ++e) tab_t[e] = 8, tab_l[e] = bps + e, tab_t[e + 8] = 24, tab_l[e + 8] = bps + 1 + e;
          tab_t[8] = 0xff;
          tab_t[9] = 16, tab_l[9] = bps + 1;
        }
        if (st == 12) tab_t[2] = 0x45, tab_l[2] = 2, tab_t[3] = 0x40, tab_l[3] = 3;
        if (st == 14) {
          entry_bits = 4;
          tab_t[3] = 0x10, tab_l[3] = 2;
          tab_t[4] = 0x45, tab_l[4] = 2;
          tab_t[5] = 0x40, tab_l[5] = 3;
          tab_t[12] = 0xc5, tab_l[12] = 4;
          tab_t[13] = 0xc0, tab_l[13] = 5;
          tab_t[15] = 0xe0, tab_l[15] = 6;
        }
        require(st == 8 || st == 9 || st == 10 || st == 11 || n_ == 1, kErrIO,
                "Error: single-sample .pgen encoding with more than one sample: " + pgen);
      }
      for (uint64_t vb = 0; vb < nvb; ++vb) {
        const uint64_t j0 = vb * kVblock, ct = std::min(kVblock, m_ - j0);
        if (st >= 8) {
          const uint64_t nbytes = (ct * entry_bits + 7) / 8;
          require(at + nbytes <= hlen, kErrIO, "Error: truncated .pgen header: " + pgen);
          for (uint64_t r = 0; r < ct; ++r) {
            const uint64_t bit = r * entry_bits;
            const uint32_t e = (h[at + bit / 8] >> (bit % 8)) & ((1u << entry_bits) - 1);
            require(tab_t[e] != 0xff, kErrIO, "Error: invalid record code in " + pgen);
            vrtype_[j0 + r] = tab_t[e];
            fpos_[j0 + r] = pos;
            pos += tab_l[e];
          }
          at += nbytes;
        } else {
          const bool wide = st >= 4;
          const uint64_t tbytes = wide ? ct : (ct + 1) / 2;
          const int lb = 1 + (int)(st & 3);
          require(at + tbytes + ct * lb <= hlen, kErrIO, "Error: truncated .pgen header: " + pgen);
          for (uint64_t r = 0; r < ct; ++r)
            vrtype_[j0 + r] = wide ? h[at + r] : (uint8_t)((h[at + r / 2] >> (4 * (r & 1))) & 15);
          at += tbytes;
          for (uint64_t r = 0; r < ct; ++r) {
            fpos_[j0 + r] = pos;
            pos += rdn(h + at + r * lb, lb);
          }
          at += ct * lb;
        }
        if (nonref == 3) at += (ct + 7) / 8;
      }
      fpos_[m_] = pos;
      require(pos <= fsize, kErrIO, "Error: .pgen records extend past the end of " + pgen);
      for (uint64_t j = 0; j < m_; ++j) {
        require((vrtype_[j] & 8) == 0, kErrIO,
                "multiallelic hardcalls in variant " + std::to_string(j + 1) + " of " + pgen +
                    " (only bi-allelic variants are read)");
        if (vrtype_[j] & 0x60) dosage_ = true;
      }
    }
    require(!dosage_ || hardcall, kErrIO,
            pgen + " stores dosages; this reader takes the hardcall track only (set hardcall, the reference's "
                   "--hardcall)");
  } catch (...) {
    if (map_) munmap(map_, map_len_);
    map_ = nullptr;
    throw;
  }
}

PgenSource::~PgenSource() {
  if (map_) munmap(map_, map_len_);
}

void PgenSource::record(uint64_t j, const uint8_t** p, const uint8_t** end) const {
  if (const_width_) {
    *p = base_ + const_offset_ + j * const_width_;
    *end = *p + const_width_;
  } else {
    *p = base_ + fpos_[j];
    *end = base_ + fpos_[j + 1];
  }
}

uint64_t PgenSource::ldbase_of(uint64_t j) const {
  // the last earlier variant that is not LD-compressed
  for (uint64_t i = j; i-- > 0;)
    if ((vrtype_[i] & 6) != 2) return i;
  malformed(j);
}

void PgenSource::decode_one(uint64_t j, uint8_t* g, Scratch& s) const {
  const uint64_t bps = bytes_per_snp(n_);
  const uint8_t* p;
  const uint8_t* end;
  record(j, &p, &end);
  const uint32_t vt = const_width_ ? const_vrtype_ : vrtype_[j];
  const uint32_t main = vt & 7;
  auto apply_difflist = [&](uint8_t* dst) {
    const uint32_t len = varint(p, end, j);
    if (len == 0) return;
    if (len > n_ / 8) malformed(j);
    const uint32_t groups = (len + 63) / 64;
    const int sb = 1 + (63 - __builtin_clzll(n_)) / 8;
    const uint8_t* first = p;
    const uint64_t index_bytes = (uint64_t)groups * (sb + 1) - 1;
    const uint64_t geno_bytes = (len + 3) / 4;
    if ((uint64_t)(end - p) < index_bytes + geno_bytes) malformed(j);
    const uint8_t* genos = p + index_bytes;
    p = genos + geno_bytes;
    uint64_t id = 0;
    for (uint32_t e = 0; e < len; ++e) {
      if (e % 64 == 0)
        id = rdn(first + (uint64_t)(e / 64) * sb, sb);
      else
        id += varint(p, end, j);
      if (id >= n_) malformed(j);
      set2(dst, id, (genos[e / 4] >> (2 * (e % 4))) & 3);
    }
  };
  if ((main & 6) == 2) {
    // LD-compressed: the base variant's genotypes, then the differences
    const uint64_t b = ldbase_of(j);
    if (s.ld_idx != b) {
      s.ld.resize(bps);
      decode_one(b, s.ld.data(), s);
      s.ld_idx = b;
    }
    std::memcpy(g, s.ld.data(), bps);
    apply_difflist(g);
    if (main == 3)
      for (uint64_t i = 0; i < bps; ++i) g[i] = kSwap02.t[g[i]];  // genotypes 0 and 2 swapped
    return;
  }
  if (main == 0) {
    if ((uint64_t)(end - p) < bps) malformed(j);
    std::memcpy(g, p, bps);
  } else if (main == 1) {
    const uint64_t nbits = (n_ + 7) / 8;
    if ((uint64_t)(end - p) < 1 + nbits) malformed(j);
    const uint32_t code = *p++;
    const uint32_t lo = code >> 2, hi = lo + (code & 3);
    if (hi > 3 || hi == lo) malformed(j);
    // 4 samples per output byte: a nibble of the bit array -> 4 genotypes
    uint8_t nib[16];
    for (int v = 0; v < 16; ++v) {
      uint8_t o = 0;
      for (int e = 0; e < 4; ++e) o |= (uint8_t)((((v >> e) & 1) ? hi : lo) << (2 * e));
      nib[v] = o;
    }
    for (uint64_t i = 0; i < bps; ++i) g[i] = nib[(p[i >> 1] >> (4 * (i & 1))) & 15];
    p += nbits;
    apply_difflist(g);
  } else if (main == 5) {
    std::memset(g, 0, bps);
  } else {
    // 4, 6, 7: one genotype everywhere, then the differences
    const uint8_t fill = main == 4 ? 0x00 : main == 6 ? 0xaa : 0xff;
    std::memset(g, fill, bps);
    apply_difflist(g);
  }
}

void PgenSource::decode_rows(uint64_t j0, uint64_t n, const int* order, uint8_t* out) const {
  const uint64_t bps = bytes_per_snp(n_);
  const uint32_t tail = (uint32_t)(n_ & 3);
  const uint8_t tail_mask = tail ? (uint8_t)((1u << (2 * tail)) - 1) : 0xff;
  std::vector<std::string> errors;
  std::atomic<bool> failed{false};
#pragma omp parallel
  {
    Scratch s;
    std::vector<uint8_t> g(bps);
#pragma omp for schedule(static)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
      if (failed.load(std::memory_order_relaxed)) continue;
      const uint64_t j = order ? (uint64_t)order[j0 + (uint64_t)r] : j0 + (uint64_t)r;
      uint8_t* row = out + (uint64_t)r * bps;
      try {
        require(j < m_, kErrArgument, "pgen: variant index out of range");
        if (plink1_) {
          const uint8_t* p;
          const uint8_t* e;
          record(j, &p, &e);
          std::memcpy(row, p, bps);
        } else {
          decode_one(j, g.data(), s);
          for (uint64_t i = 0; i < bps; ++i) row[i] = kToBed.t[g[i]];
        }
        row[bps - 1] &= tail_mask;
      } catch (const std::exception& ex) {
#pragma omp critical
        {
          failed = true;
          errors.emplace_back(ex.what());
        }
      }
    }
  }
  if (failed) fail(kErrIO, errors.front());
}

std::vector<int> pgen_perm_order(uint64_t nsnps, uint32_t nbatches, uint64_t blocksize, uint32_t nthreads,
                                 int seed) {
  // compute_pgen_perm (src/FilePgen.cpp:91-151): every block of a batch is
  // split over the reader threads; thread t reads, over all blocks, a
  // contiguous partition of the file's variants, shuffled once (one engine,
  // seeded with --seed, partitions in thread order).
  if (nsnps == 0) return {};
  require(nsnps < (1ull << 31), kErrArgument, "pgen permutation: too many SNPs");
  nbatches = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nbatches, nsnps));
  blocksize = std::max<uint64_t>(1, blocksize);
  nthreads = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nthreads, std::min<uint64_t>(blocksize, nsnps)));
  auto for_each_block = [&](auto&& body) {
    for (uint64_t b = 0; b < nbatches; ++b) {
      const uint64_t bs = b * nsnps / nbatches, be = (b + 1) * nsnps / nbatches;
      for (uint64_t s0 = bs; s0 < be; s0 += blocksize) {
        const uint64_t len = std::min(blocksize, be - s0);
        const uint64_t active = std::min<uint64_t>(nthreads, len);
        body(s0, len, active, len / active, len % active);
      }
    }
  };
  std::vector<uint64_t> count(nthreads, 0);
  for_each_block([&](uint64_t, uint64_t, uint64_t active, uint64_t base, uint64_t extra) {
    for (uint64_t t = 0; t < active; ++t) count[t] += base + (t < extra ? 1 : 0);
  });
  std::vector<std::vector<int>> src(nthreads);
  uint64_t s = 0;
  for (uint32_t t = 0; t < nthreads; ++t) {
    src[t].resize(count[t]);
    std::iota(src[t].begin(), src[t].end(), (int)s);
    s += count[t];
  }
  std::default_random_engine rng;
  rng.seed(seed);
  for (auto& v : src) std::shuffle(v.begin(), v.end(), rng);
  std::vector<uint64_t> next(nthreads, 0);
  std::vector<int> order(nsnps, -1);
  for_each_block([&](uint64_t s0, uint64_t, uint64_t active, uint64_t base, uint64_t extra) {
    uint64_t off = 0;
    for (uint64_t t = 0; t < active; ++t) {
      const uint64_t len = base + (t < extra ? 1 : 0);
      for (uint64_t q = 0; q < len; ++q) {
        require(next[t] < src[t].size(), kErrInternal, "BUG: exhausted PGEN permutation partition.");
        order[s0 + off + q] = src[t][next[t]++];
      }
      off += len;
    }
  });
  return order;
}

}  // namespace pcaone
