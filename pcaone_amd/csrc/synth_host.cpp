// Host side of the synthetic Balding-Nichols PLINK generator.
//
// Model: ancestral A1 frequency p ~ U(freq_lo, freq_hi) per SNP; population q
// has F_q ~ U(fst_lo, fst_hi) and frequency p_q ~ Beta(p(1-F_q)/F_q,
// (1-p)(1-F_q)/F_q); genotype ~ Binomial(2, p_q) with samples split into
// contiguous, equally sized populations.  A SNP whose observed genotypes are
// all identical is redrawn with the next attempt counter (up to 64 times).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "common.hpp"
#include "synth.hpp"
#include "synth_host.hpp"

namespace pcaone {
namespace synth {

namespace {

// Sequential SplitMix64 stream used for the variable-length gamma draws.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0,1)
  double normal() {                                                       // Box-Muller
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Marsaglia & Tsang (2000), "A simple method for generating gamma variables".
  double gamma(double a) {
    if (a < 1.0) {
      const double g = gamma(a + 1.0);
      return g * std::pow(uniform(), 1.0 / a);
    }
    const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x, v;
      do {
        x = normal();
        v = 1.0 + c * x;
      } while (v <= 0.0);
      v = v * v * v;
      const double u = uniform();
      if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
  double beta(double a, double b) {
    const double x = gamma(a), y = gamma(b);
    return x / (x + y);
  }
};

inline double unit(uint64_t h) { return ((double)(h >> 11) + 0.5) * 0x1.0p-53; }

inline uint64_t to_thr(double p) {
  if (!(p > 0.0)) return 0;
  if (p >= 1.0) return 1ULL << 32;
  return (uint64_t)std::llround(p * 4294967296.0);
}

}  // namespace

void validate(const Spec& s) {
  require(s.nsamples > 0 && s.nsnps > 0 && s.npops > 0, kErrArgument, "synth: empty shape");
  require(s.npops <= s.nsamples, kErrArgument, "synth: more populations than samples");
  require(s.nsnps < (1ULL << 32), kErrArgument, "synth: nsnps must be < 2^32");
  require(s.freq_lo >= 0 && s.freq_hi <= 1 && s.freq_lo <= s.freq_hi, kErrArgument, "synth: bad frequency range");
  require(s.fst_lo >= 0 && s.fst_hi < 1 && s.fst_lo <= s.fst_hi, kErrArgument, "synth: bad Fst range");
  require(s.miss_rate >= 0 && s.miss_rate <= 1 && s.high_miss_rate >= 0 && s.high_miss_rate <= 1 &&
              s.high_miss_frac >= 0 && s.high_miss_frac <= 1,
          kErrArgument, "synth: bad missingness");
}

std::vector<double> pop_fst(const Spec& s) {
  std::vector<double> f(s.npops);
  for (uint32_t q = 0; q < s.npops; ++q) f[q] = s.fst_lo + (s.fst_hi - s.fst_lo) * unit(hash3(s.seed, kStreamPop, q));
  return f;
}

std::vector<uint8_t> sample_flags(const Spec& s) {
  std::vector<uint8_t> fl(s.nsamples, 0);
  if (s.high_miss_frac > 0)
    for (uint64_t i = 0; i < s.nsamples; ++i) fl[i] = unit(hash3(s.seed, kStreamSample, i)) < s.high_miss_frac;
  return fl;
}

MissSpec miss_spec(const Spec& s) { return MissSpec{to_thr(s.miss_rate), to_thr(s.high_miss_rate)}; }

void snp_thresholds(const Spec& s, const std::vector<double>& fst, uint64_t j, uint32_t attempt, uint64_t* thr) {
  SplitMix rng{hash3(s.seed, kStreamSnp + attempt, j)};
  const double p = s.freq_lo + (s.freq_hi - s.freq_lo) * rng.uniform();
  for (uint32_t q = 0; q < s.npops; ++q) {
    double pq = p;
    if (fst[q] > 0) {
      const double f = fst[q];
      pq = rng.beta(p * (1 - f) / f, (1 - p) * (1 - f) / f);
    }
    thr[q] = to_thr(pq);
  }
}

uint32_t pop_of(const Spec& s, uint64_t i) { return (uint32_t)(i * s.npops / s.nsamples); }

// Fills one SNP row (bytes_per_snp(N) bytes; padding bits of the last byte are
// zero as PLINK writes them).  Returns true if the SNP is polymorphic among
// its observed genotypes.
bool snp_row(const Spec& s, const MissSpec& ms, const std::vector<uint8_t>& flags, uint64_t j, uint32_t attempt,
             const uint64_t* thr, uint8_t* row) {
  const uint64_t bps = bytes_per_snp(s.nsamples);
  std::memset(row, 0, bps);
  uint32_t seen = 0;  // bit0: 0 copies, bit1: 1 copy, bit2: 2 copies observed
  for (uint64_t i = 0; i < s.nsamples; ++i) {
    const uint32_t code = genotype_code(s.seed, attempt, j, i, thr[pop_of(s, i)], ms, flags[i] != 0);
    row[i >> 2] |= (uint8_t)(code << (2 * (i & 3)));
    if (code == 3u) seen |= 1u;
    else if (code == 2u) seen |= 2u;
    else if (code == 0u) seen |= 4u;
  }
  return seen != 0 && seen != 1u && seen != 4u;
}

void generate(const Spec& s, uint8_t* packed, uint32_t* attempts) {
  validate(s);
  const auto fst = pop_fst(s);
  const auto flags = sample_flags(s);
  const MissSpec ms = miss_spec(s);
  const uint64_t bps = bytes_per_snp(s.nsamples);
  bool bad = false;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t jj = 0; jj < (int64_t)s.nsnps; ++jj) {
    const uint64_t j = (uint64_t)jj;
    std::vector<uint64_t> thr(s.npops);
    uint32_t t = 0;
    for (;; ++t) {
      snp_thresholds(s, fst, j, t, thr.data());
      if (snp_row(s, ms, flags, j, t, thr.data(), packed + j * bps)) break;
      if (t == 63) {
#pragma omp atomic write
        bad = true;
        break;
      }
    }
    if (attempts) attempts[j] = t;
  }
  require(!bad, kErrNumeric, "synth: a SNP stayed monomorphic after 64 attempts");
}

void write_plink(const Spec& s, const std::string& prefix) {
  const uint64_t bps = bytes_per_snp(s.nsamples);
  std::vector<uint8_t> packed(bps * s.nsnps);
  generate(s, packed.data(), nullptr);
  {
    std::ofstream bed(prefix + ".bed", std::ios::binary);
    require(bed.good(), kErrIO, "synth: cannot write " + prefix + ".bed");
    const uint8_t magic[3] = {0x6c, 0x1b, 0x01};
    bed.write((const char*)magic, 3);
    bed.write((const char*)packed.data(), (std::streamsize)packed.size());
    require(bed.good(), kErrIO, "synth: short write on " + prefix + ".bed");
  }
  {
    FILE* f = std::fopen((prefix + ".bim").c_str(), "w");
    require(f != nullptr, kErrIO, "synth: cannot write " + prefix + ".bim");
    for (uint64_t j = 0; j < s.nsnps; ++j)
      std::fprintf(f, "1\tsnp%llu\t0\t%llu\tA\tC\n", (unsigned long long)j, (unsigned long long)(j + 1));
    std::fclose(f);
  }
  {
    FILE* f = std::fopen((prefix + ".fam").c_str(), "w");
    require(f != nullptr, kErrIO, "synth: cannot write " + prefix + ".fam");
    for (uint64_t i = 0; i < s.nsamples; ++i)
      std::fprintf(f, "pop%u\tind%llu\t0\t0\t0\t-9\n", pop_of(s, i), (unsigned long long)i);
    std::fclose(f);
  }
}

}  // namespace synth
}  // namespace pcaone
