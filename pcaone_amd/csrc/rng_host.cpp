// Parallel, bit-identical reproduction of the reference's random test matrix.
//
// The reference fills Omega (N x L, column-major) with successive draws of
// std::normal_distribution<double>{0,1} (or uniform_real_distribution{-1,1})
// from std::default_random_engine seeded with --seed (src/Halko.cpp:15-23,
// src/RSVD.hpp:25-59).  With libstdc++ that engine is minstd_rand0
// (x <- 16807 x mod 2^31-1); generate_canonical<double> consumes exactly two
// engine outputs per uniform, and the normal distribution is the Marsaglia
// polar method, one attempt = two uniforms = four engine outputs, each
// accepted attempt giving y*mult then x*mult.  Because every attempt has a
// fixed width, attempt j starts at engine position 4j, which a modular power
// reaches directly; threads evaluate blocks of attempts and the accepted pairs
// are concatenated in attempt order.  The arithmetic is the same expression
// sequence as the library's, so the values are identical (checked against
// the library draw in the tests).
#include "rng_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace pcaone {
namespace rng {

namespace {
constexpr uint64_t kA = 16807, kM = 2147483647;  // minstd_rand0

uint64_t powmod(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= kM;
  while (e) {
    if (e & 1) r = r * b % kM;
    b = b * b % kM;
    e >>= 1;
  }
  return r;
}

// x * 16807 mod (2^31 - 1) without a division: 2^31 = 1 (mod 2^31 - 1).
inline uint64_t step(uint64_t x) {
  const uint64_t t = x * kA;
  uint64_t r = (t & kM) + (t >> 31);
  return r >= kM ? r - kM : r;
}

uint64_t seed_state(int seed) {
  // linear_congruential_engine::seed: s mod m, or 1 if that is 0
  const uint64_t s = (uint64_t)(unsigned long)seed % kM;
  return s == 0 ? 1 : s;
}

// generate_canonical<double, 53> for an engine with range [1, 2^31-2]:
// two draws, sum += (x - 1) * tmp with tmp = 1 then R, divided by R^2.
// The library computes tmp *= r in long double and stores it as double; the
// two values it takes are constants, formed the same way once.
const long double kR = (long double)(kM - 1) - 1.0L + 1.0L;  // max - min + 1 = 2^31 - 2
const double kTmp1 = (double)(1.0 * kR);
const double kTmp2 = (double)(kTmp1 * kR);

inline double canonical(uint64_t& x) {
  double sum = 0.0;
  x = step(x);
  sum += (double)(x - 1) * 1.0;
  x = step(x);
  sum += (double)(x - 1) * kTmp1;
  double ret = sum / kTmp2;
  if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
  return ret;
}
}  // namespace

void omega(int seed, bool gaussian, uint64_t nrows, uint64_t ncols, double* out) {
  const uint64_t total = nrows * ncols;
  const uint64_t x0 = seed_state(seed);
  if (!gaussian) {
    // uniform_real_distribution{-1,1}: u * (b - a) + a, one uniform per element
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)total; ++e) {
      uint64_t x = x0 * powmod(kA, 2 * (uint64_t)e) % kM;
      out[e] = canonical(x) * (1.0 - -1.0) + -1.0;
    }
    return;
  }
  // A round covers the attempts the remaining values need on average (the
  // polar method accepts pi/4 of them) plus 2%, split into 256 blocks.  Each
  // block evaluates its attempts once, keeping its values in a block-local
  // buffer; the buffers are then concatenated in block order at their prefix
  // offsets.
  const int nb = 256;
  std::vector<uint64_t> cnt(nb + 1);
  // block buffers of the calling thread, kept across calls (no page faults on
  // every draw); the reference makes the parallel region share this thread's copy
  static thread_local std::vector<std::vector<double>> tl_vals;
  std::vector<std::vector<double>>& vals = tl_vals;
  vals.resize(nb);
  uint64_t filled = 0, next = 0;  // next attempt index
  while (filled < total) {
    const uint64_t want = (uint64_t)((double)(total - filled) / 2.0 / 0.785398 * 1.02) + 64;
    const uint64_t B = (want + nb - 1) / nb;
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < nb; ++b) {
      uint64_t x = x0 * powmod(kA, 4 * (next + (uint64_t)b * B)) % kM;
      std::vector<double>& v = vals[b];
      v.resize(2 * B);
      uint64_t o = 0;
      for (uint64_t j = 0; j < B; ++j) {
        const double u1 = canonical(x);
        const double u2 = canonical(x);
        const double px = 2.0 * u1 - 1.0;
        const double py = 2.0 * u2 - 1.0;
        const double r2 = px * px + py * py;
        if (r2 > 1.0 || r2 == 0.0) continue;
        const double mult = std::sqrt(-2 * std::log(r2) / r2);
        v[o++] = py * mult * 1.0 + 0.0;  // ret * stddev + mean
        v[o++] = px * mult * 1.0 + 0.0;
      }
      cnt[b + 1] = o;
    }
    cnt[0] = 0;
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    const uint64_t room = total - filled;
    double* dst = out + filled;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b) {
      if (cnt[b] >= room) continue;
      const uint64_t n = std::min(cnt[b + 1], room) - cnt[b];
      std::copy(vals[b].begin(), vals[b].begin() + (int64_t)n, dst + cnt[b]);
    }
    filled += std::min(cnt[nb], room);
    next += (uint64_t)nb * B;
  }
}

}  // namespace rng
}  // namespace pcaone
