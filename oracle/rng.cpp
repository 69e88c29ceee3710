// Test-infrastructure only.  The reference draws its random test matrix and
// its in-core SNP shuffle from libstdc++ <random>; re-deriving those streams
// is not possible, so the oracle calls the same library routines here.
//   Omega:   src/Halko.cpp:15-23 + src/RSVD.hpp:25-59 (std::default_random_engine
//            seeded with --seed; normal_distribution<double>{0,1} or
//            uniform_real_distribution<double>{-1,1}; filled in column-major
//            order, one draw per element).
//   shuffle: src/RSVD.hpp:61-70 (std::shuffle of 0..M-1 with a default-seeded
//            std::default_random_engine).
//   parts:   src/FilePgen.cpp:124-126 (one engine seeded with --seed shuffles
//            consecutive partitions in turn).
#include <algorithm>
#include <numeric>
#include <random>

extern "C" {

void oracle_omega(int seed, int gaussian, long nrows, long ncols, double* out) {
  auto rng = std::default_random_engine{};
  rng.seed(seed);
  const long n = nrows * ncols;
  if (gaussian) {
    std::normal_distribution<double> d{0, 1};
    for (long i = 0; i < n; ++i) out[i] = d(rng);
  } else {
    std::uniform_real_distribution<double> d{-1, 1};
    for (long i = 0; i < n; ++i) out[i] = d(rng);
  }
}

void oracle_shuffle(long m, int* out) {
  std::iota(out, out + m, 0);
  auto rng = std::default_random_engine{};
  std::shuffle(out, out + m, rng);
}

void oracle_shuffle_parts(int seed, long nparts, const long* sizes, int* data) {
  auto rng = std::default_random_engine{};
  rng.seed(seed);
  for (long p = 0; p < nparts; ++p) {
    std::shuffle(data, data + sizes[p], rng);
    data += sizes[p];
  }
}

}  // extern "C"
