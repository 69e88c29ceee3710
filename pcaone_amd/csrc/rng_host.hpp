#pragma once
#include <cstdint>

namespace pcaone {
namespace rng {

// The reference's Omega (nrows x ncols, column-major): libstdc++
// default_random_engine seeded with `seed`, normal{0,1} if gaussian else
// uniform{-1,1}, drawn in parallel with identical values.
void omega(int seed, bool gaussian, uint64_t nrows, uint64_t ncols, double* out);

}  // namespace rng
}  // namespace pcaone
