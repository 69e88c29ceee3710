#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "synth.hpp"

namespace pcaone {
namespace synth {

struct Spec {
  uint64_t seed;
  uint32_t nsamples;
  uint32_t npops;
  uint64_t nsnps;
  double freq_lo, freq_hi;  // ancestral A1 frequency ~ U(freq_lo, freq_hi)
  double fst_lo, fst_hi;    // per-population Fst ~ U(fst_lo, fst_hi)
  double miss_rate;         // missing completely at random, every genotype
  double high_miss_frac;    // fraction of samples flagged for extra missingness
  double high_miss_rate;    // missing rate applied on top for flagged samples
};

void validate(const Spec& s);
std::vector<double> pop_fst(const Spec& s);
std::vector<uint8_t> sample_flags(const Spec& s);
MissSpec miss_spec(const Spec& s);
uint32_t pop_of(const Spec& s, uint64_t i);
void snp_thresholds(const Spec& s, const std::vector<double>& fst, uint64_t j, uint32_t attempt, uint64_t* thr);
bool snp_row(const Spec& s, const MissSpec& ms, const std::vector<uint8_t>& flags, uint64_t j, uint32_t attempt,
             const uint64_t* thr, uint8_t* row);
void generate(const Spec& s, uint8_t* packed, uint32_t* attempts);
void write_plink(const Spec& s, const std::string& prefix);

}  // namespace synth
}  // namespace pcaone
