// Deterministic Balding-Nichols genotype generator shared by the host writer
// (synth_host.cpp) and the device kernel (synth_kernel.hip).
//
// Every random quantity is a pure function of (seed, stream, counter) through
// a 64-bit finalizer hash, so the host and the GPU produce byte-identical
// PLINK payloads.  Per-SNP population frequencies (Beta draws, which need
// libm) are computed on the host only and handed to the device as 2^32-scaled
// integer thresholds; the per-genotype work is integer-only.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define PCAONE_HD __host__ __device__
#else
#define PCAONE_HD
#endif

namespace pcaone {
namespace synth {

enum : uint64_t {
  kStreamSnp = 0x100,      // + attempt: per-SNP ancestral/population frequencies
  kStreamGeno = 0x10000,   // + attempt: the two allele draws of one genotype
  kStreamMiss = 0x20000,   // missing-completely-at-random draw of one genotype
  kStreamSample = 0x30000, // per-sample "high missingness" selection
  kStreamPop = 0x40000,    // per-population Fst
};

PCAONE_HD inline uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

PCAONE_HD inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return fmix64(fmix64(seed * 0x9E3779B97F4A7C15ULL + stream) ^ (idx * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
}

// PLINK 2-bit code for a genotype carrying `c` copies of allele A1:
// 2 copies -> 0b00, 1 copy -> 0b10, 0 copies -> 0b11; missing -> 0b01.
PCAONE_HD inline uint32_t code_of_count(uint32_t c) { return c == 2u ? 0u : (c == 1u ? 2u : 3u); }

struct MissSpec {
  uint64_t mcar_thr;   // P(missing) * 2^32 for every genotype (0 = none)
  uint64_t high_thr;   // P(missing) * 2^32 for genotypes of flagged samples
};

// Code of sample i at SNP j.  `thr` = 2^32-scaled A1 frequency of the
// sample's population; `flagged` = sample selected for high missingness.
PCAONE_HD inline uint32_t genotype_code(uint64_t seed, uint32_t attempt, uint64_t j, uint64_t i, uint64_t thr,
                                        const MissSpec& ms, bool flagged) {
  const uint64_t key = (j << 32) | i;
  const uint64_t h = hash3(seed, kStreamGeno + attempt, key);
  const uint32_t c = (uint32_t)((uint64_t)(uint32_t)h < thr) + (uint32_t)((h >> 32) < thr);
  uint32_t code = code_of_count(c);
  if (ms.mcar_thr | ms.high_thr) {
    const uint64_t hm = hash3(seed, kStreamMiss, key);
    if ((uint64_t)(uint32_t)hm < ms.mcar_thr || (flagged && (hm >> 32) < ms.high_thr)) code = 1u;
  }
  return code;
}

}  // namespace synth
}  // namespace pcaone
