// PLINK 2 .pgen input (the hardcall path of src/FilePgen.cpp:27-287).
//
// The records are decoded on the host into the PLINK 1 2-bit codes that the
// device path takes from a .bed (code 00 for two ALT copies, 10 for one, 11
// for none, 01 missing), so every kernel, the allele frequencies and the
// centering are shared with the .bed input.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace pcaone {

class PgenSource {
 public:
  // Opens <prefix>.pgen (and <prefix>.pgen.pgi for the index-in-a-side-file
  // modes) and reads its variant index; nsamples comes from the .psam.
  // hardcall: read the hardcall track of files that also carry dosages (the
  // reference's --hardcall); without it such files are rejected.
  PgenSource(const std::string& pgen, uint64_t nsamples, bool hardcall);
  ~PgenSource();
  PgenSource(const PgenSource&) = delete;
  PgenSource& operator=(const PgenSource&) = delete;

  uint64_t nsamples() const { return n_; }
  uint64_t nsnps() const { return m_; }
  bool dosage_present() const { return dosage_; }
  int mode() const { return mode_; }  // third byte of the file

  // Rows j0..j0+n of the file (or rows order[j0 + r] when order is given) as
  // .bed rows of ceil(N/4) bytes each, trailing bits zero.  Parallel over
  // rows (OpenMP).
  void decode_rows(uint64_t j0, uint64_t n, const int* order, uint8_t* out) const;

 private:
  struct Scratch;
  void decode_one(uint64_t j, uint8_t* geno, Scratch& s) const;  // plink2 2-bit genotypes of variant j
  void record(uint64_t j, const uint8_t** p, const uint8_t** end) const;
  uint64_t ldbase_of(uint64_t j) const;

  uint64_t n_ = 0, m_ = 0;
  int mode_ = 0;
  bool dosage_ = false;
  void* map_ = nullptr;
  size_t map_len_ = 0;
  const uint8_t* base_ = nullptr;
  // fixed-width modes: one record width and vrtype; variable-width: per variant
  uint64_t const_width_ = 0, const_offset_ = 0;
  uint32_t const_vrtype_ = 0;
  bool plink1_ = false;
  std::vector<uint8_t> vrtype_;
  std::vector<uint64_t> fpos_;  // M + 1 record offsets
};

// .psam sample count (data lines: not starting with '#'), as FilePgen's
// constructor counts them (src/FilePgen.hpp:18-25).
uint64_t psam_samples(const std::string& psam);

// compute_pgen_perm (src/FilePgen.cpp:91-151): the logical SNP order of an
// out-of-core winSVD run on a .pgen (no file is rewritten).  Position i of the
// run reads file variant order[i].
std::vector<int> pgen_perm_order(uint64_t nsnps, uint32_t nbatches, uint64_t blocksize, uint32_t nthreads,
                                 int seed);

}  // namespace pcaone
