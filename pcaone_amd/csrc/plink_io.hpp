#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "context.hpp"
#include "rsvd.hpp"

namespace pcaone {

struct PlinkInfo {
  uint64_t nsamples = 0, nsnps_file = 0;
  std::vector<uint32_t> keep;  // kept SNP indices when --maf filtered
};

// FileBed constructor: dimensions and the .bed magic (no device work).
PlinkInfo plink_dims(const std::string& prefix);
PlinkInfo read_plink(Context& ctx, const std::string& prefix, const Param& params);
// Data::write_residuals (src/Data.cpp:242-291): <out>.residuals.
void write_residuals(Context& ctx, const PcaResult& r, const Param& params, const std::string& out);
void write_outputs(const Context& ctx, const PcaResult& r, const Param& params, const std::string& bfile,
                   const std::string& out, const PlinkInfo& info);

}  // namespace pcaone
