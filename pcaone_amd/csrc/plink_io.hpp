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
  bool pgen = false;           // a .pgen/.pvar/.psam set (outputs from .pvar and .psam)
};

// FileBed constructor: dimensions and the .bed magic (no device work).
PlinkInfo plink_dims(const std::string& prefix);
PlinkInfo read_plink(Context& ctx, const std::string& prefix, const Param& params);
// FilePgen::FilePgen + read_all (src/FilePgen.hpp:14-35, src/FilePgen.cpp:27-89):
// N from the .psam, M from the .pgen; the hardcalls decoded on the host into
// .bed rows and uploaded as read_plink uploads them.
PlinkInfo pgen_dims(const std::string& prefix);
PlinkInfo read_pgen(Context& ctx, const std::string& prefix, const Param& params);
// The .pgen opened for streamed -m runs (records decoded per block load).
PlinkInfo stream_pgen(Context& ctx, const std::string& prefix, const Param& params, uint64_t block_snps);
// Data::write_residuals (src/Data.cpp:242-291): <out>.residuals.
void write_residuals(Context& ctx, const PcaResult& r, const Param& params, const std::string& out);
void write_outputs(const Context& ctx, const PcaResult& r, const Param& params, const std::string& bfile,
                   const std::string& out, const PlinkInfo& info);

}  // namespace pcaone
