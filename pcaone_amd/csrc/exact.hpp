// --svd 3 (exact PCA) and --svd 0 (IRAM) drivers on the resident genotypes.
#pragma once
#include "context.hpp"
#include "rsvd.hpp"

namespace pcaone {

constexpr int PCAONE_PREC_FP64_ID = 0;  // PCAONE_PREC_FP64

// src/Main.cpp:180-217: U N x k, S = sqrt(M evals), E = evals, V M x k,
// flip_UV with U as the base.
PcaResult run_pca_full(Context& ctx, const Param& params, RunInfo* info);
// src/Arnoldi.cpp:48-118 (in-core, no EMU): U, S, E = S^2 / M, V.
PcaResult run_pca_arnoldi(Context& ctx, const Param& params, RunInfo* info);

}  // namespace pcaone
