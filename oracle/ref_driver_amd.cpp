// Test-infrastructure only: the reference's own PLINK + randomized-SVD path
// with libpcaone_amd bound in through pcaone_amd/binding/
// pcaone_reference_binding.hpp, as a PCAone maintainer would bind it.  It
// shows on a GPU that the binding compiles against the reference's headers,
// links with its objects and gives the reference's results.
//
//   PCAONE_AMD_LEVEL=1  run_pca_with_halko_amd (the whole run on the device)
//   PCAONE_AMD_LEVEL=2  (default) AmdFileBed + AmdNormal/AmdFancyRsvdOpData
//                       under the reference's own computeUSV, EM loop and
//                       write_eigs_files (src/Halko.cpp:46-97, 271-345)
//   PCAONE_AMD_PRECISION=<PCAONE_PREC_*> (default 0, FP64)
//
// Dispatch as src/Main.cpp:48-58, :168-179, :222-231, for -b or -p input,
// in-core --svd 1/2 (level 1 also takes -m).
#define _DECLARE_TOOLBOX_HERE

#include <cstdlib>
#include <string>

#include "Cmd.hpp"
#include "Data.hpp"
#include "Halko.hpp"
#include "Utils.hpp"
#include "pcaone_reference_binding.hpp"

int main(int argc, char* argv[]) {
  Param params(argc, argv);
  cao.cao.open(params.fileout + ".log");
  cao.is_screen = params.verbose > 0;
  cao.print(params.ss.str());
  if (params.file_t != FileType::PLINK && params.file_t != FileType::PGEN)
    cao.error("this driver only reads PLINK (-b) or PLINK 2 (-p) input");
  if (params.svd_t != SvdType::PCAoneAlg1 && params.svd_t != SvdType::PCAoneAlg2)
    cao.error("this driver only runs --svd 1 or --svd 2");
  const char* lv = std::getenv("PCAONE_AMD_LEVEL");
  const char* pr = std::getenv("PCAONE_AMD_PRECISION");
  const int level = lv ? std::atoi(lv) : 2;
  const int precision = pr ? std::atoi(pr) : PCAONE_PREC_FP64;
  cao.print(tick.date(), "program started (libpcaone_amd binding, level", level, ")");
  if (level == 1) {
    pcaone_amd_binding::run_pca_with_halko_amd(params, precision);
  } else {
    if (params.out_of_core) cao.error("level 2 of this driver runs in-core only");
    pcaone_ctx* ctx = nullptr;
    if (pcaone_open(0, &ctx) != PCAONE_OK) cao.error(pcaone_last_error(nullptr));
    const bool pgen = params.file_t == FileType::PGEN;
    Data* data = pgen ? (Data*)new pcaone_amd_binding::AmdFilePgen(params, ctx)
                      : (Data*)new pcaone_amd_binding::AmdFileBed(params, ctx);
    data->prepare();
    pcaone_amd_binding::run_pca_with_halko_level2(data, params, ctx, precision);
    delete data;
    pcaone_close(ctx);
    if (pgen)
      make_plink2_eigenvec_from_psam(params.k, params.fileout + ".eigvecs2", params.fileout + ".eigvecs",
                                     params.filein + ".psam");
    else
      make_plink2_eigenvec_file(params.k, params.fileout + ".eigvecs2", params.fileout + ".eigvecs",
                                params.filein + ".fam");
  }
  cao.print(tick.date(), "total elapsed wall time:", tick.abstime(), " seconds");
  return 0;
}
