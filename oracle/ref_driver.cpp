// Test-infrastructure only: a small `main` that drives the reference's own
// PLINK + randomized-SVD path (FileBed, Data::prepare, run_pca_with_halko,
// make_plink2_eigenvec_file) so that path can be compiled from its few source
// files without the reference's other readers (PGEN/BGEN/BEAGLE/CSV/LD/...).
//
// It follows the dispatch of src/Main.cpp:48-58 (params, log, thread env),
// :120-151 (out-of-core .bed permutation for --svd 2, then FileBed), :168
// (prepare), :178-179 (run_pca_with_halko), :222-231 (read time, .eigvecs2)
// and :234-240 (temp-file removal), restricted to `-b` input with --svd 1/2.
// Nothing here is shipped or measured as the product; it produces golden
// fixtures and the CPU baseline binary under oracle/_ref/.
#define _DECLARE_TOOLBOX_HERE

#include <cstdlib>
#include <filesystem>
#include <string>
#include <thread>

#include "Cmd.hpp"
#include "Data.hpp"
#include "FilePlink.hpp"
#include "Halko.hpp"
#include "Utils.hpp"

int main(int argc, char* argv[]) {
  Param params(argc, argv);
  cao.cao.open(params.fileout + ".log");
  cao.is_screen = params.verbose > 0;
  cao.print(params.ss.str());
  if (params.file_t != FileType::PLINK) cao.error("this driver only reads PLINK (-b) input");
  if (params.svd_t != SvdType::PCAoneAlg1 && params.svd_t != SvdType::PCAoneAlg2)
    cao.error("this driver only runs --svd 1 or --svd 2");
  cao.print(tick.date(), "program started");

  const bool ooc_permutation = params.perm && params.out_of_core;
  Data* data = nullptr;
  if (ooc_permutation) {
    tick.clock();
    PermMat perm = permute_plink(params.filein, params.fileout, params.buffer, params.bands);
    data = new FileBed(params);
    data->perm = perm;
    cao.print(tick.date(), "elapsed time of permuting data:", tick.reltime(), " seconds");
  } else {
    data = new FileBed(params);
  }
  data->prepare();
  run_pca_with_halko(data, params);
  cao.print(tick.date(), "total elapsed reading time: ", data->readtime, " seconds");
  delete data;
  make_plink2_eigenvec_file(params.k, params.fileout + ".eigvecs2", params.fileout + ".eigvecs",
                            params.filein + ".fam");
  if (ooc_permutation && params.verbose < 3) {
    for (const char* suf : {".bed", ".bim", ".fam"}) std::filesystem::remove(params.filein + suf);
  }
  cao.print(tick.date(), "total elapsed wall time:", tick.abstime(), " seconds");
  return 0;
}
