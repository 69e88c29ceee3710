// Reference-side binding of libpcaone_amd: what a PCAone maintainer adds to
// the reference tree (next to src/Halko.hpp) to run its randomized SVD on an
// MI355X.  Header-only; it includes the reference's own headers, so it is
// compiled inside the reference's build (g++ -std=c++17, Eigen) and links
// against libpcaone_amd.so.
//
// Level 1: run_pca_with_halko_amd replaces the PLINK branch of main
//   (src/Main.cpp:150-231): read, randomized SVD and output files on the device.
// Level 2: AmdNormalRsvdOpData / AmdFancyRsvdOpData override computeGandH
//   (src/Halko.cpp:99-224) and keep the reference's computeUSV
//   (src/Halko.cpp:46-97) with its host Householder QR and JacobiSVD;
//   AmdFileBed replaces FileBed (src/FilePlink.hpp:10-30) and AmdFilePgen
//   replaces FilePgen (src/FilePgen.hpp:14-48) as the Data provider.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "Halko.hpp"
#include "RSVD.hpp"
#include "Utils.hpp"
#include "pcaone_amd.h"

// A failed call becomes the reference's own error path (src/Logger.hpp:85-94).
#define AMD_CHECK(ctx, call)                                         \
  do {                                                               \
    if ((call) != PCAONE_OK) cao.error(pcaone_last_error(ctx));      \
  } while (0)

namespace pcaone_amd_binding {

inline pcaone_params to_amd(const Param& p, int precision = PCAONE_PREC_BF16) {
  pcaone_params a;
  pcaone_params_default(&a);
  a.svd = p.svd_t == SvdType::PCAoneAlg1 ? 1 : 2;  // src/Cmd.hpp:14
  a.k = p.k;
  a.oversamples = p.oversamples;
  a.maxp = p.maxp;
  a.bands = p.bands;
  a.maxiter = p.maxiter;
  a.tol = p.tol;
  a.tolem = p.tolem;
  a.maf = p.maf;
  a.seed = p.seed;
  a.rand = (int)p.rand;
  a.ploidy = p.ploidy;
  a.noshuffle = p.noshuffle;
  a.emu = p.emu;
  a.printv = p.printv;
  a.verbose = (int)p.verbose;
  a.memory = p.memory;
  a.buffer = p.buffer;
  a.hardcall = p.hardcall;
  a.threads = p.threads;
  a.precision = precision;
  return a;
}

// Level 1: replaces FileBed/FilePgen::read_all + run_pca_with_halko +
// write_eigs_files (+ make_plink2_eigenvec_file / _from_psam) for PLINK or
// PLINK 2 input with --svd 1/2.
inline void run_pca_with_halko_amd(const Param& params, int precision = PCAONE_PREC_BF16) {
  pcaone_ctx* ctx = nullptr;
  if (pcaone_open(0, &ctx) != PCAONE_OK) cao.error(pcaone_last_error(nullptr));
  const pcaone_params a = to_amd(params, precision);
  pcaone_run_info info;
  if (params.file_t == FileType::PGEN)
    AMD_CHECK(ctx, pcaone_run_pgen(ctx, params.filein.c_str(), params.fileout.c_str(), &a, &info));
  else
    AMD_CHECK(ctx, pcaone_run_bfile(ctx, params.filein.c_str(), params.fileout.c_str(), &a, &info));
  cao.print(tick.date(), "device time of the products:", info.g_ms + info.h_ms, "ms");
  pcaone_close(ctx);
}

// The Data provider: the .bed payload goes to HBM undecoded; F, p_miss and
// the counts are filled for the rest of the reference.
class AmdFileBed : public Data {
 public:
  AmdFileBed(const Param& p, pcaone_ctx* c) : Data(p), ctx(c) {
    uint32_t n = 0;
    uint64_t m = 0;
    AMD_CHECK(nullptr, pcaone_bed_dims(params.filein.c_str(), &n, &m));  // FileBed's line counts
    nsamples = n;
    nsnps = (uint)m;
  }
  void read_all() override {
    uint32_t n = 0;
    uint64_t m = 0, miss = 0;
    AMD_CHECK(ctx, pcaone_bed_read(ctx, params.filein.c_str(), params.maf, &n, &m));
    nsamples = n;
    nsnps = (uint)m;  // after --maf (src/Data.cpp:88-105)
    F.resize(nsnps);
    AMD_CHECK(ctx, pcaone_allele_freq(ctx, F.data()));
    AMD_CHECK(ctx, pcaone_dims(ctx, &n, &m, &miss));
    p_miss = (double)miss / ((double)n * (double)m);
  }
  void check_file_offset_first_var() override {}
  // With -m the whole payload is resident in HBM: blocks need no reading.
  void read_block_initial(uint64, uint64, bool) override {}
  void read_block_update(uint64, uint64, const Mat2D&, const Mat1D&, const Mat2D&, bool) override {}
  pcaone_ctx* ctx;
};

// The PLINK 2 Data provider: the .pgen records are decoded on the host into
// .bed rows and sent to HBM; the rest as AmdFileBed.
class AmdFilePgen : public Data {
 public:
  AmdFilePgen(const Param& p, pcaone_ctx* c) : Data(p), ctx(c) {
    uint32_t n = 0;
    uint64_t m = 0;
    AMD_CHECK(nullptr, pcaone_pgen_dims(params.filein.c_str(), &n, &m));  // .psam lines, .pgen header
    nsamples = n;
    nsnps = (uint)m;
  }
  void read_all() override {
    uint32_t n = 0;
    uint64_t m = 0, miss = 0;
    AMD_CHECK(ctx, pcaone_pgen_read(ctx, params.filein.c_str(), params.maf, params.hardcall, &n, &m));
    nsamples = n;
    nsnps = (uint)m;
    F.resize(nsnps);
    AMD_CHECK(ctx, pcaone_allele_freq(ctx, F.data()));
    AMD_CHECK(ctx, pcaone_dims(ctx, &n, &m, &miss));
    p_miss = (double)miss / ((double)n * (double)m);
  }
  void check_file_offset_first_var() override {}
  void read_block_initial(uint64, uint64, bool) override {}
  void read_block_update(uint64, uint64, const Mat2D&, const Mat1D&, const Mat2D&, bool) override {}
  pcaone_ctx* ctx;
};

// The data-side effects of computeGandH at pi == 0 (src/Halko.cpp:107-118,
// 167-186): imputation from the current U, S, V, then standardize_E.
inline void amd_epoch0_side_effects(RsvdOpData& op, pcaone_ctx* ctx) {
  if (op.update) AMD_CHECK(ctx, pcaone_set_impute(ctx, op.U.data(), op.S.data(), op.V.data(), (uint32_t)op.S.size()));
  if (op.standardize) AMD_CHECK(ctx, pcaone_standardize(ctx, op.data->params.ploidy));
}

// --svd 1 (src/Halko.cpp:99-153).  G and H cross back to the host each pass.
class AmdNormalRsvdOpData : public NormalRsvdOpData {
 public:
  AmdNormalRsvdOpData(Data* d, int k, int os, pcaone_ctx* c, int precision = PCAONE_PREC_BF16)
      : NormalRsvdOpData(d, k, os), ctx(c) {
    AMD_CHECK(ctx, pcaone_configure(ctx, precision, (uint32_t)(ranks() + oversamples())));
  }
  void computeGandH(Mat2D& G, Mat2D& H, int pi) override {
    if (pi == 0) {
      initOmg();
      amd_epoch0_side_effects(*this, ctx);
      AMD_CHECK(ctx, pcaone_set_omega(ctx, Omg.data()));  // Omg = Omg2
    } else if (!data->params.out_of_core) {
      AMD_CHECK(ctx, pcaone_update_omega(ctx, 0));  // Omg = thin-Q(H), flipOmg(Omg2, Omg)
    } else {
      AMD_CHECK(ctx, pcaone_thin_q_omega(ctx, 0));  // src/Halko.cpp:133-136: no flipOmg
    }
    AMD_CHECK(ctx, pcaone_sketch(ctx, 0, (uint64_t)rows(), 0));  // G = A Omg, H = A^T G
    AMD_CHECK(ctx, pcaone_get_g(ctx, 0, (uint64_t)rows(), G.data()));
    AMD_CHECK(ctx, pcaone_get_h(ctx, 0, H.data()));
  }
  pcaone_ctx* ctx;
};

// --svd 2, in-core (src/Halko.cpp:155-224): the same window schedule, with H1
// and H2 kept on the device.
class AmdFancyRsvdOpData : public FancyRsvdOpData {
 public:
  AmdFancyRsvdOpData(Data* d, int k, int os, pcaone_ctx* c, int precision = PCAONE_PREC_BF16)
      : FancyRsvdOpData(d, k, os), ctx(c) {
    AMD_CHECK(ctx, pcaone_configure(ctx, precision, (uint32_t)(ranks() + oversamples())));
  }
  void computeGandH(Mat2D& G, Mat2D& H, int pi) override {
    if (data->params.out_of_core) cao.error("AmdFancyRsvdOpData: use run_pca_with_halko_amd for -m");
    const uint64 bands = data->params.bands, M = data->nsnps;
    if (pi == 0) initOmg();
    if (std::pow(2, pi) >= bands) {
      AMD_CHECK(ctx, pcaone_zero_h(ctx, 1));
      AMD_CHECK(ctx, pcaone_zero_h(ctx, 2));
    }
    if (pi == 0) {
      amd_epoch0_side_effects(*this, ctx);
      AMD_CHECK(ctx, pcaone_set_omega(ctx, Omg.data()));
      bandsize_ = 1;
      blocksize_ = (uint64)std::ceil((double)M / bands);
      if (data->params.perm && !(update && data->params.emu)) {
        Mat2D cols(1, M);  // the reference's own shuffle (src/RSVD.hpp:61-78) decides the order
        PCAone::permute_matrix(cols, data->perm);
        std::vector<int32_t> order(data->perm.indices().data(), data->perm.indices().data() + M);
        AMD_CHECK(ctx, pcaone_permute_snps(ctx, order.data()));
      }
    }
    bandsize_ = std::min<uint64>(bandsize_ * 2, bands);
    for (uint64 b = 0, i = 1; b < bands; ++b, ++i) {
      const uint64 start = b * blocksize_;
      const uint64 stop = (b + 1) * blocksize_ >= M ? M - 1 : (b + 1) * blocksize_ - 1;
      if (start <= stop) AMD_CHECK(ctx, pcaone_sketch(ctx, start, stop + 1, i <= bandsize_ / 2 ? 1 : 2));
      const bool adjacent = (pi > 0 && (b + 1) == std::pow(2, pi - 1) && std::pow(2, pi) < bands);
      if ((b + 1) < bandsize_ && !adjacent) continue;
      if (!((i == bandsize_) || (i == bandsize_ / 2) || adjacent)) continue;
      AMD_CHECK(ctx, pcaone_update_omega(ctx, 1));  // H = H1 + H2, Omg = thin-Q(H), flipOmg
      if (i == bandsize_) {
        AMD_CHECK(ctx, pcaone_zero_h(ctx, 1));
        i = 0;
      } else {
        AMD_CHECK(ctx, pcaone_zero_h(ctx, 2));
      }
    }
    AMD_CHECK(ctx, pcaone_get_g(ctx, 0, M, G.data()));
    AMD_CHECK(ctx, pcaone_get_h(ctx, 0, H.data()));
  }
  pcaone_ctx* ctx;

 private:
  uint64 bandsize_ = 1, blocksize_ = 1;
};

// run_pca_with_halko (src/Halko.cpp:271-345) for genetic input with the
// level-2 operators: the reference's computeUSV, EM loop and output files.
inline void run_pca_with_halko_level2(Data* data, const Param& params, pcaone_ctx* ctx,
                                      int precision = PCAONE_PREC_BF16) {
  RsvdOpData* rsvd;
  if (params.svd_t == SvdType::PCAoneAlg2)
    rsvd = new AmdFancyRsvdOpData(data, params.k, params.oversamples, ctx, precision);
  else
    rsvd = new AmdNormalRsvdOpData(data, params.k, params.oversamples, ctx, precision);
  if (!params.missme) {
    rsvd->setFlags(false, true);
    rsvd->computeUSV(params.maxp, params.tol);
  } else {
    rsvd->setFlags(false, false);
    rsvd->computeUSV(params.maxp, params.tol);
    flip_UV(rsvd->U, rsvd->V, false);
    for (uint i = 0; i < params.maxiter; ++i) {
      rsvd->setFlags(true, false);
      const Mat2D Vpre = rsvd->V;
      rsvd->computeUSV(params.maxp, params.tol);
      flip_UV(rsvd->U, rsvd->V, false);
      const double diff = 1.0 - mev(rsvd->V, Vpre);
      cao.print(tick.date(), "individual allele frequencies estimated iter =", i + 1, ", diff =", diff);
      if (diff < params.tolem) break;
    }
    if (params.emu) {
      rsvd->setFlags(true, true);
      rsvd->computeUSV(params.maxp, params.tol);
      flip_UV(rsvd->U, rsvd->V, false);
    }
  }
  data->write_eigs_files(rsvd->S.array().square() / data->nsnps, rsvd->S, rsvd->U, rsvd->V);
  delete rsvd;
}

}  // namespace pcaone_amd_binding
