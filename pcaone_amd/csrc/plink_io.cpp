// PLINK input and PCAone output files.
//   read_plink        FileBed::FileBed + read_all (src/FilePlink.hpp:10-30,
//                     src/FilePlink.cpp:26-120): the .bed payload goes to HBM
//                     undecoded; allele frequencies are counted on the device.
//   MAF filter        Data::filter_snps_resize_F (src/Data.cpp:88-105)
//   read_pgen         FilePgen::FilePgen + read_all (src/FilePgen.hpp:14-35,
//                     src/FilePgen.cpp:27-89): hardcalls decoded on the host
//                     (pgen_io.cpp) into .bed rows, then the same upload
//   write_outputs     Data::write_eigs_files (src/Data.cpp:211-240),
//                     save_snps_in_mbim (src/Data.cpp:110-170; .pvar lines
//                     as .bim lines, src/Utils.cpp:405-430),
//                     make_plink2_eigenvec_file (src/Utils.cpp:225-238) and
//                     make_plink2_eigenvec_from_psam (src/Utils.cpp:240-265)
#include "plink_io.hpp"

#include <fstream>
#include <sstream>
#include <vector>

#include "common.hpp"
#include "pgen_io.hpp"

namespace pcaone {

namespace {

uint64_t count_lines(const std::string& path) {
  std::ifstream f(path);
  require(f.is_open(), kErrIO, "can not open " + path);
  uint64_t n = 0;
  std::string line;
  while (std::getline(f, line)) ++n;
  return n;
}

// Eigen IOFormat(6, DontAlignCols, "\t", "\n"): stream precision 6.
std::string fmt(double x) {
  std::ostringstream o;
  o.precision(6);
  o << x;
  return o.str();
}

void write_matrix(const std::string& path, const std::vector<double>& A, uint64_t rows, int cols) {
  std::ofstream o(path);
  require(o.is_open(), kErrIO, "can not open " + path);
  std::string buf;
  for (uint64_t r = 0; r < rows; ++r) {
    buf.clear();
    for (int c = 0; c < cols; ++c) {
      if (c) buf += '\t';
      buf += fmt(A[(uint64_t)c * rows + r]);
    }
    buf += '\n';
    o << buf;
  }
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ' ' || ch == '\t') {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  out.push_back(cur);
  return out;
}

std::vector<std::string> tokens(const std::string& s) {
  // split on runs of blanks and tabs
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ' ' || ch == '\t') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// pvar_line_to_bim_line (src/Utils.cpp:405-430): CHROM ID 0 POS ALT(first) REF
std::string pvar_to_bim(const std::string& line, const std::string& path) {
  const auto t = tokens(line);
  require(t.size() >= 5, kErrIO, "the input pvar file is not valid!\n => " + path + "\n" + line);
  const std::string alt = t[4].substr(0, t[4].find(','));
  return t[0] + "\t" + t[2] + "\t0\t" + t[1] + "\t" + alt + "\t" + t[3];
}

void filter_by_maf(Context& ctx, const Param& params, PlinkInfo& info) {
  // Data::filter_snps_resize_F (src/Data.cpp:88-105)
  if (!params.filterSNP) return;
  std::vector<uint32_t> keep;
  for (uint64_t j = 0; j < ctx.M; ++j) {
    const double f = ctx.F[j];
    const double maf = f > 0.5 ? 1 - f : f;  // MAF(a) (src/Common.hpp:26)
    if (maf > params.maf) keep.push_back((uint32_t)j);
  }
  require(!keep.empty(), kErrArgument, "no SNPs left after filtering!");
  ctx.filter_snps(keep);
  info.keep = keep;
}

}  // namespace

PlinkInfo plink_dims(const std::string& prefix) {
  // FileBed::FileBed (src/FilePlink.hpp:10-30): N and M from the line counts
  // of .fam and .bim, then the .bed opened and its magic checked
  PlinkInfo info;
  info.nsamples = count_lines(prefix + ".fam");
  info.nsnps_file = count_lines(prefix + ".bim");
  require(info.nsamples > 0 && info.nsnps_file > 0, kErrIO, "empty .fam or .bim");
  require(info.nsamples < (1ull << 32), kErrIO, "too many samples");
  std::ifstream bed(prefix + ".bed", std::ios::binary);
  require(bed.is_open(), kErrIO, "Cannot open bed file.");
  unsigned char header[3] = {0, 0, 0};
  bed.read(reinterpret_cast<char*>(header), 3);
  require(bed.gcount() == 3 && header[0] == 0x6c && header[1] == 0x1b && header[2] == 0x01, kErrIO,
          "Incorrect magic number in plink bed file.");
  return info;
}

PlinkInfo read_plink(Context& ctx, const std::string& prefix, const Param& params) {
  PlinkInfo info = plink_dims(prefix);
  std::ifstream bed(prefix + ".bed", std::ios::binary);
  require(bed.is_open(), kErrIO, "Cannot open bed file.");
  bed.seekg(3, std::ios::beg);
  const uint64_t bps = bytes_per_snp(info.nsamples);
  ctx.reserve((uint32_t)info.nsamples, info.nsnps_file);
  const uint64_t chunk = std::max<uint64_t>(1, (64ull << 20) / bps);
  std::vector<uint8_t> buf(std::min(chunk, info.nsnps_file) * bps);
  for (uint64_t j0 = 0; j0 < info.nsnps_file; j0 += chunk) {
    const uint64_t m = std::min(chunk, info.nsnps_file - j0);
    bed.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(m * bps));
    require((uint64_t)bed.gcount() == m * bps, kErrIO, "truncated .bed file");
    ctx.upload_rows(buf.data(), j0, m);
  }
  ctx.finalize();
  filter_by_maf(ctx, params, info);
  return info;
}

PlinkInfo pgen_dims(const std::string& prefix) {
  PlinkInfo info;
  info.pgen = true;
  info.nsamples = psam_samples(prefix + ".psam");
  require(info.nsamples > 0, kErrIO, "no samples in " + prefix + ".psam");
  info.nsnps_file = PgenSource(prefix + ".pgen", info.nsamples, true).nsnps();
  return info;
}

PlinkInfo read_pgen(Context& ctx, const std::string& prefix, const Param& params) {
  PlinkInfo info;
  info.pgen = true;
  info.nsamples = psam_samples(prefix + ".psam");
  require(info.nsamples > 0, kErrIO, "no samples in " + prefix + ".psam");
  const PgenSource src(prefix + ".pgen", info.nsamples, params.hardcall);
  info.nsnps_file = src.nsnps();
  const uint64_t bps = bytes_per_snp(info.nsamples);
  ctx.reserve((uint32_t)info.nsamples, info.nsnps_file);
  const uint64_t chunk = std::max<uint64_t>(1, (64ull << 20) / bps);
  std::vector<uint8_t> buf(std::min(chunk, info.nsnps_file) * bps);
  for (uint64_t j0 = 0; j0 < info.nsnps_file; j0 += chunk) {
    const uint64_t m = std::min(chunk, info.nsnps_file - j0);
    src.decode_rows(j0, m, nullptr, buf.data());
    ctx.upload_rows(buf.data(), j0, m);
  }
  ctx.finalize();
  ctx.pgen_input = true;
  filter_by_maf(ctx, params, info);
  return info;
}

PlinkInfo stream_pgen(Context& ctx, const std::string& prefix, const Param& params, uint64_t block_snps) {
  PlinkInfo info;
  info.pgen = true;
  info.nsamples = psam_samples(prefix + ".psam");
  require(info.nsamples > 0, kErrIO, "no samples in " + prefix + ".psam");
  auto src = std::make_shared<const PgenSource>(prefix + ".pgen", info.nsamples, params.hardcall);
  info.nsnps_file = src->nsnps();
  ctx.stream_pgen(std::move(src), block_snps);
  return info;
}

void write_residuals(Context& ctx, const PcaResult& r, const Param& params, const std::string& out) {
  // src/Data.cpp:242-291: uint32 nsnps, uint32 nsamples, then per SNP the
  // nsamples float32 values of its centered residual column, at the SNP's
  // position in the file order (the shuffle undone, src/Data.cpp:264-268).
  const uint64_t N = ctx.N, M = ctx.M;
  const int k = params.ld_stats == 0 ? r.k : 0;
  std::vector<double> vcur;
  if (k > 0) {
    require(r.V.size() == M * (uint64_t)r.k && r.U.size() == N * (uint64_t)r.k, kErrState,
            "residuals need U and V of the run");
    vcur.resize(M * (uint64_t)k);
    for (int c = 0; c < k; ++c)
      for (uint64_t i = 0; i < M; ++i) vcur[(uint64_t)c * M + i] = r.V[(uint64_t)c * M + ctx.perm[i]];
  }
  ctx.residuals_prepare(k, r.U.data(), r.S.data(), vcur.data(), params.out_of_core);
  std::ofstream f(out + ".residuals", std::ios::binary);
  require(f.is_open(), kErrIO, "can not open " + out + ".residuals");
  const uint32_t hdr[2] = {(uint32_t)M, (uint32_t)N};
  f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  const uint64_t rows = std::max<uint64_t>(1, std::min<uint64_t>(M, (256ull << 20) / (N * 4)));
  std::vector<float> buf(rows * N);
  for (uint64_t j0 = 0; j0 < M; j0 += rows) {
    const uint64_t j1 = std::min(M, j0 + rows);
    ctx.residuals_rows(j0, j1, buf.data());
    for (uint64_t j = j0; j < j1; ++j) {
      f.seekp((std::streamoff)(8 + (uint64_t)ctx.perm[j] * N * 4));
      f.write(reinterpret_cast<const char*>(buf.data() + (j - j0) * N), (std::streamsize)(N * 4));
    }
  }
  require(f.good(), kErrIO, "write failed: " + out + ".residuals");
}

void write_outputs(const Context& ctx, const PcaResult& r, const Param& params, const std::string& bfile,
                   const std::string& out, const PlinkInfo& info) {
  const uint64_t N = ctx.N, M = ctx.M;
  const int k = r.k;
  {
    std::ofstream o(out + ".sigvals");
    require(o.is_open(), kErrIO, "can not open " + out + ".sigvals");
    o << '#' << N << ',' << (r.S.empty() ? 0 : M) << '\n';
    for (double s : r.S) o << fmt(s) << '\n';
  }
  {
    std::ofstream o(out + ".eigvals");
    for (double e : r.E) o << fmt(e) << '\n';
  }
  write_matrix(out + ".eigvecs", r.U, N, k);
  if (params.printv || params.ld) {
    // .mbim: .bim line (a .pvar line as a .bim line) + allele frequency, kept SNPs only
    const std::string meta = bfile + (info.pgen ? ".pvar" : ".bim");
    std::ifstream bim(meta);
    std::ofstream mb(out + ".mbim");
    if (bim.is_open() && mb.is_open()) {
      std::string line;
      uint64_t src = 0, kept = 0;
      while (std::getline(bim, line) && kept < M) {
        if (info.pgen) {
          if (!line.empty() && line[0] == '#') continue;
          line = pvar_to_bim(line, meta);
        }
        if (info.keep.empty() || info.keep[kept] == src) {
          mb << line << '\t' << fmt(ctx.F[kept]) << '\n';
          ++kept;
        }
        ++src;
      }
    }
    if (params.printv) write_matrix(out + ".loadings", r.V, M, k);
  }
  // .eigvecs2: "#FID IID PC1.." + the .eigvecs rows
  std::ifstream ev(out + ".eigvecs");
  std::ofstream o2(out + ".eigvecs2");
  o2 << "#FID\tIID";
  for (int i = 0; i < k; ++i) o2 << "\tPC" << i + 1;
  o2 << '\n';
  std::string l1, l2;
  if (info.pgen) {
    // src/Utils.cpp:240-265: '##' lines skipped up to the '#' column header;
    // FID and IID when it starts '#FID', else FID 0 and the first column as IID
    std::ifstream psam(bfile + ".psam");
    std::string header;
    while (std::getline(psam, header))
      if (!header.empty() && header[0] == '#' && (header.size() < 2 || header[1] != '#')) break;
    const bool has_fid = header.compare(0, 4, "#FID") == 0;
    while (std::getline(psam, l1)) {
      if (l1.empty() || l1[0] == '#') continue;
      const auto tok = tokens(l1);
      std::getline(ev, l2);
      std::string fid = "0", iid;
      if (has_fid && tok.size() >= 2) {
        fid = tok[0];
        iid = tok[1];
      } else if (!tok.empty()) {
        iid = tok[0];
      }
      o2 << fid << '\t' << iid << '\t' << l2 << '\n';
    }
    return;
  }
  std::ifstream fam(bfile + ".fam");
  while (std::getline(fam, l1)) {
    auto tok = split_ws(l1);
    std::getline(ev, l2);
    o2 << tok[0] << '\t' << (tok.size() > 1 ? tok[1] : "") << '\t' << l2 << '\n';
  }
}

}  // namespace pcaone
