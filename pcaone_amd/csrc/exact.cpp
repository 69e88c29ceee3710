// --svd 3 (exact PCA, src/Main.cpp:180-217) and --svd 0 (IRAM,
// src/Arnoldi.cpp:48-118) on the device-resident genotypes.
//
// Exact PCA: the Gram matrix of the standardized genotypes, A A^T / M (N x N,
// N <= M) or A^T A / M (M x M), built block by block with the FP64 packed
// products (a block of standardized sample columns as G~ through the H
// product, or a block of SNP rows as Omega through the G product), then a
// symmetric eigen-decomposition on the device (rocSOLVER dsyevd).
//
// IRAM: the top-k eigenpairs of the same operator, applied one vector at a time
// through the products, by a thick-restart Lanczos (Krylov-Schur) iteration
// with full reorthogonalization on the host: the restarted Lanczos process
// that Spectra's SymEigsSolver also runs (implicit restarts with exact shifts
// keep the same Krylov subspace as a thick restart; Wu and Simon, "Thick-restart
// Lanczos method for large symmetric eigenvalue problems", SIAM J. Matrix Anal.
// Appl. 22, 2000; Stewart, "A Krylov-Schur algorithm for large eigenproblems",
// SIAM J. Matrix Anal. Appl. 23, 2001).  Its stop rule is Spectra's
// (SymEigsBase::num_converged): every one of the k wanted Ritz pairs has a
// residual below itol x max(eps^(2/3), |theta|).
#include "exact.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <random>

#include "common.hpp"
#include "kernels.hpp"
#include "linalg_host.hpp"

namespace pcaone {

namespace {

constexpr int kGramBlock = 64;  // Gram columns per FP64 product

// flip_UV with U as the base (src/Utils.cpp:118-133): the largest |U| entry of
// each column (the first one on ties) is made positive, V follows.
void flip_uv_ubase(std::vector<double>& U, std::vector<double>& V, uint64_t n, uint64_t m, int k) {
  for (int c = 0; c < k; ++c) {
    uint64_t x = 0;
    double best = -1.0;
    for (uint64_t i = 0; i < n; ++i) {
      const double a = std::fabs(U[(uint64_t)c * n + i]);
      if (a > best) best = a, x = i;
    }
    if (U[(uint64_t)c * n + x] < 0) {
      for (uint64_t i = 0; i < n; ++i) U[(uint64_t)c * n + i] = -U[(uint64_t)c * n + i];
      for (uint64_t i = 0; i < m; ++i) V[(uint64_t)c * m + i] = -V[(uint64_t)c * m + i];
    }
  }
}

// The product of the standardized matrix with k vectors on one side, giving the
// other side: wide = true: X (N x k) -> A X (M x k); false: Y (M x k) -> A^T Y.
// (A is M x N here: SNP rows.)  Column-major host arrays.
void apply_a(Context& ctx, bool from_samples, const double* X, int k, double* out) {
  const uint64_t N = ctx.N, M = ctx.M;
  if (from_samples) {
    std::vector<double> om(N * (uint64_t)ctx.L, 0.0);
    std::copy(X, X + N * (uint64_t)k, om.begin());
    ctx.set_omega(om.data());
    ctx.sketch(0, M, 0);
    std::vector<double> g(M * (uint64_t)ctx.L);
    ctx.download_g(0, M, g.data());
    std::copy(g.begin(), g.begin() + M * (uint64_t)k, out);
  } else {
    std::vector<double> gt(M * (uint64_t)ctx.lp, 0.0);
    for (int c = 0; c < k; ++c)
      for (uint64_t j = 0; j < M; ++j) gt[j * ctx.lp + c] = X[(uint64_t)c * M + j];
    PCAONE_HIP(hipMemcpyAsync(ctx.gt(), gt.data(), gt.size() * 8, hipMemcpyHostToDevice, ctx.stream));
    ctx.h_from_g(0);
    std::vector<double> h(N * (uint64_t)ctx.L);
    ctx.download_h(0, h.data());
    std::copy(h.begin(), h.begin() + N * (uint64_t)k, out);
  }
}

}  // namespace

PcaResult run_pca_full(Context& ctx, const Param& params, RunInfo* info) {
  const auto t0 = std::chrono::steady_clock::now();
  ctx.reset_columns();
  const uint64_t N = ctx.N, M = ctx.M;
  require(N > 0 && M > 0, kErrState, "no data");
  require(!ctx.streamed(), kErrArgument, "--svd 0 and --svd 3 need the genotypes resident (not streamed)");
  ctx.configure(PCAONE_PREC_FP64_ID, kGramBlock, {});
  ctx.standardize(params.ploidy);  // src/Main.cpp:181-182
  const bool wide = N <= M;
  const uint64_t n = wide ? N : M;
  require(n < (1ull << 31), kErrArgument, "--svd 3: the Gram matrix dimension must fit 32 bits");
  const int ncomp = (int)std::min<uint64_t>(params.k, std::min(N, M));
  DevBuf<double> K(n * n);
  if (wide) {
    // K[:, s0:s0 + nc] = A_std^T (standardized sample columns s0..) / M
    for (uint64_t s0 = 0; s0 < N; s0 += kGramBlock) {
      const int nc = (int)std::min<uint64_t>(kGramBlock, N - s0);
      launch_sample_columns(ctx.xs(), ctx.pitch_s(), M, s0, nc, ctx.lp, ctx.sc(), ctx.cf(), ctx.gt(), ctx.stream);
      ctx.h_from_g(0);
      launch_rows_to_cols(ctx.h(), N, ctx.lp, nc, 1.0 / (double)M, K.get(), n, s0, ctx.stream);
    }
  } else {
    // K[:, t0:t0 + nc] = A_std (standardized SNP rows t0..)^T / M
    for (uint64_t s0 = 0; s0 < M; s0 += kGramBlock) {
      const int nc = (int)std::min<uint64_t>(kGramBlock, M - s0);
      launch_snp_rows(ctx.xs(), ctx.pitch_s(), N, s0, nc, ctx.L, ctx.sc(), ctx.cf(), ctx.omega_scratch(), ctx.stream);
      ctx.set_omega_scratch();
      ctx.sketch(0, M, 0);
      launch_rows_to_cols(ctx.gt(), M, ctx.lp, nc, 1.0 / (double)M, K.get(), n, s0, ctx.stream);
    }
  }
  // SelfAdjointEigenSolver (src/Main.cpp:187): eigenvalues ascending
  DevBuf<double> D(n), E(n);
  DevBuf<rocblas_int> dinfo(1);
  rocblas_handle h = nullptr;
  require(rocblas_create_handle(&h) == rocblas_status_success, kErrDevice, "rocblas_create_handle failed");
  rocblas_set_stream(h, ctx.stream);
  const rocblas_status st =
      rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, K.get(), (rocblas_int)n, D.get(),
                       E.get(), dinfo.get());
  rocblas_destroy_handle(h);
  require(st == rocblas_status_success, kErrDevice, "rocsolver_dsyevd failed");
  rocblas_int hinfo = 0;
  std::vector<double> w(n);
  PCAONE_HIP(hipMemcpyAsync(&hinfo, dinfo.get(), sizeof(hinfo), hipMemcpyDeviceToHost, ctx.stream));
  PCAONE_HIP(hipMemcpyAsync(w.data(), D.get(), n * 8, hipMemcpyDeviceToHost, ctx.stream));
  std::vector<double> top(n * (uint64_t)ncomp);
  for (int i = 0; i < ncomp; ++i)  // the largest first
    PCAONE_HIP(hipMemcpyAsync(top.data() + (uint64_t)i * n, K.get() + (n - 1 - i) * n, n * 8, hipMemcpyDeviceToHost,
                              ctx.stream));
  PCAONE_HIP(hipStreamSynchronize(ctx.stream));
  require(hinfo == 0, kErrNumeric, "failed eigendecomposition of the sample covariance matrix.");
  PcaResult res;
  res.k = ncomp;
  res.E.resize(ncomp);
  res.S.resize(ncomp);
  for (int i = 0; i < ncomp; ++i) {
    res.E[i] = std::max(0.0, w[n - 1 - i]);
    res.S[i] = std::sqrt(res.E[i] * (double)M);
  }
  // the other side: V = A_std U / s (wide) or U = A_std^T V / s (tall)
  const int Lv = (int)round_up((uint64_t)ncomp, 16);
  ctx.configure(PCAONE_PREC_FP64_ID, Lv, {});
  std::vector<double> other((wide ? M : N) * (uint64_t)ncomp);
  apply_a(ctx, wide, top.data(), ncomp, other.data());
  const uint64_t no = wide ? M : N;
  for (int c = 0; c < ncomp; ++c)
    if (res.S[c] > 0)
      for (uint64_t i = 0; i < no; ++i) other[(uint64_t)c * no + i] /= res.S[c];
  if (wide) {
    res.U = std::move(top);
    res.V = std::move(other);
  } else {
    res.V = std::move(top);
    res.U = std::move(other);
  }
  flip_uv_ubase(res.U, res.V, N, M, ncomp);  // src/Main.cpp:215
  if (info) {
    info->epochs = 1;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return res;
}

PcaResult run_pca_arnoldi(Context& ctx, const Param& params, RunInfo* info) {
  const auto t0 = std::chrono::steady_clock::now();
  ctx.reset_columns();
  const uint64_t N = ctx.N, M = ctx.M;
  require(N > 0 && M > 0, kErrState, "no data");
  require(!ctx.streamed(), kErrArgument, "--svd 0 and --svd 3 need the genotypes resident (not streamed)");
  require(!params.emu, kErrArgument, "--svd 0 with --emu does not run on this path");
  const int nev = (int)params.k;
  // PartialSVDSolver (external/Spectra/contrib/PartialSVDSolver.h): the
  // eigen-problem of the smaller Gram side; m = min(N, M)
  const bool wide = N <= M;
  const uint64_t n = wide ? N : M;
  const int ncv = (int)std::min<uint64_t>(params.ncv, n);
  require(nev >= 1 && nev < ncv, kErrArgument, "--svd 0 needs k < ncv <= min(N, M)");
  // The operator runs in the FP64 form: the Lanczos residuals must fall to
  // itol (1e-6) of the eigenvalues, below the rounding of the reduced forms'
  // operand pieces, which would make the operator visibly non-symmetric.
  ctx.configure(PCAONE_PREC_FP64_ID, 16, params.tuning);
  if (!params.ld) ctx.standardize(params.ploidy);  // src/Arnoldi.cpp:60
  std::vector<double> xbuf(n * 16, 0.0), ybuf(n);
  uint64_t nops = 0;
  const auto op = [&](const double* x, double* y) {
    if (wide) {
      // y = A_std^T (A_std x): G = A Omega, H = A^T G with Omega = x
      std::copy(x, x + n, xbuf.begin());
      ctx.set_omega(xbuf.data());
      ctx.sketch(0, M, 0);
      std::vector<double> h(N * (uint64_t)ctx.L);
      ctx.download_h(0, h.data());
      std::copy(h.begin(), h.begin() + N, y);
    } else {
      // y = A_std (A_std^T x) on SNP vectors
      std::vector<double> mid(N);
      apply_a(ctx, false, x, 1, mid.data());
      apply_a(ctx, true, mid.data(), 1, y);
    }
    ++nops;
  };
  // Krylov basis V (n x ncv, column-major) and the projected matrix T = V^T Op V
  std::vector<double> V(n * (uint64_t)(ncv + 1)), w(n);
  la::Mat T(ncv);
  std::mt19937_64 rng(params.seed);
  std::normal_distribution<double> nd;
  for (uint64_t i = 0; i < n; ++i) V[i] = nd(rng);
  auto col = [&](int j) { return V.data() + (uint64_t)j * n; };
  auto dot = [&](const double* a, const double* b) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  };
  {
    const double nr = std::sqrt(dot(col(0), col(0)));
    for (uint64_t i = 0; i < n; ++i) col(0)[i] /= nr;
  }
  int kept = 0;  // basis vectors carried over by the last restart (Ritz vectors)
  double beta = 0.0;
  std::vector<double> theta;
  la::Mat S;
  bool converged = false;
  uint32_t restarts = 0;
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  for (;;) {
    for (int j = kept; j < ncv; ++j) {
      op(col(j), w.data());
      // full reorthogonalization (twice): T[:, j] = V^T w
      std::vector<double> hcoef(j + 1, 0.0);
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i <= j; ++i) {
          const double c = dot(col(i), w.data());
          hcoef[i] += c;
          for (uint64_t r = 0; r < n; ++r) w[r] -= c * col(i)[r];
        }
      for (int i = 0; i <= j; ++i) {
        // the column of T; the arrowhead of a thick restart sits in rows < kept
        if (i >= kept || j == kept) T(i, j) = T(j, i) = hcoef[i];
      }
      beta = std::sqrt(dot(w.data(), w.data()));
      double scale = 0.0;
      for (int i = 0; i <= j; ++i) scale = std::max(scale, std::fabs(hcoef[i]));
      double* nx = col(j + 1);
      if (beta <= 1e-13 * scale) {
        // an invariant subspace: continue from a fresh direction orthogonal to
        // the basis; the coupling (and the residual of every Ritz pair) is 0
        beta = 0.0;
        for (uint64_t r = 0; r < n; ++r) nx[r] = nd(rng);
        for (int pass = 0; pass < 2; ++pass)
          for (int i = 0; i <= j; ++i) {
            const double c = dot(col(i), nx);
            for (uint64_t r = 0; r < n; ++r) nx[r] -= c * col(i)[r];
          }
        const double nr = std::sqrt(dot(nx, nx));
        for (uint64_t r = 0; r < n; ++r) nx[r] /= nr;
      } else {
        for (uint64_t r = 0; r < n; ++r) nx[r] = w[r] / beta;
      }
      if (j + 1 < ncv) T(j + 1, j) = T(j, j + 1) = beta;
    }
    // Ritz pairs of T, largest algebraic first
    la::sym_eig_desc(T, theta, S);
    int nconv = 0;
    for (int i = 0; i < nev; ++i) {
      const double resid = std::fabs(beta * S(ncv - 1, i));
      if (resid < params.itol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
    }
    if (nconv >= nev) {
      converged = true;
      break;
    }
    if (++restarts > params.imaxiter) break;
    // thick restart: keep the leading Ritz vectors, then the residual direction
    kept = std::min(ncv - 1, nev + (ncv - nev) / 2);
    std::vector<double> Y(n * (uint64_t)kept, 0.0);
    for (int c = 0; c < kept; ++c)
      for (int q = 0; q < ncv; ++q) {
        const double s = S(q, c);
        if (s == 0.0) continue;
        for (uint64_t r = 0; r < n; ++r) Y[(uint64_t)c * n + r] += s * col(q)[r];
      }
    const std::vector<double> resid(col(ncv), col(ncv) + n);
    std::copy(Y.begin(), Y.end(), V.begin());
    std::copy(resid.begin(), resid.end(), col(kept));
    T = la::Mat(ncv);
    for (int c = 0; c < kept; ++c) {
      T(c, c) = theta[c];
      T(c, kept) = T(kept, c) = beta * S(ncv - 1, c);
    }
  }
  require(converged, kErrNumeric, "the nconv is not equal to k.");  // src/Arnoldi.cpp:62
  // Ritz vectors of the k wanted pairs
  std::vector<double> Y(n * (uint64_t)nev, 0.0);
  for (int c = 0; c < nev; ++c)
    for (int q = 0; q < ncv; ++q)
      for (uint64_t r = 0; r < n; ++r) Y[(uint64_t)c * n + r] += S(q, c) * col(q)[r];
  PcaResult res;
  res.k = nev;
  res.S.resize(nev);
  res.E.resize(nev);
  for (int c = 0; c < nev; ++c) {
    res.S[c] = std::sqrt(std::max(theta[c], 0.0));
    res.E[c] = res.S[c] * res.S[c] / (double)M;  // src/Arnoldi.cpp:67
  }
  // matrix_U / matrix_V of PartialSVDSolver: the other side is A Y / s
  const int Lv = (int)round_up((uint64_t)nev, 16);
  ctx.configure(PCAONE_PREC_FP64_ID, Lv, params.tuning);
  std::vector<double> other((wide ? M : N) * (uint64_t)nev);
  apply_a(ctx, wide, Y.data(), nev, other.data());
  const uint64_t no = wide ? M : N;
  for (int c = 0; c < nev; ++c)
    for (uint64_t i = 0; i < no; ++i) other[(uint64_t)c * no + i] /= res.S[c];
  if (wide) {
    res.U = std::move(Y);
    res.V = std::move(other);
  } else {
    res.V = std::move(Y);
    res.U = std::move(other);
  }
  if (info) {
    info->epochs = (int)restarts;
    info->total_epochs = (int)nops;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return res;
}

}  // namespace pcaone
