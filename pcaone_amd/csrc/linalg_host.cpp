// Host FP64 linear algebra on L x L matrices (L = k + oversamples <= 128).
// Row-major storage throughout.
#include "linalg_host.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "common.hpp"

namespace pcaone {
namespace la {

Mat chol_upper(const Mat& A) {
  const int n = A.n;
  Mat R(n);
  for (int j = 0; j < n; ++j) {
    double d = A(j, j);
    for (int k = 0; k < j; ++k) d -= R(k, j) * R(k, j);
    if (!(d > 0.0) || !std::isfinite(d))
      fail(kErrNumeric, "Cholesky breakdown at column " + std::to_string(j) + " (Gram matrix not positive definite)");
    const double rjj = std::sqrt(d);
    R(j, j) = rjj;
    for (int i = j + 1; i < n; ++i) {
      double s = A(j, i);
      for (int k = 0; k < j; ++k) s -= R(k, j) * R(k, i);
      R(j, i) = s / rjj;
    }
  }
  return R;
}

bool chol_try(const Mat& A, Mat& R, double& min_ratio) {
  const int n = A.n;
  R = Mat(n);
  min_ratio = 1.0;
  double dmax = 0.0;
  for (int j = 0; j < n; ++j) dmax = std::max(dmax, A(j, j));
  for (int j = 0; j < n; ++j) {
    double d = A(j, j);
    for (int k = 0; k < j; ++k) d -= R(k, j) * R(k, j);
    if (!(d > 0.0) || !std::isfinite(d) || !(dmax > 0.0)) {
      min_ratio = 0.0;
      return false;
    }
    min_ratio = std::min(min_ratio, d / dmax);
    const double rjj = std::sqrt(d);
    R(j, j) = rjj;
    for (int i = j + 1; i < n; ++i) {
      double s = A(j, i);
      for (int k = 0; k < j; ++k) s -= R(k, j) * R(k, i);
      R(j, i) = s / rjj;
    }
  }
  return true;
}

void householder_q(double* X, uint64_t n, int L, int ld) {
  require(n >= (uint64_t)L && L > 0, kErrArgument, "householder_q: needs rows >= columns");
  std::vector<double> A(n * (uint64_t)L);  // column-major copy: reflectors work on contiguous columns
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    for (int j = 0; j < L; ++j) A[(uint64_t)j * n + i] = X[(uint64_t)i * ld + j];
  std::vector<double> tau(L, 0.0);
  [[maybe_unused]] const bool par = n * (uint64_t)L > 65536;
  // dgeqr2: H_j = I - tau_j v_j v_j^T with v_j = [1; A(j+1:n, j)]
  for (int j = 0; j < L; ++j) {
    double* a = A.data() + (uint64_t)j * n;
    const double c0 = a[j];
    double tail = 0.0;
    for (uint64_t i = j + 1; i < n; ++i) tail += a[i] * a[i];
    double beta = c0, t = 0.0;
    if (tail > std::numeric_limits<double>::min()) {
      beta = std::sqrt(c0 * c0 + tail);
      if (c0 >= 0.0) beta = -beta;
      const double inv = 1.0 / (c0 - beta);
      for (uint64_t i = j + 1; i < n; ++i) a[i] *= inv;
      t = (beta - c0) / beta;
    } else {
      for (uint64_t i = j + 1; i < n; ++i) a[i] = 0.0;
    }
    a[j] = beta;
    tau[j] = t;
#pragma omp parallel for schedule(static) if (par)
    for (int c = j + 1; c < L; ++c) {
      double* b = A.data() + (uint64_t)c * n;
      double w = b[j];
      for (uint64_t i = j + 1; i < n; ++i) w += a[i] * b[i];
      w *= t;
      b[j] -= w;
      for (uint64_t i = j + 1; i < n; ++i) b[i] -= w * a[i];
    }
  }
  // dorg2r: Q = H_0 ... H_{L-1} [I; 0], formed in place from the last reflector
  for (int j = L - 1; j >= 0; --j) {
    double* a = A.data() + (uint64_t)j * n;
    const double t = tau[j];
    a[j] = 1.0;
#pragma omp parallel for schedule(static) if (par)
    for (int c = j + 1; c < L; ++c) {
      double* b = A.data() + (uint64_t)c * n;
      double w = 0.0;
      for (uint64_t i = j; i < n; ++i) w += a[i] * b[i];
      w *= t;
      for (uint64_t i = j; i < n; ++i) b[i] -= w * a[i];
    }
    for (uint64_t i = j + 1; i < n; ++i) a[i] *= -t;
    a[j] = 1.0 - t;
    for (int i = 0; i < j; ++i) a[i] = 0.0;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    for (int j = 0; j < L; ++j) X[(uint64_t)i * ld + j] = A[(uint64_t)j * n + i];
}

// Rows (matmul) and columns (inv_upper) are independent: at the larger L they
// are spread over a few host threads (more only add start-up latency at these
// sizes); each element's arithmetic order is unchanged.
[[maybe_unused]] constexpr int kParallelN = 48, kThreads = 8;  // (unused in the device pass)

Mat inv_upper(const Mat& R) {
  const int n = R.n;
  Mat X(n);
#pragma omp parallel for schedule(dynamic, 4) num_threads(kThreads) if (n >= kParallelN)
  for (int j = n - 1; j >= 0; --j) {
    X(j, j) = 1.0 / R(j, j);
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R(i, k) * X(k, j);
      X(i, j) = -s / R(i, i);
    }
  }
  return X;
}

Mat matmul(const Mat& A, const Mat& B) {
  const int n = A.n;
  Mat C(n);
#pragma omp parallel for schedule(static) num_threads(kThreads) if (n >= kParallelN)
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const double a = A(i, k);
      if (a == 0.0) continue;
      for (int j = 0; j < n; ++j) C(i, j) += a * B(k, j);
    }
  return C;
}

Mat transpose(const Mat& A) {
  Mat T(A.n);
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < A.n; ++j) T(j, i) = A(i, j);
  return T;
}

// Householder reduction of a symmetric matrix to tridiagonal form followed by
// the implicit-shift QL iteration (the EISPACK tred2/tql2 method).  V holds
// the accumulated orthogonal transformation; on return A = V diag(d) V^T.
// V is addressed column-major (element (i, j) at j n + i; the symmetric input
// reads the same either way), so the method's inner loops over a column, and
// the QL rotations of two eigenvector columns, run over contiguous memory.
static void tred2(int n, std::vector<double>& V, std::vector<double>& d, std::vector<double>& e) {
  auto v = [&](int i, int j) -> double& { return V[(size_t)j * n + i]; };
  for (int j = 0; j < n; ++j) d[j] = v(n - 1, j);
  for (int i = n - 1; i > 0; --i) {
    double scale = 0.0, h = 0.0;
    for (int k = 0; k < i; ++k) scale += std::fabs(d[k]);
    if (scale == 0.0) {
      e[i] = d[i - 1];
      for (int j = 0; j < i; ++j) {
        d[j] = v(i - 1, j);
        v(i, j) = 0.0;
        v(j, i) = 0.0;
      }
    } else {
      for (int k = 0; k < i; ++k) {
        d[k] /= scale;
        h += d[k] * d[k];
      }
      double f = d[i - 1];
      double g = std::sqrt(h);
      if (f > 0) g = -g;
      e[i] = scale * g;
      h -= f * g;
      d[i - 1] = f - g;
      for (int j = 0; j < i; ++j) e[j] = 0.0;
      for (int j = 0; j < i; ++j) {
        f = d[j];
        v(j, i) = f;
        g = e[j] + v(j, j) * f;
        for (int k = j + 1; k <= i - 1; ++k) {
          g += v(k, j) * d[k];
          e[k] += v(k, j) * f;
        }
        e[j] = g;
      }
      f = 0.0;
      for (int j = 0; j < i; ++j) {
        e[j] /= h;
        f += e[j] * d[j];
      }
      const double hh = f / (h + h);
      for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
      for (int j = 0; j < i; ++j) {
        f = d[j];
        g = e[j];
        for (int k = j; k <= i - 1; ++k) v(k, j) -= (f * e[k] + g * d[k]);
        d[j] = v(i - 1, j);
        v(i, j) = 0.0;
      }
    }
    d[i] = h;
  }
  for (int i = 0; i < n - 1; ++i) {
    v(n - 1, i) = v(i, i);
    v(i, i) = 1.0;
    const double h = d[i + 1];
    if (h != 0.0) {
      for (int k = 0; k <= i; ++k) d[k] = v(k, i + 1) / h;
      for (int j = 0; j <= i; ++j) {
        double g = 0.0;
        for (int k = 0; k <= i; ++k) g += v(k, i + 1) * v(k, j);
        for (int k = 0; k <= i; ++k) v(k, j) -= g * d[k];
      }
    }
    for (int k = 0; k <= i; ++k) v(k, i + 1) = 0.0;
  }
  for (int j = 0; j < n; ++j) {
    d[j] = v(n - 1, j);
    v(n - 1, j) = 0.0;
  }
  v(n - 1, n - 1) = 1.0;
  e[0] = 0.0;
}

// sqrt(a^2 + b^2): the plain formula where the squares neither overflow nor
// lose precision to underflow, std::hypot's careful scaling otherwise.
static inline double norm2(double a, double b) {
  const double s = a * a + b * b;
  return (s > 1e-280 && s < 1e280) ? std::sqrt(s) : std::hypot(a, b);
}

static void tql2(int n, std::vector<double>& V, std::vector<double>& d, std::vector<double>& e) {
  auto v = [&](int i, int j) -> double& { return V[(size_t)j * n + i]; };
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  double f = 0.0, tst1 = 0.0;
  const double eps = std::ldexp(1.0, -52);
  for (int l = 0; l < n; ++l) {
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    int m = l;
    while (m < n) {
      if (std::fabs(e[m]) <= eps * tst1) break;
      ++m;
    }
    if (m == n) m = n - 1;
    if (m > l) {
      int iter = 0;
      do {
        if (++iter > 200) fail(kErrNumeric, "symmetric eigensolver did not converge");
        double g = d[l];
        double p = (d[l + 1] - g) / (2.0 * e[l]);
        double r = norm2(p, 1.0);
        if (p < 0) r = -r;
        d[l] = e[l] / (p + r);
        d[l + 1] = e[l] * (p + r);
        const double dl1 = d[l + 1];
        double h = g - d[l];
        for (int i = l + 2; i < n; ++i) d[i] -= h;
        f += h;
        p = d[m];
        double c = 1.0, c2 = c, c3 = c;
        const double el1 = e[l + 1];
        double s = 0.0, s2 = 0.0;
        for (int i = m - 1; i >= l; --i) {
          c3 = c2;
          c2 = c;
          s2 = s;
          g = c * e[i];
          h = c * p;
          r = norm2(p, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = p / r;
          p = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
          for (int k = 0; k < n; ++k) {
            h = v(k, i + 1);
            v(k, i + 1) = s * v(k, i) + c * h;
            v(k, i) = c * v(k, i) - s * h;
          }
        }
        p = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * p;
        d[l] = c * p;
      } while (std::fabs(e[l]) > eps * tst1);
    }
    d[l] = d[l] + f;
    e[l] = 0.0;
  }
}

void sym_eig_desc(const Mat& A, std::vector<double>& w, Mat& Vout) {
  const int n = A.n;
  std::vector<double> V(A.a), d(n), e(n);
  // symmetrize (inputs are symmetric up to rounding)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) V[(size_t)i * n + j] = V[(size_t)j * n + i] = 0.5 * (A(i, j) + A(j, i));
  if (n == 1) {
    w.assign(1, A(0, 0));
    Vout = Mat(1);
    Vout(0, 0) = 1.0;
    return;
  }
  tred2(n, V, d, e);
  tql2(n, V, d, e);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] > d[b]; });
  w.resize(n);
  Vout = Mat(n);
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    for (int i = 0; i < n; ++i) Vout(i, j) = V[(size_t)idx[j] * n + i];  // column-major V
  }
}

// Column signs s such that Q diag(s) equals the thin Q of Householder QR with
// reflectors beta = -sign(alpha) * ||x|| (LAPACK / Eigen convention), given
// the top n x n block Y of an orthonormal factor with positive-diagonal R.
// Method: the modified LU factorization of Y - diag(s) of Ballard, Demmel,
// Grigori, Jacquelin, Nguyen, Solomonik, "Reconstructing Householder vectors
// from tall-skinny QR" (IPDPS 2014), which picks s_i = -sign(pivot_i).
std::vector<double> householder_signs(const Mat& Y) {
  const int n = Y.n;
  Mat W = Y;
  std::vector<double> s(n);
  for (int i = 0; i < n; ++i) {
    s[i] = W(i, i) >= 0.0 ? -1.0 : 1.0;
    W(i, i) -= s[i];
    const double piv = W(i, i);
    for (int r = i + 1; r < n; ++r) {
      W(r, i) /= piv;
      const double lr = W(r, i);
      if (lr == 0.0) continue;
      for (int c = i + 1; c < n; ++c) W(r, c) -= lr * W(i, c);
    }
  }
  return s;
}

}  // namespace la
}  // namespace pcaone
