#pragma once
#include <cstdint>
#include <vector>

namespace pcaone {
namespace la {

// Dense square row-major FP64 matrix.
struct Mat {
  int n = 0;
  std::vector<double> a;
  Mat() = default;
  explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)i * n + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * n + j]; }
  static Mat identity(int n_) {
    Mat I(n_);
    for (int i = 0; i < n_; ++i) I(i, i) = 1.0;
    return I;
  }
};

Mat chol_upper(const Mat& A);  // A = R^T R, R upper, positive diagonal
// The same without throwing: false on breakdown.  min_ratio = min over j of
// the pivot d_j / max_i A_ii: the squared length of column j's component
// orthogonal to the columns before it, relative to the longest column.  It is
// a lower bound of the smallest eigenvalue over the largest up to a factor n,
// so a near-zero or dependent column shows as min_ratio ~ 0 (CholeskyQR of X
// loses about u / min_ratio of orthogonality).
bool chol_try(const Mat& A, Mat& R, double& min_ratio);
Mat inv_upper(const Mat& R);
Mat matmul(const Mat& A, const Mat& B);
Mat transpose(const Mat& A);
void sym_eig_desc(const Mat& A, std::vector<double>& w, Mat& V);  // A = V diag(w) V^T, w descending
std::vector<double> householder_signs(const Mat& Y);
// Thin Q (n x L) of the Householder QR of X (n x L, row-major, row pitch ld),
// in place: LAPACK dgeqr2 + dorg2r, with the reflector convention
// beta = -sign(alpha) ||x|| that Eigen's HouseholderQR also uses.  Never breaks
// down: rank-deficient X still gives L orthonormal columns.
void householder_q(double* X, uint64_t n, int L, int ld);

}  // namespace la
}  // namespace pcaone
