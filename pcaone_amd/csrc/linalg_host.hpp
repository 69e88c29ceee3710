#pragma once
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
Mat inv_upper(const Mat& R);
Mat matmul(const Mat& A, const Mat& B);
Mat transpose(const Mat& A);
void sym_eig_desc(const Mat& A, std::vector<double>& w, Mat& V);  // A = V diag(w) V^T, w descending
std::vector<double> householder_signs(const Mat& Y);

}  // namespace la
}  // namespace pcaone
