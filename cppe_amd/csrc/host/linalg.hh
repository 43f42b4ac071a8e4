// Small dense types of the host API (the reference uses Eigen; it is not available here).
#pragma once
#include <array>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace libcppe {

using Vector = std::vector<double>;

// Row-major dense matrix.
struct Matrix {
  size_t rows = 0, cols = 0;
  std::vector<double> data;
  Matrix() = default;
  Matrix(size_t r, size_t c) : rows(r), cols(c), data(r * c, 0.0) {}
  double& operator()(size_t r, size_t c) { return data[r * cols + c]; }
  double operator()(size_t r, size_t c) const { return data[r * cols + c]; }
};

using Matrix3 = std::array<double, 9>;  // row-major 3x3

inline Matrix3 triangle_to_mat(const std::vector<double>& y) {  // math.cc:7-17
  if (y.size() != 6) throw std::runtime_error("triangle_to_mat needs 6 values");
  return Matrix3{y[0], y[1], y[2], y[1], y[3], y[4], y[2], y[4], y[5]};
}

inline std::vector<double> mat_to_triangle(const Matrix3& m) {  // math.cc:22-31
  return {m[0], m[1], m[2], m[4], m[5], m[8]};
}

inline Vector operator+(const Vector& a, const Vector& b) {
  if (a.size() != b.size()) throw std::invalid_argument("vector sizes differ");
  Vector r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] + b[i];
  return r;
}

}  // namespace libcppe
