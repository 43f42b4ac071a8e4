#include "bmatrix.hh"

#include <cmath>
#include <stdexcept>

namespace libcppe {

BMatrix::BMatrix(std::vector<Potential> polsites, const PeOptions& options)
      : m_polsites(std::move(polsites)), m_options(options) {
  double damping = m_options.damp_induced ? m_options.damping_factor_induced : 0.0;  // bmatrix.cc:18-21
  // no polarizable sites: B is the empty operator, nothing goes to the device
  if (!m_polsites.empty()) m_dev = std::make_unique<DeviceSystem>(m_polsites, m_options, damping, true);
}

static void check_size(const Vector& v, size_t n) {
  if (v.size() != 3 * n) throw std::invalid_argument("vector size does not match 3 * number of polarizable sites");
}

Vector BMatrix::apply(const Vector& x) {
  check_size(x, size());
  Vector out(x.size());
  if (out.empty()) return out;
  hip_check(cppe_hip_apply_b(m_dev->get(), x.data(), out.data()));
  return out;
}

Vector BMatrix::apply_diagonal_inverse(const Vector& x) {
  check_size(x, size());
  Vector out(x.size());
  if (out.empty()) return out;
  hip_check(cppe_hip_apply_alpha(m_dev->get(), x.data(), out.data()));
  return out;
}

Vector BMatrix::apply_diagonal(const Vector& x) {
  check_size(x, size());
  Vector out(x.size());
  if (out.empty()) return out;
  hip_check(cppe_hip_apply_alpha_inv(m_dev->get(), x.data(), out.data()));
  return out;
}

Matrix BMatrix::to_dense_matrix() {
  const size_t n = size();
  std::vector<int> index(n);
  for (size_t i = 0; i < n; ++i) index[i] = m_polsites[i].index;
  Matrix B(3 * n, 3 * n);
  if (n == 0) return B;
  hip_check(cppe_hip_dense_b(m_dev->get(), index.data(), B.data.data()));
  return B;
}

Matrix BMatrix::direct_inverse() {
  const size_t n = size();
  std::vector<int> index(n);
  for (size_t i = 0; i < n; ++i) index[i] = m_polsites[i].index;
  Matrix B(3 * n, 3 * n);
  if (n == 0) return B;
  hip_check(cppe_hip_dense_b_inverse(m_dev->get(), index.data(), B.data.data()));
  return B;
}

}  // namespace libcppe
