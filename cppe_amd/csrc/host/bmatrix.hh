// Classical response operator B = alpha^-1 - T2 over polarizable sites (src/core/bmatrix.hh:10-43).
#pragma once
#include <stdexcept>
#include <memory>
#include <vector>

#include "device_system.hh"
#include "linalg.hh"
#include "pe_options.hh"
#include "potential.hh"

namespace libcppe {

class BMatrix {
 public:
  BMatrix(std::vector<Potential> polsites, const PeOptions& options);
  Vector apply(const Vector& induced_moments);        // bmatrix.cc:9-40
  Vector apply_diagonal_inverse(const Vector& in);    // alpha x  (bmatrix.cc:42-51)
  Vector apply_diagonal(const Vector& in);            // alpha^-1 x (bmatrix.cc:53-62)
  Matrix to_dense_matrix();                            // bmatrix.cc:64-92
  Matrix direct_inverse();                             // bmatrix.cc:94-97
  DeviceSystem& device() {
    if (!m_dev) throw std::runtime_error("B matrix has no polarizable sites");
    return *m_dev;
  }
  size_t size() const { return m_polsites.size(); }

 private:
  std::vector<Potential> m_polsites;
  PeOptions m_options;
  std::unique_ptr<DeviceSystem> m_dev;
};

}  // namespace libcppe
