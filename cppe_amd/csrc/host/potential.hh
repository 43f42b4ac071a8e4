// Site data (src/core/potential.hh:11-127): packed multipoles, 3x3 polarizability, 0-based
// exclusions and the potfile index.
#pragma once
#include <algorithm>
#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "linalg.hh"

namespace libcppe {

inline int multipole_components(int k) { return (k + 1) * (k + 2) / 2; }  // math.cc:99

class Multipole {
 public:
  explicit Multipole(unsigned k) : m_k(k) {}
  void add_value(double v) { m_values.push_back(v); }
  void remove_trace() {  // potential.hh:23-34
    if (m_k == 2) {
      double trace = (m_values[0] + m_values[3] + m_values[5]) / 3.0;
      m_values[0] -= trace;
      m_values[3] -= trace;
      m_values[5] -= trace;
    } else if (m_k > 2) {
      throw std::runtime_error("remove_trace() not implemented for multipoles of order > 2");
    }
  }
  std::vector<double>& get_values() { return m_values; }
  const std::vector<double>& get_values() const { return m_values; }
  std::vector<double> get_values_vec() const { return m_values; }
  unsigned m_k;

 private:
  std::vector<double> m_values;
};

class Polarizability {
 public:
  Polarizability() { m_values.fill(0.0); }
  explicit Polarizability(const std::vector<double>& values) {
    if (values.size() != 6)
      throw std::runtime_error(
            "Invalid number of elements to construct polarizability. Vector must have exactly 6 elements.");
    m_values = triangle_to_mat(values);
  }
  void make_isotropic() {  // potential.hh:60-66
    double t = get_isotropic_value();
    m_values.fill(0.0);
    m_values[0] = m_values[4] = m_values[8] = t;
  }
  double get_isotropic_value() const { return (m_values[0] + m_values[4] + m_values[8]) / 3.0; }
  std::vector<double> get_values_vec() const { return mat_to_triangle(m_values); }
  const Matrix3& get_matrix() const { return m_values; }

 private:
  Matrix3 m_values;
};

class Potential {
 public:
  Potential(double x, double y, double z, std::string element, int idx)
        : m_x(x), m_y(y), m_z(z), m_element(std::move(element)), index(idx) {}
  double m_x, m_y, m_z;
  std::string m_element;
  int index;

  void add_multipole(const Multipole& m) { m_multipoles.push_back(m); }
  void add_exclusion(int e) { m_exclusions.push_back(e); }  // 0-based
  bool excludes_site(int other) const {
    return std::find(m_exclusions.begin(), m_exclusions.end(), other) != m_exclusions.end();
  }
  std::vector<int>& get_exclusions() { return m_exclusions; }
  const std::vector<int>& get_exclusions() const { return m_exclusions; }
  std::vector<Multipole>& get_multipoles() { return m_multipoles; }
  const std::vector<Multipole>& get_multipoles() const { return m_multipoles; }
  void set_polarizability(const Polarizability& p) {
    m_polarizability = p;
    m_is_polarizable = true;
  }
  int max_multipole_order() const {
    int k = 0;
    for (auto& m : m_multipoles)
      if ((int)m.m_k > k) k = (int)m.m_k;
    return k;
  }
  Polarizability& get_polarizability() { return m_polarizability; }
  const Polarizability& get_polarizability() const { return m_polarizability; }
  bool is_polarizable() const { return m_is_polarizable; }
  std::array<double, 3> get_site_position() const { return {m_x, m_y, m_z}; }

 private:
  std::vector<Multipole> m_multipoles;
  Polarizability m_polarizability;
  bool m_is_polarizable = false;
  std::vector<int> m_exclusions;
};

std::vector<Potential> get_polarizable_sites(const std::vector<Potential>& potentials);

}  // namespace libcppe
