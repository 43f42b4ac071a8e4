// QM core atoms (src/core/molecule.hh:9-21): charge = atomic number.
#pragma once
#include <array>
#include <vector>

namespace libcppe {

struct Atom {
  int atomic_number;
  int charge;
  double m_x = 0.0, m_y = 0.0, m_z = 0.0;
  explicit Atom(int an) : atomic_number(an), charge(an) {}
  Atom(int an, double x, double y, double z) : atomic_number(an), charge(an), m_x(x), m_y(y), m_z(z) {}
  std::array<double, 3> get_position() const { return {m_x, m_y, m_z}; }
};

using Molecule = std::vector<Atom>;

}  // namespace libcppe
