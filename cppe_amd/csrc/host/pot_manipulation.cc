#include "pot_manipulation.hh"

#include <cmath>
#include <set>
#include <stdexcept>

namespace libcppe {

std::vector<Potential> PotManipulator::manipulate(const PeOptions& opt) {
  if (opt.iso_pol)  // pot_manipulation.cc:15-22
    for (Potential& p : m_potentials) p.get_polarizability().make_isotropic();
  if (!opt.pe_border) return m_potentials;
  std::set<int> changed;  // pot_manipulation.cc:27-38
  for (const Atom& a : m_mol)
    for (const Potential& p : m_potentials) {
      double dx = a.m_x - p.m_x, dy = a.m_y - p.m_y, dz = a.m_z - p.m_z;
      if (std::sqrt(dx * dx + dy * dy + dz * dz) <= opt.border_rmin) changed.insert(p.index);
    }
  if (changed.empty()) return m_potentials;
  if (opt.border_type != "remove")
    throw std::runtime_error("pe_border with border_type '" + opt.border_type + "' is not supported by this build");
  std::vector<Potential> out;  // pot_manipulation.cc:40-48
  for (const Potential& p : m_potentials) {
    if (changed.count(p.index))
      m_printer("     Removing all parameters on site: " + std::to_string(p.index));
    else
      out.push_back(p);
  }
  return out;
}

}  // namespace libcppe
