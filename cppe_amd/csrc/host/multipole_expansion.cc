#include "multipole_expansion.hh"

#include "device_system.hh"

namespace libcppe {

double MultipoleExpansion::interaction_energy() {
  const int n = (int)m_potentials.size();
  std::vector<double> pos(3 * (size_t)n), mul(10 * (size_t)n, 0.0), ch, apos;
  std::vector<int> kmax(n, -1);
  for (int i = 0; i < n; ++i) {
    const Potential& p = m_potentials[i];
    pos[3 * i]         = p.m_x;
    pos[3 * i + 1]     = p.m_y;
    pos[3 * i + 2]     = p.m_z;
    for (const Multipole& m : p.get_multipoles()) {
      if (m.m_k > 2) throw std::runtime_error("multipoles above order 2 are not supported");
      const int off = m.m_k == 0 ? 0 : (m.m_k == 1 ? 1 : 4);
      const auto& v = m.get_values();
      for (int c = 0; c < multipole_components(m.m_k); ++c) mul[10 * (size_t)i + off + c] = v.at(c);
      if ((int)m.m_k > kmax[i]) kmax[i] = (int)m.m_k;
    }
  }
  for (const Atom& a : m_mol) {
    ch.push_back((double)a.charge);
    apos.insert(apos.end(), {a.m_x, a.m_y, a.m_z});
  }
  double e = 0.0;
  if (n == 0 || m_mol.empty()) return e;
  hip_check(cppe_hip_nuclear_multipole_energy(n, pos.data(), kmax.data(), mul.data(), (int)m_mol.size(), ch.data(),
                                              apos.data(), &e));
  return e;
}

}  // namespace libcppe
