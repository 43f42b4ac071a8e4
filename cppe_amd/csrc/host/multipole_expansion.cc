#include "multipole_expansion.hh"

#include "device_system.hh"

namespace libcppe {

namespace {
// Site and atom arrays of the device calls: positions, highest order per site, 10 values per site
// (q, mu, packed quadrupole), atom charges and positions.
struct PackedSites {
  std::vector<double> pos, mul, ch, apos;
  std::vector<int> kmax;
};

PackedSites pack(const std::vector<Potential>& pots, const Molecule& mol) {
  const int n = (int)pots.size();
  PackedSites P;
  P.pos.resize(3 * (size_t)n);
  P.mul.assign(10 * (size_t)n, 0.0);
  P.kmax.assign(n, -1);
  std::vector<double>& pos = P.pos;
  std::vector<double>& mul = P.mul;
  std::vector<int>& kmax   = P.kmax;
  const std::vector<Potential>& m_potentials = pots;
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
  for (const Atom& a : mol) {
    P.ch.push_back((double)a.charge);
    P.apos.insert(P.apos.end(), {a.m_x, a.m_y, a.m_z});
  }
  return P;
}
}  // namespace

double MultipoleExpansion::interaction_energy() {  // multipole_expansion.cc:10-33
  const int n = (int)m_potentials.size();
  double e    = 0.0;
  if (n == 0 || m_mol.empty()) return e;
  PackedSites P = pack(m_potentials, m_mol);
  hip_check(cppe_hip_nuclear_multipole_energy(n, P.pos.data(), P.kmax.data(), P.mul.data(), (int)m_mol.size(),
                                              P.ch.data(), P.apos.data(), &e));
  return e;
}

Matrix MultipoleExpansion::nuclear_gradient() {  // multipole_expansion.cc:35-62
  const int n = (int)m_potentials.size();
  Matrix g(m_mol.size(), 3);
  if (n == 0 || m_mol.empty()) return g;
  PackedSites P = pack(m_potentials, m_mol);
  hip_check(cppe_hip_nuclear_multipole_gradient(n, P.pos.data(), P.kmax.data(), P.mul.data(), (int)m_mol.size(),
                                                P.ch.data(), P.apos.data(), g.data.data()));
  return g;
}

}  // namespace libcppe
