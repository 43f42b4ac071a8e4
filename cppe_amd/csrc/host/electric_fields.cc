#include "electric_fields.hh"

#include <iomanip>
#include <algorithm>
#include <sstream>
#include <stdexcept>

#include "bmatrix.hh"

namespace libcppe {

NuclearFields::NuclearFields(Molecule mol, std::vector<Potential> potentials)
      : m_mol(std::move(mol)), m_polsites(get_polarizable_sites(potentials)) {}

Vector NuclearFields::compute() {  // electric_fields.cc:58-75
  const int n = (int)m_polsites.size();
  std::vector<double> pos(3 * (size_t)n), ch, apos;
  for (int i = 0; i < n; ++i) {
    pos[3 * i]     = m_polsites[i].m_x;
    pos[3 * i + 1] = m_polsites[i].m_y;
    pos[3 * i + 2] = m_polsites[i].m_z;
  }
  for (const Atom& a : m_mol) {
    ch.push_back((double)a.charge);
    apos.insert(apos.end(), {a.m_x, a.m_y, a.m_z});
  }
  Vector F(3 * (size_t)n, 0.0);
  if (n == 0) return F;
  hip_check(cppe_hip_nuclear_fields(n, pos.data(), (int)m_mol.size(), ch.data(), apos.data(), F.data()));
  return F;
}

Matrix NuclearFields::nuclear_gradient() {  // electric_fields.cc:77-104
  const int n = (int)m_polsites.size();
  Matrix g(3 * m_mol.size(), 3 * (size_t)n);
  if (n == 0 || m_mol.empty()) return g;
  std::vector<double> pos(3 * (size_t)n), ch, apos;
  for (int i = 0; i < n; ++i) {
    pos[3 * i]     = m_polsites[i].m_x;
    pos[3 * i + 1] = m_polsites[i].m_y;
    pos[3 * i + 2] = m_polsites[i].m_z;
  }
  for (const Atom& a : m_mol) {
    ch.push_back((double)a.charge);
    apos.insert(apos.end(), {a.m_x, a.m_y, a.m_z});
  }
  hip_check(cppe_hip_nuclear_field_gradient(n, pos.data(), (int)m_mol.size(), ch.data(), apos.data(), g.data.data()));
  return g;
}

MultipoleFields::MultipoleFields(std::vector<Potential> potentials, const PeOptions& options)
      : m_potentials(std::move(potentials)), m_options(options) {}

Vector MultipoleFields::compute() {  // electric_fields.cc:106-189
  const int n = (int)m_potentials.size();
  std::vector<double> q(n, 0.0), mu(3 * (size_t)n, 0.0), th(6 * (size_t)n, 0.0);
  int max_order = 0;
  for (int i = 0; i < n; ++i) {
    const Potential& p = m_potentials[i];
    int k              = p.max_multipole_order();
    if (k > max_order) max_order = k;
    if (k > 2) throw std::runtime_error("MultipoleFields::compute_tree only supported up to 2nd order.");
    for (const Multipole& m : p.get_multipoles()) {
      const auto& v = m.get_values();
      if (m.m_k == 0) q[i] = v.at(0);
      if (m.m_k == 1)
        for (int c = 0; c < 3; ++c) mu[3 * (size_t)i + c] = v.at(c);
      if (m.m_k == 2)
        for (int c = 0; c < 6; ++c) th[6 * (size_t)i + c] = v.at(c);
    }
  }
  Vector F(3 * (size_t)n, 0.0);
  if (n == 0) return F;
  double damping = m_options.damp_multipole ? m_options.damping_factor_multipole : 0.0;
  DeviceSystem dev(m_potentials, m_options, damping, false);
  hip_check(cppe_hip_multipole_fields(dev.get(), q.data(), max_order > 0 ? mu.data() : nullptr,
                                      max_order > 1 ? th.data() : nullptr, F.data()));
  return F;
}

InducedMoments::InducedMoments(std::vector<Potential> potentials, const PeOptions& options)
      : m_polsites(get_polarizable_sites(potentials)), m_options(options) {}

BMatrix& InducedMoments::bmatrix() {
  if (!m_bmat) m_bmat = std::make_shared<BMatrix>(m_polsites, m_options);
  return *m_bmat;
}

Vector InducedMoments::compute(const Vector& rhs, Vector guess, bool make_guess) {
  const size_t n3 = 3 * m_polsites.size();
  if (rhs.size() != n3) throw std::invalid_argument("rhs size does not match 3 * number of polarizable sites");
  if (!make_guess && guess.size() != n3) throw std::invalid_argument("guess size does not match");
  Vector mu(n3, 0.0);
  if (n3 == 0) {  // the reference's loop takes one step on empty vectors: residual norm 0
    m_last_iters = 1;
    m_last_norms.assign(1, 0.0);
    m_printer("0 --- Norm: 0.000000000000");
    return mu;
  }
  m_last_norms.assign(std::max(m_options.maxiter, 1), 0.0);
  int iters = 0;
  int rc    = cppe_hip_solve(bmatrix().device().get(), rhs.data(), make_guess ? nullptr : guess.data(),
                             make_guess ? 1 : 0, m_options.induced_thresh, m_options.maxiter, mu.data(), &iters,
                             m_last_norms.data());
  m_last_iters = iters;
  m_last_norms.resize(iters);
  // one line per iteration, the reference's format (electric_fields.cc:222-225), replayed after
  // the device solve
  for (int k = 0; k < iters; ++k) {
    std::stringstream ss;
    ss.precision(12);
    ss << std::fixed << m_last_norms[k];
    m_printer(std::to_string(k) + " --- Norm: " + ss.str());
  }
  if (rc == CPPE_HIP_ERR_NOCONV) throw std::runtime_error("Failed to converge induced dipole moments.");
  hip_check(rc);
  return mu;
}

std::vector<Vector> InducedMoments::compute_multi(const std::vector<Vector>& rhs, bool make_guess) {
  const size_t n3 = 3 * m_polsites.size();
  const int k     = (int)rhs.size();
  if (k == 0) return {};
  if (n3 == 0) {  // as compute() on empty vectors, once per right-hand side
    std::vector<Vector> out;
    for (int a = 0; a < k; ++a) out.push_back(compute(rhs[a], Vector(), true));
    return out;
  }
  std::vector<double> r((size_t)k * n3), mu((size_t)k * n3, 0.0);
  for (int a = 0; a < k; ++a) {
    if (rhs[a].size() != n3) throw std::invalid_argument("rhs size does not match 3 * number of polarizable sites");
    std::copy(rhs[a].begin(), rhs[a].end(), r.begin() + (size_t)a * n3);
  }
  const int maxit = std::max(m_options.maxiter, 1);
  std::vector<int> iters(k, 0);
  std::vector<double> norms((size_t)k * maxit, 0.0);
  int rc = cppe_hip_solve_multi(bmatrix().device().get(), k, r.data(), nullptr, make_guess ? 1 : 0,
                                m_options.induced_thresh, m_options.maxiter, mu.data(), iters.data(), norms.data());
  for (int a = 0; a < k; ++a) {
    m_last_iters = iters[a];
    m_last_norms.assign(norms.begin() + (size_t)a * maxit, norms.begin() + (size_t)a * maxit + iters[a]);
    for (int it = 0; it < iters[a]; ++it) {
      std::stringstream ss;
      ss.precision(12);
      ss << std::fixed << m_last_norms[it];
      m_printer(std::to_string(it) + " --- Norm: " + ss.str());
    }
  }
  if (rc == CPPE_HIP_ERR_NOCONV) throw std::runtime_error("Failed to converge induced dipole moments.");
  hip_check(rc);
  std::vector<Vector> out(k);
  for (int a = 0; a < k; ++a) out[a].assign(mu.begin() + (size_t)a * n3, mu.begin() + (size_t)(a + 1) * n3);
  return out;
}

}  // namespace libcppe
