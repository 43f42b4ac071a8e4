#include "device_system.hh"

namespace libcppe {

void hip_check(int rc) {
  if (rc == CPPE_HIP_OK) return;
  std::string msg = cppe_hip_last_error();
  if (rc == CPPE_HIP_ERR_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

cppe_hip_options make_options(const PeOptions& opt, double damping) {
  cppe_hip_options o{};
  // anything other than "fmm" is direct summation (bmatrix.cc:26-30)
  o.summation = opt.summation_induced_fields == "fmm" ? CPPE_HIP_SUM_FMM : CPPE_HIP_SUM_DIRECT;
  o.ncrit     = opt.tree_ncrit;
  o.order     = opt.tree_expansion_order;
  o.theta     = opt.theta;
  o.damping   = damping;
  if (opt.solver_precision == "fp64")
    o.precision = CPPE_HIP_PREC_FP64;
  else if (opt.solver_precision == "mixed")
    o.precision = CPPE_HIP_PREC_MIXED;
  else
    throw std::invalid_argument("solver_precision must be 'fp64' or 'mixed'.");
  return o;
}

DeviceSystem::DeviceSystem(const std::vector<Potential>& sites, const PeOptions& opt, double damping,
                           bool with_alpha) {
  m_n = (int)sites.size();
  std::vector<double> pos(3 * (size_t)m_n), a6, iso(m_n, 0.0);
  std::vector<int> off(m_n + 1, 0), idx;
  if (with_alpha) a6.resize(6 * (size_t)m_n);
  for (int i = 0; i < m_n; ++i) {
    const Potential& p = sites[i];
    pos[3 * i]         = p.m_x;
    pos[3 * i + 1]     = p.m_y;
    pos[3 * i + 2]     = p.m_z;
    if (p.is_polarizable()) {
      iso[i] = p.get_polarizability().get_isotropic_value();
      if (with_alpha) {
        auto v = p.get_polarizability().get_values_vec();
        for (int k = 0; k < 6; ++k) a6[6 * (size_t)i + k] = v[k];
      }
    }
    for (int e : p.get_exclusions()) idx.push_back(e);
    off[i + 1] = (int)idx.size();
  }
  if (idx.empty()) idx.push_back(-1);
  auto o = make_options(opt, damping);
  if (opt.solver_preconditioner != "jacobi" && opt.solver_preconditioner != "local")
    throw std::invalid_argument("solver_preconditioner must be 'jacobi' or 'local'.");
  hip_check(cppe_hip_system_create(&m_sys, m_n, pos.data(), with_alpha ? a6.data() : nullptr, iso.data(),
                                   off.data(), idx.data(), &o));
  if (with_alpha && opt.solver_preconditioner == "local")
    hip_check(cppe_hip_set_preconditioner(m_sys, CPPE_HIP_PRECOND_LOCAL, 5.6691783738, 0.8));
}

DeviceSystem::~DeviceSystem() {
  if (m_sys) cppe_hip_system_destroy(m_sys);
}

}  // namespace libcppe
