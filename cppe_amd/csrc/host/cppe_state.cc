#include "cppe_state.hh"

#include <iomanip>
#include <map>
#include <numeric>
#include <sstream>

#include "device_system.hh"
#include "multipole_expansion.hh"
#include "pot_manipulation.hh"
#include "potfile_reader.hh"

namespace libcppe {

const PrintCallback default_printer = [](std::string s) { std::cout << std::setprecision(12) << s << std::endl; };

CppeState::CppeState(const PeOptions& options, Molecule mol, PrintCallback printer)
      : m_mol(std::move(mol)), m_options(options), m_printer(std::move(printer)) {  // cppe_state.cc:15-32
  std::vector<Potential> potentials = PotfileReader(m_options.potfile).read();
  PotManipulator manip(potentials, m_mol);
  manip.set_print_callback(m_printer);
  set_potentials(manip.manipulate(m_options));
  std::map<std::string, std::vector<std::string>> terms{{"Electrostatic", {"Electronic", "Nuclear", "Multipoles"}},
                                                        {"Polarization", {"Electronic", "Nuclear", "Multipoles"}}};
  for (auto& t : terms)
    for (auto& nm : t.second) m_pe_energy[t.first][nm] = 0.0;
}

void CppeState::set_potentials(std::vector<Potential> potentials) {  // cppe_state.cc:34-55
  m_potentials        = std::move(potentials);
  m_polarizable_sites = 0;
  for (auto& p : m_potentials) m_polarizable_sites += p.is_polarizable();
  m_induced_moments.assign(3 * m_polarizable_sites, 0.0);
  m_positions             = Matrix(m_potentials.size(), 3);
  m_positions_polarizable = Matrix(m_polarizable_sites, 3);
  size_t k                = 0;
  for (size_t i = 0; i < m_potentials.size(); ++i) {
    const Potential& p = m_potentials[i];
    m_positions(i, 0)  = p.m_x;
    m_positions(i, 1)  = p.m_y;
    m_positions(i, 2)  = p.m_z;
    if (p.is_polarizable()) {
      m_positions_polarizable(k, 0) = p.m_x;
      m_positions_polarizable(k, 1) = p.m_y;
      m_positions_polarizable(k, 2) = p.m_z;
      ++k;
    }
  }
  // the response operator belongs to the old sites; m_make_guess is left as it was (the
  // reference's set_potentials zeroes the moments only, so a later solve starts from zero)
  m_solver.reset();
}

void CppeState::calculate_static_energies_and_fields() {  // cppe_state.cc:57-71
  MultipoleExpansion mexp(m_mol, m_potentials);
  m_pe_energy["Electrostatic"]["Nuclear"] = mexp.interaction_energy();
  NuclearFields nfields(m_mol, m_potentials);
  m_nuc_fields = nfields.compute();
  MultipoleFields mfields(m_potentials, m_options);
  m_multipole_fields = mfields.compute();
}

InducedMoments& CppeState::solver() {
  if (!m_solver) m_solver = std::make_shared<InducedMoments>(m_potentials, m_options);
  m_solver->set_print_callback(m_printer);
  return *m_solver;
}

static double device_dot(const Vector& a, const Vector& b) {
  if (a.size() != b.size()) throw std::invalid_argument("vector sizes differ");
  double r = 0.0;
  hip_check(cppe_hip_dot((long long)a.size(), a.data(), b.data(), &r));
  return r;
}

void CppeState::update_induced_moments(Vector elec_fields, bool elec_only) {  // cppe_state.cc:83-109
  Vector total = elec_only ? elec_fields : elec_fields + m_nuc_fields + m_multipole_fields;
  m_induced_moments = solver().compute(total, m_induced_moments, m_make_guess);
  if (m_make_guess) m_make_guess = false;
  m_pe_energy["Polarization"]["Electronic"] = -0.5 * device_dot(m_induced_moments, elec_fields);
  if (!elec_only) {
    m_pe_energy["Polarization"]["Nuclear"]    = -0.5 * device_dot(m_induced_moments, m_nuc_fields);
    m_pe_energy["Polarization"]["Multipoles"] = -0.5 * device_dot(m_induced_moments, m_multipole_fields);
  }
}

Matrix CppeState::nuclear_interaction_energy_gradient() {  // cppe_state.cc:73-76
  return MultipoleExpansion(m_mol, m_potentials).nuclear_gradient();
}

Matrix CppeState::nuclear_field_gradient() {  // cppe_state.cc:78-81
  return NuclearFields(m_mol, m_potentials).nuclear_gradient();
}

Matrix CppeState::induced_moments_eef() {  // cppe_state.cc:111-128
  // the three unit-field solves as one device solve of three right-hand sides
  const size_t n = m_polarizable_sites;
  Matrix ret(3 * n, 3);
  std::vector<Vector> F(3, Vector(3 * n, 0.0));
  for (int a = 0; a < 3; ++a)
    for (size_t s = 0; s < n; ++s) F[a][3 * s + a] = 1.0;
  const std::vector<Vector> mu = solver().compute_multi(F, true);
  for (int a = 0; a < 3; ++a)
    for (size_t r = 0; r < 3 * n; ++r) ret(r, a) = mu[a][r];
  return ret;
}

double CppeState::get_total_energy_for_category(const std::string& category) {
  auto& cat = m_pe_energy[category];
  double acc = 0.0;
  for (auto& kv : cat) acc += kv.second;
  return acc;
}

double CppeState::get_total_energy() {
  return get_total_energy_for_category("Electrostatic") + get_total_energy_for_category("Polarization");
}

std::string CppeState::get_energy_summary_string() {  // cppe_state.cc:145-180
  size_t w = 30;
  std::string off(3, ' '), off2(6, ' ');
  std::stringstream ss;
  ss << std::string(2 * w + 10, '-') << std::endl;
  ss << "Polarizable Embedding Summary:";
  ss << std::setprecision(12) << std::endl << std::endl;
  ss << off << "Electrostatics:" << std::endl;
  ss << off2 << "Electronic:" << std::string(w - 11, ' ') << m_pe_energy["Electrostatic"]["Electronic"] << std::endl;
  ss << off2 << "Nuclear:" << std::string(w - 8, ' ') << m_pe_energy["Electrostatic"]["Nuclear"] << std::endl;
  ss << off2 << "Multipole:" << std::string(w - 10, ' ') << m_pe_energy["Electrostatic"]["Multipoles"] << std::endl;
  double es = get_total_energy_for_category("Electrostatic");
  ss << off2 << "Total:" << std::string(w - 6, ' ') << es << std::endl << std::endl;
  ss << off << "Polarization:" << std::endl;
  ss << off2 << "Electronic:" << std::string(w - 11, ' ') << m_pe_energy["Polarization"]["Electronic"] << std::endl;
  ss << off2 << "Nuclear:" << std::string(w - 8, ' ') << m_pe_energy["Polarization"]["Nuclear"] << std::endl;
  ss << off2 << "Multipole:" << std::string(w - 10, ' ') << m_pe_energy["Polarization"]["Multipoles"] << std::endl;
  double pol = get_total_energy_for_category("Polarization");
  ss << off2 << "Total:" << std::string(w - 6, ' ') << pol << std::endl << std::endl;
  ss << off << "Total Energy:" << std::string(w - 10, ' ') << es + pol << std::endl;
  ss << std::string(2 * w + 10, '-') << std::endl << std::endl;
  return ss.str();
}

}  // namespace libcppe
