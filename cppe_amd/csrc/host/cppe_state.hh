// Stateful PE API (src/core/cppe_state.{hh,cc}); fields, energies and the induced-moment solve
// run on the gfx950 library. Between update_induced_moments calls the response operator (sites,
// octree, interaction lists) stays resident on the device and the previous moments warm-start CG.
#pragma once
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <unordered_map>

#include "electric_fields.hh"
#include "linalg.hh"
#include "molecule.hh"
#include "pe_options.hh"
#include "potential.hh"

namespace libcppe {

using PeEnergy      = std::unordered_map<std::string, std::unordered_map<std::string, double>>;
using PrintCallback = std::function<void(std::string)>;

extern const PrintCallback default_printer;

class CppeState {
 public:
  CppeState() = default;
  explicit CppeState(const PeOptions& options, Molecule mol, PrintCallback printer = default_printer);

  void set_options(const PeOptions& options) { m_options = options; }
  PeOptions get_options() const { return m_options; }
  void set_molecule(Molecule mol) { m_mol = std::move(mol); }
  void set_potentials(std::vector<Potential> potentials);
  std::vector<Potential> get_potentials() const { return m_potentials; }
  Matrix get_positions() const { return m_positions; }
  Matrix get_positions_polarizable() const { return m_positions_polarizable; }

  void calculate_static_energies_and_fields();
  std::vector<double> get_induced_moments() const { return m_induced_moments; }
  Vector get_induced_moments_vec() const { return m_induced_moments; }
  void update_induced_moments(Vector elec_fields, bool elec_only = false);
  Matrix induced_moments_eef();
  Matrix nuclear_interaction_energy_gradient();  // cppe_state.cc:73-76, natoms x 3
  Matrix nuclear_field_gradient();               // cppe_state.cc:78-81, (3 natoms) x (3 n_polsites)

  size_t get_polarizable_site_number() const { return m_polarizable_sites; }
  Vector get_static_fields() const { return m_nuc_fields + m_multipole_fields; }
  Vector get_nuclear_fields() const { return m_nuc_fields; }
  Vector get_multipole_fields() const { return m_multipole_fields; }

  PeEnergy m_pe_energy;
  double get_total_energy_for_category(const std::string& category);
  double get_total_energy();
  std::string get_energy_summary_string();

  // iterations / residual norms of the last induced-moment solve
  int last_iterations() const { return m_solver ? m_solver->last_iterations() : 0; }

 private:
  Molecule m_mol;
  std::vector<Potential> m_potentials;
  size_t m_polarizable_sites = 0;
  Vector m_nuc_fields, m_multipole_fields, m_induced_moments;
  Matrix m_positions, m_positions_polarizable;
  PeOptions m_options;
  bool m_make_guess = true;
  PrintCallback m_printer = default_printer;
  std::shared_ptr<InducedMoments> m_solver;
  InducedMoments& solver();
};

}  // namespace libcppe
