// Nuclear-multipole electrostatic energy (src/core/multipole_expansion.hh:8-22).
#pragma once
#include <vector>

#include "linalg.hh"
#include "molecule.hh"
#include "potential.hh"

namespace libcppe {

class MultipoleExpansion {
 public:
  MultipoleExpansion(Molecule core, std::vector<Potential> potentials)
        : m_mol(std::move(core)), m_potentials(std::move(potentials)) {}
  double interaction_energy();  // multipole_expansion.cc:10-33
  Matrix nuclear_gradient();    // multipole_expansion.cc:35-62, natoms x 3

 private:
  Molecule m_mol;
  std::vector<Potential> m_potentials;
};

}  // namespace libcppe
