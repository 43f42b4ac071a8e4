// Potential preprocessing (src/core/pot_manipulation.{hh,cc}): iso_pol and the QM/MM border
// "remove" option. The "redist" border option is not provided by this build.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "molecule.hh"
#include "pe_options.hh"
#include "potential.hh"

namespace libcppe {

class PotManipulator {
 public:
  PotManipulator(std::vector<Potential> potentials, Molecule mol)
        : m_potentials(std::move(potentials)), m_mol(std::move(mol)) {}
  void set_print_callback(std::function<void(std::string)> printer) { m_printer = std::move(printer); }
  std::vector<Potential> manipulate(const PeOptions& options);

 private:
  std::vector<Potential> m_potentials;
  Molecule m_mol;
  std::function<void(std::string)> m_printer = [](std::string) {};
};

}  // namespace libcppe
