// Static fields and the induced-moment solver (src/core/electric_fields.hh:14-80), computed by
// the gfx950 library.
#pragma once
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "device_system.hh"
#include "linalg.hh"
#include "molecule.hh"
#include "pe_options.hh"
#include "potential.hh"

namespace libcppe {

class NuclearFields {
 public:
  NuclearFields(Molecule mol, std::vector<Potential> potentials);
  Vector compute();  // 3 * n_polsites, AoS
  Matrix nuclear_gradient();  // electric_fields.cc:77-104, (3 natoms) x (3 n_polsites)

 private:
  Molecule m_mol;
  std::vector<Potential> m_polsites;
};

class MultipoleFields {
 public:
  MultipoleFields(std::vector<Potential> potentials, const PeOptions& options);
  Vector compute();  // 3 * n_sites, AoS: F(q) - F(mu) + F(theta/2 diag)

 private:
  std::vector<Potential> m_potentials;
  PeOptions m_options;
};

class BMatrix;

class InducedMoments {
 public:
  InducedMoments(std::vector<Potential> potentials, const PeOptions& options);
  void set_print_callback(std::function<void(std::string)> printer) { m_printer = std::move(printer); }
  // electric_fields.cc:191-239
  Vector compute(const Vector& rhs, Vector guess, bool make_guess);
  Vector compute(const Vector& total_fields, bool make_guess) {
    return compute(total_fields, Vector(total_fields.size(), 0.0), make_guess);
  }
  // several right-hand sides solved together (one device CG sharing every B apply); the printer
  // receives each right-hand side's lines in turn, as consecutive compute calls would print them
  std::vector<Vector> compute_multi(const std::vector<Vector>& rhs, bool make_guess);
  // iterations and residual norms of the last solve
  int last_iterations() const { return m_last_iters; }
  const std::vector<double>& last_norms() const { return m_last_norms; }
  BMatrix& bmatrix();

 private:
  std::vector<Potential> m_polsites;
  PeOptions m_options;
  std::function<void(std::string)> m_printer = [](std::string s) { std::cout << s << std::endl; };
  std::shared_ptr<BMatrix> m_bmat;
  int m_last_iters = 0;
  std::vector<double> m_last_norms;
};

}  // namespace libcppe
