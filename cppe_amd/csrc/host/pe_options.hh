// Options of the PE path; same fields, defaults and key list as the reference
// (src/core/pe_options.hh:8-69).
#pragma once
#include <string>
#include <vector>

namespace libcppe {

struct PeOptions {
  std::string potfile{"potential.pot"};
  bool iso_pol = false;
  double induced_thresh = 1e-8;
  int maxiter = 50;
  bool damp_induced = false;
  double damping_factor_induced = 2.1304;
  bool damp_multipole = false;
  double damping_factor_multipole = 2.1304;
  std::string summation_induced_fields{"direct"};  // "direct" or "fmm"
  int tree_ncrit = 64;
  int tree_expansion_order = 5;
  double theta = 0.5;
  bool pe_border = false;
  std::string border_type{"remove"};
  double border_rmin = 2.2;
  int border_nredist = -1;
  int border_redist_order = 1;
  bool border_redist_pol = false;
  // Arithmetic of the induced-dipole solve on the device: "fp64" (plain form) or "mixed"
  // (FP32 field applies inside FP64 iterative refinement; same FP64 stopping test).
  // No counterpart in the reference (see extension_option_keys).
  std::string solver_precision{"fp64"};
  // CG preconditioner of the induced-dipole solve: "jacobi" (alpha r, the reference's) or "local"
  // (alpha r plus the first Neumann correction over pairs closer than 3 A; cppe_hip.h).
  // No counterpart in the reference (see extension_option_keys).
  std::string solver_preconditioner{"jacobi"};
};

static const std::vector<std::string> valid_option_keys{"potfile",
                                                        "iso_pol",
                                                        "induced_thresh",
                                                        "maxiter",
                                                        "damp_induced",
                                                        "damping_factor_induced",
                                                        "damp_multipole",
                                                        "damping_factor_multipole",
                                                        "summation_induced_fields",
                                                        "tree_expansion_order",
                                                        "tree_ncrit",
                                                        "theta",
                                                        "pe_border",
                                                        "border_type",
                                                        "border_rmin",
                                                        "border_nredist",
                                                        "border_redist_order",
                                                        "border_redist_pol"};

// Keys this build accepts beyond the reference's list (valid_option_keys stays identical to it).
static const std::vector<std::string> extension_option_keys{"solver_precision", "solver_preconditioner"};

}  // namespace libcppe
