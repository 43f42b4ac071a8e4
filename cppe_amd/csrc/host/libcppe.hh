// Umbrella header of the host API (the reference's src/libcppe.hh).
#pragma once
#include "bmatrix.hh"
#include "cppe_state.hh"
#include "electric_fields.hh"
#include "molecule.hh"
#include "multipole_expansion.hh"
#include "pe_options.hh"
#include "pot_manipulation.hh"
#include "potential.hh"
#include "potfile_reader.hh"
