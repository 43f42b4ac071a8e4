// pybind11 module "pycppe" with the reference's Python names and arguments
// (src/python_iface/export_*.cc), served by the gfx950 library.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>

#include <algorithm>

#include "libcppe.hh"

namespace py = pybind11;
using namespace libcppe;

using MapSD  = std::unordered_map<std::string, double>;
using MapSSD = std::unordered_map<std::string, std::unordered_map<std::string, double>>;
PYBIND11_MAKE_OPAQUE(MapSD);
PYBIND11_MAKE_OPAQUE(MapSSD);
PYBIND11_MAKE_OPAQUE(std::vector<libcppe::Atom>);

namespace {

py::array_t<double> to_np(const std::vector<double>& v) {
  py::array_t<double> a(v.size());
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

py::array_t<double> to_np(const Matrix& m) {
  py::array_t<double> a({m.rows, m.cols});
  std::copy(m.data.begin(), m.data.end(), a.mutable_data());
  return a;
}

template <size_t N>
py::array_t<double> to_np(const std::array<double, N>& v) {
  py::array_t<double> a(N);
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

Vector from_np(py::array_t<double, py::array::c_style | py::array::forcecast> a) {
  return Vector(a.data(), a.data() + a.size());
}

bool is_key(const std::vector<std::string>& keys, const std::string& k) {
  return std::find(keys.begin(), keys.end(), k) != keys.end();
}

// export_state.cc:21-69: unknown keys -> std::invalid_argument (ValueError in Python)
PeOptions dict_to_options(const py::dict& d) {
  PeOptions o;
  for (const auto& item : d) {
    std::string key = item.first.cast<std::string>();
    auto v          = item.second;
    if (!is_key(valid_option_keys, key) && !is_key(extension_option_keys, key))
      throw std::invalid_argument("Option key '" + key + "' is invalid.");
    if (key == "potfile") o.potfile = v.cast<std::string>();
    else if (key == "iso_pol") o.iso_pol = v.cast<bool>();
    else if (key == "induced_thresh") o.induced_thresh = v.cast<double>();
    else if (key == "maxiter") o.maxiter = v.cast<int>();
    else if (key == "damp_induced") o.damp_induced = v.cast<bool>();
    else if (key == "damp_multipole") o.damp_multipole = v.cast<bool>();
    else if (key == "damping_factor_induced") o.damping_factor_induced = v.cast<double>();
    else if (key == "damping_factor_multipole") o.damping_factor_multipole = v.cast<double>();
    else if (key == "summation_induced_fields") o.summation_induced_fields = v.cast<std::string>();
    else if (key == "tree_ncrit") o.tree_ncrit = v.cast<int>();
    else if (key == "tree_expansion_order") o.tree_expansion_order = v.cast<int>();
    else if (key == "theta") o.theta = v.cast<double>();
    else if (key == "pe_border") o.pe_border = v.cast<bool>();
    else if (key == "border_type") o.border_type = v.cast<std::string>();
    else if (key == "border_rmin") o.border_rmin = v.cast<double>();
    else if (key == "border_nredist") o.border_nredist = v.cast<int>();
    else if (key == "border_redist_order") o.border_redist_order = v.cast<int>();
    else if (key == "border_redist_pol") o.border_redist_pol = v.cast<bool>();
    else if (key == "solver_precision") o.solver_precision = v.cast<std::string>();
    else if (key == "solver_preconditioner") o.solver_preconditioner = v.cast<std::string>();
  }
  return o;
}

py::dict options_to_dict(const PeOptions& o) {  // export_state.cc:78-123
  py::dict d;
  d["potfile"]                  = o.potfile;
  d["iso_pol"]                  = o.iso_pol;
  d["induced_thresh"]           = o.induced_thresh;
  d["maxiter"]                  = o.maxiter;
  d["damp_induced"]             = o.damp_induced;
  d["damp_multipole"]           = o.damp_multipole;
  d["damping_factor_induced"]   = o.damping_factor_induced;
  d["damping_factor_multipole"] = o.damping_factor_multipole;
  d["summation_induced_fields"] = o.summation_induced_fields;
  d["tree_expansion_order"]     = o.tree_expansion_order;
  d["tree_ncrit"]               = o.tree_ncrit;
  d["theta"]                    = o.theta;
  d["pe_border"]                = o.pe_border;
  d["border_type"]              = o.border_type;
  d["border_rmin"]              = o.border_rmin;
  d["border_nredist"]           = o.border_nredist;
  d["border_redist_order"]      = o.border_redist_order;
  d["border_redist_pol"]        = o.border_redist_pol;
  if (o.solver_precision != "fp64") d["solver_precision"] = o.solver_precision;
  if (o.solver_preconditioner != "jacobi") d["solver_preconditioner"] = o.solver_preconditioner;
  return d;
}

}  // namespace

PYBIND11_MODULE(pycppe, m) {
  m.doc()                 = "CPPE polarizable-embedding API on the MI355X (gfx950) HIP library";
  m.attr("__build_type__") = "Release";
  m.attr("__parallel__")   = true;

  // molecule (export_molecule.cc)
  py::class_<Atom>(m, "Atom", "Atom class")
        .def(py::init<int, double, double, double>())
        .def_readwrite("atomic_number", &Atom::atomic_number)
        .def_readwrite("charge", &Atom::charge)
        .def_readwrite("x", &Atom::m_x)
        .def_readwrite("y", &Atom::m_y)
        .def_readwrite("z", &Atom::m_z)
        .def_property_readonly("position", [](const Atom& a) { return to_np(a.get_position()); });
  py::bind_vector<std::vector<Atom>>(m, "Molecule");

  // multipole / potential (export_potential.cc)
  py::class_<Multipole>(m, "Multipole")
        .def(py::init<unsigned>())
        .def_readwrite("k", &Multipole::m_k)
        .def_property_readonly("values", [](const Multipole& mm) { return to_np(mm.get_values()); })
        .def("remove_trace", &Multipole::remove_trace)
        .def("add_value", &Multipole::add_value);
  py::class_<Polarizability>(m, "Polarizability")
        .def(py::init<>())
        .def_property_readonly("values", [](const Polarizability& p) { return to_np(p.get_values_vec()); })
        .def_property_readonly("isotropic_value", &Polarizability::get_isotropic_value);
  py::class_<Potential>(m, "Potential", "Potential (Site)")
        .def(py::init<double, double, double, std::string, int>())
        .def_readwrite("x", &Potential::m_x, "x coordinate")
        .def_readwrite("y", &Potential::m_y, "y coordinate")
        .def_readwrite("z", &Potential::m_z, "z coordinate")
        .def_readwrite("element", &Potential::m_element, "element")
        .def_property_readonly("is_polarizable", &Potential::is_polarizable)
        .def("excludes_site", &Potential::excludes_site)
        .def_property_readonly("exclusions", [](const Potential& p) { return p.get_exclusions(); })
        .def_property_readonly("position", [](const Potential& p) { return to_np(p.get_site_position()); })
        .def_property_readonly("multipoles", [](const Potential& p) { return p.get_multipoles(); })
        .def_property_readonly("polarizability", [](const Potential& p) { return p.get_polarizability(); })
        .def_readwrite("index", &Potential::index, "site index");

  py::class_<PotfileReader>(m, "PotfileReader").def(py::init<std::string>()).def("read", &PotfileReader::read);

  // state (export_state.cc)
  py::bind_map<MapSD>(m, "UnorderedMapStringDouble");
  py::bind_map<MapSSD>(m, "UnorderedMapMapStringDouble");
  py::class_<CppeState, std::shared_ptr<CppeState>>(m, "CppeState")
        .def(py::init([](py::dict options, Molecule mol, PrintCallback printer) {
               return std::make_shared<CppeState>(dict_to_options(options), std::move(mol), std::move(printer));
             }),
             "Create a CppeState using a dictionary with options, molecule, and print callback",
             py::arg("options") = py::dict(), py::arg("molecule") = Molecule(),
             py::arg("printer") = default_printer)
        .def_property_readonly("positions", [](const CppeState& s) { return to_np(s.get_positions()); })
        .def_property_readonly("positions_polarizable",
                               [](const CppeState& s) { return to_np(s.get_positions_polarizable()); })
        .def("set_potentials", &CppeState::set_potentials)
        .def("calculate_static_energies_and_fields", &CppeState::calculate_static_energies_and_fields)
        .def("get_induced_moments", [](const CppeState& s) { return to_np(s.get_induced_moments()); })
        .def("induced_moments_eef", [](CppeState& s) { return to_np(s.induced_moments_eef()); })
        .def("nuclear_interaction_energy_gradient",
             [](CppeState& s) { return to_np(s.nuclear_interaction_energy_gradient()); })
        .def("nuclear_field_gradient", [](CppeState& s) { return to_np(s.nuclear_field_gradient()); })
        .def_readwrite("energies", &CppeState::m_pe_energy)
        .def_property_readonly("total_energy", &CppeState::get_total_energy)
        .def(
              "update_induced_moments",
              [](CppeState& s, py::array_t<double, py::array::c_style | py::array::forcecast> f, bool elec_only) {
                s.update_induced_moments(from_np(f), elec_only);
              },
              py::arg("elec_fields"), py::arg("elec_only") = false)
        .def_property_readonly("static_fields", [](const CppeState& s) { return to_np(s.get_static_fields()); })
        .def_property_readonly("nuclear_fields", [](const CppeState& s) { return to_np(s.get_nuclear_fields()); })
        .def_property_readonly("multipole_fields",
                               [](const CppeState& s) { return to_np(s.get_multipole_fields()); })
        .def_property_readonly("summary_string", &CppeState::get_energy_summary_string)
        .def_property_readonly("potentials", &CppeState::get_potentials)
        .def("get_polarizable_site_number", &CppeState::get_polarizable_site_number)
        .def_property_readonly("options", [](const CppeState& s) { return options_to_dict(s.get_options()); })
        .def_property_readonly("last_iterations", &CppeState::last_iterations);
  m.attr("valid_option_keys") = py::cast(valid_option_keys);

  // fields (export_fields.cc)
  py::class_<NuclearFields>(m, "NuclearFields", "Electric fields created by nuclei")
        .def(py::init<Molecule, std::vector<Potential>>())
        .def("compute", [](NuclearFields& f) { return to_np(f.compute()); })
        .def("nuclear_gradient", [](NuclearFields& f) { return to_np(f.nuclear_gradient()); });
  py::class_<MultipoleFields, std::shared_ptr<MultipoleFields>>(m, "MultipoleFields",
                                                                "Electric fields created by multipoles")
        .def(py::init([](std::vector<Potential> p, py::dict o) {
          return std::make_shared<MultipoleFields>(std::move(p), dict_to_options(o));
        }))
        .def("compute", [](MultipoleFields& f) { return to_np(f.compute()); });
  py::class_<InducedMoments, std::shared_ptr<InducedMoments>>(m, "InducedMoments")
        .def(py::init([](std::vector<Potential> p, py::dict o) {
          return std::make_shared<InducedMoments>(std::move(p), dict_to_options(o));
        }))
        .def(
              "compute",
              [](InducedMoments& im, py::array_t<double, py::array::c_style | py::array::forcecast> f,
                 bool make_guess) { return to_np(im.compute(from_np(f), make_guess)); },
              "Compute the induced moments solving the classical response equation using a CG solver",
              py::arg("total_fields"), py::arg("make_guess"))
        .def("set_print_callback", &InducedMoments::set_print_callback)
        .def_property_readonly("last_iterations", &InducedMoments::last_iterations);
  m.def("get_polarizable_sites", &get_polarizable_sites);
  py::class_<BMatrix, std::shared_ptr<BMatrix>>(m, "BMatrix")
        .def(py::init([](std::vector<Potential> p, py::dict o) {
          return std::make_shared<BMatrix>(std::move(p), dict_to_options(o));
        }))
        .def("direct_inverse", [](BMatrix& b) { return to_np(b.direct_inverse()); })
        .def("to_dense_matrix", [](BMatrix& b) { return to_np(b.to_dense_matrix()); })
        .def("apply", [](BMatrix& b, py::array_t<double, py::array::c_style | py::array::forcecast> x) {
          return to_np(b.apply(from_np(x)));
        })
        .def("apply_diagonal", [](BMatrix& b, py::array_t<double, py::array::c_style | py::array::forcecast> x) {
          return to_np(b.apply_diagonal(from_np(x)));
        })
        .def("apply_diagonal_inverse",
             [](BMatrix& b, py::array_t<double, py::array::c_style | py::array::forcecast> x) {
               return to_np(b.apply_diagonal_inverse(from_np(x)));
             });
  py::class_<MultipoleExpansion, std::shared_ptr<MultipoleExpansion>>(m, "MultipoleExpansion")
        .def(py::init<Molecule, std::vector<Potential>>())
        .def("interaction_energy", &MultipoleExpansion::interaction_energy)
        .def("nuclear_gradient", [](MultipoleExpansion& e) { return to_np(e.nuclear_gradient()); });
}
