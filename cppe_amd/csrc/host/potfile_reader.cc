// Restates the reference reader's behaviour (src/core/potfile_reader.cc:31-179): AA -> bohr with
// 1.8897261246, missing multipole sites filled with zeros, quadrupole trace removed on read,
// EXCLISTS header count, non-positive exclusion entries skipped.
#include "potfile_reader.hh"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace libcppe {

namespace {
constexpr double kAng2Bohr = 1.8897261246;

std::vector<std::string> tokens(const std::string& line) {
  std::istringstream is(line);
  std::vector<std::string> out;
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

std::string reduce(const std::string& line) {
  auto t = tokens(line);
  std::string r;
  for (size_t i = 0; i < t.size(); ++i) r += (i ? " " : "") + t[i];
  return r;
}

bool file_exists(const std::string& name) {
  std::ifstream f(name.c_str());
  return f.good();
}
}  // namespace

PotfileReader::PotfileReader(std::string potfile_name) : m_potfile(std::move(potfile_name)) {
  if (!file_exists(m_potfile)) throw std::runtime_error("Potential file does not exist.");
}

std::vector<Potential> PotfileReader::read() {
  std::ifstream infile(m_potfile);
  std::vector<Potential> potentials;
  std::string line;
  int num_sites = 0;
  while (std::getline(infile, line)) {
    if (line.find("@COORDINATES") != std::string::npos) {
      std::getline(infile, line);
      num_sites = std::stoi(tokens(line).at(0));
      std::getline(infile, line);
      std::string unit = reduce(line);
      double conversion;
      if (unit == "AA")
        conversion = kAng2Bohr;
      else if (unit == "AU")
        conversion = 1.0;
      else
        throw std::runtime_error("Invalid unit for potential file.");
      potentials.reserve(num_sites);
      for (int i = 0; i < num_sites; ++i) {
        std::getline(infile, line);
        auto t = tokens(line);
        if (t.size() < 4) throw std::runtime_error("Invalid coordinate line in potential file.");
        potentials.emplace_back(std::stod(t[1]) * conversion, std::stod(t[2]) * conversion,
                                std::stod(t[3]) * conversion, t[0], i);
      }
    }
    if (line.find("ORDER") != std::string::npos) {
      auto t = tokens(line);
      if (t.size() == 2) {
        int order = std::stoi(t[1]);
        std::getline(infile, line);
        int num_multipoles = std::stoi(line);
        int site_before    = -1;
        const int nc       = multipole_components(order);
        for (int n_mul = 0; n_mul < num_multipoles; ++n_mul) {
          std::getline(infile, line);
          t            = tokens(line);
          int site_num = std::stoi(t.at(0)) - 1;
          if (site_num != site_before + 1) {
            for (int d = 1; d < site_num - site_before; ++d) {
              Multipole mul(order);
              for (int v = 0; v < nc; ++v) mul.add_value(0.0);
              potentials.at(site_before + d).add_multipole(mul);
            }
          }
          Multipole mul(order);
          for (int v = 1; v <= nc; ++v) mul.add_value(std::stod(t.at(v)));
          mul.remove_trace();
          potentials.at(site_num).add_multipole(mul);
          site_before = site_num;
          if (n_mul == num_multipoles - 1 && site_num != num_sites - 1) {
            for (int d = 1; d < num_sites - site_num; ++d) {
              Multipole m0(order);
              for (int v = 0; v < nc; ++v) m0.add_value(0.0);
              potentials.at(site_num + d).add_multipole(m0);
            }
          }
        }
      } else if (t.size() == 3) {
        int order1 = std::stoi(t[1]), order2 = std::stoi(t[2]);
        if (order1 != 1 || order2 != 1)
          throw std::runtime_error("Only dipole-dipole polarizabilities are currently supported.");
        std::getline(infile, line);
        int num_pol = std::stoi(line);
        for (int p = 0; p < num_pol; ++p) {
          std::getline(infile, line);
          t            = tokens(line);
          int site_num = std::stoi(t.at(0)) - 1;
          std::vector<double> vals;
          for (int v = 1; v <= multipole_components(order1 + order2); ++v) vals.push_back(std::stod(t.at(v)));
          Polarizability pol{vals};
          if (potentials.at(site_num).is_polarizable())
            throw std::runtime_error("Potential can only have one polarizability.");
          potentials[site_num].set_polarizability(pol);
        }
      } else {
        throw std::runtime_error("Invalid number in potfile ORDER.");
      }
    }
    if (line.find("EXCLISTS") != std::string::npos) {
      std::getline(infile, line);
      int num_excl = std::stoi(tokens(line).at(0));
      for (int i = 0; i < num_excl; ++i) {
        std::getline(infile, line);
        auto t       = tokens(line);
        int site_num = std::stoi(t.at(0)) - 1;
        for (size_t k = 1; k < t.size(); ++k) {
          int e = std::stoi(t[k]) - 1;
          if (e < 0) continue;
          potentials.at(site_num).add_exclusion(e);
        }
      }
    }
  }
  return potentials;
}

std::vector<Potential> get_polarizable_sites(const std::vector<Potential>& potentials) {
  std::vector<Potential> r;  // electric_fields.cc:242-247
  for (const auto& p : potentials)
    if (p.is_polarizable()) r.push_back(p);
  return r;
}

}  // namespace libcppe
