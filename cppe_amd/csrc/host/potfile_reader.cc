// Restates the reference reader's behaviour (src/core/potfile_reader.cc:31-179): AA -> bohr with
// 1.8897261246, missing multipole sites filled with zeros, quadrupole trace removed on read,
// EXCLISTS header count, non-positive exclusion entries skipped.
//
// The file is read whole, split into lines once, and the data lines of each section are parsed in
// parallel (std::thread, at most 16 workers) with std::from_chars, which rounds exactly as the
// reference's std::stod / std::stoi; the section walk and everything order-dependent (zero fill
// of absent multipole sites, duplicate polarizability check, exclusion order) stay sequential.
#include "potfile_reader.hh"

#include <algorithm>
#include <charconv>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string_view>
#include <thread>

namespace libcppe {

namespace {
constexpr double kAng2Bohr = 1.8897261246;

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

// Whitespace tokenizer over one line (the reference splits with istringstream >>).
struct Tok {
  const char* p;
  const char* e;
  Tok(std::string_view s) : p(s.data()), e(s.data() + s.size()) {}
  bool next(std::string_view& t) {
    while (p < e && is_space(*p)) ++p;
    if (p == e) return false;
    const char* b = p;
    while (p < e && !is_space(*p)) ++p;
    t = std::string_view(b, (size_t)(p - b));
    return true;
  }
  std::string_view need() {
    std::string_view t;
    if (!next(t)) throw std::invalid_argument("missing value in potential file");
    return t;
  }
};

// std::stod / std::stoi semantics on one token: optional '+', longest numeric prefix, throws
// std::invalid_argument when nothing converts.
double to_double(std::string_view t) {
  const char* b = t.data();
  const char* e = b + t.size();
  if (b < e && *b == '+') ++b;
  double v = 0.0;
  const auto r = std::from_chars(b, e, v);
  if (r.ec == std::errc::invalid_argument) throw std::invalid_argument("stod");
  if (r.ec == std::errc::result_out_of_range) throw std::out_of_range("stod");
  return v;
}

int to_int(std::string_view t) {
  const char* b = t.data();
  const char* e = b + t.size();
  if (b < e && *b == '+') ++b;
  int v       = 0;
  const auto r = std::from_chars(b, e, v);
  if (r.ec == std::errc::invalid_argument) throw std::invalid_argument("stoi");
  if (r.ec == std::errc::result_out_of_range) throw std::out_of_range("stoi");
  return v;
}

std::string reduce(std::string_view line) {
  Tok tk(line);
  std::string r;
  std::string_view t;
  while (tk.next(t)) {
    if (!r.empty()) r += ' ';
    r.append(t.data(), t.size());
  }
  return r;
}

bool file_exists(const std::string& name) {
  std::ifstream f(name.c_str());
  return f.good();
}

// Lines of the whole file, as std::getline returns them ('\n' removed).
struct Lines {
  std::string buf;
  std::vector<size_t> start;  // nlines + 1 (the last entry is one past the final '\n')
  size_t size() const { return start.size() - 1; }
  std::string_view operator[](size_t k) const {
    const size_t b = start[k], e = start[k + 1] - 1;
    return std::string_view(buf.data() + b, e - b);
  }
};

Lines load(const std::string& path) {
  Lines L;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Potential file does not exist.");
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  L.buf.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = L.buf.empty() ? 0 : std::fread(&L.buf[0], 1, L.buf.size(), f);
  std::fclose(f);
  L.buf.resize(got);
  if (!L.buf.empty() && L.buf.back() != '\n') L.buf.push_back('\n');
  L.start.push_back(0);
  for (size_t p = 0; p < L.buf.size();) {
    const void* q = std::memchr(L.buf.data() + p, '\n', L.buf.size() - p);
    const size_t e = (size_t)((const char*)q - L.buf.data());
    L.start.push_back(e + 1);
    p = e + 1;
  }
  return L;
}

// fn(k) for k in [0, n), in contiguous blocks over up to 16 threads; the first exception is
// rethrown on the calling thread.
template <typename Fn>
void parallel_for(size_t n, Fn&& fn) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({hw, 16, n / 2048 + 1});
  if (nt <= 1) {
    for (size_t k = 0; k < n; ++k) fn(k);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nt);
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      try {
        const size_t b = n * t / nt, e = n * (t + 1) / nt;
        for (size_t k = b; k < e; ++k) fn(k);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

bool has(std::string_view line, const char* key) { return line.find(key) != std::string_view::npos; }
}  // namespace

PotfileReader::PotfileReader(std::string potfile_name) : m_potfile(std::move(potfile_name)) {
  if (!file_exists(m_potfile)) throw std::runtime_error("Potential file does not exist.");
}

std::vector<Potential> PotfileReader::read() {
  const Lines L = load(m_potfile);
  std::vector<Potential> potentials;
  int num_sites = 0;
  size_t k      = 0;
  auto getline  = [&](std::string_view& line) {
    if (k >= L.size()) throw std::runtime_error("Unexpected end of potential file.");
    line = L[k++];
  };
  std::string_view line;
  while (k < L.size()) {
    line = L[k++];
    if (has(line, "@COORDINATES")) {
      getline(line);
      num_sites = to_int(Tok(line).need());
      getline(line);
      const std::string unit = reduce(line);
      double conversion;
      if (unit == "AA")
        conversion = kAng2Bohr;
      else if (unit == "AU")
        conversion = 1.0;
      else
        throw std::runtime_error("Invalid unit for potential file.");
      if (num_sites < 0 || k + (size_t)num_sites > L.size())
        throw std::runtime_error("Unexpected end of potential file.");
      std::vector<double> xyz(3 * (size_t)num_sites);
      std::vector<std::string_view> label(num_sites);
      const size_t k0 = k;
      parallel_for((size_t)num_sites, [&](size_t i) {
        Tok tk(L[k0 + i]);
        std::string_view t[4];
        for (int c = 0; c < 4; ++c)
          if (!tk.next(t[c])) throw std::runtime_error("Invalid coordinate line in potential file.");
        label[i] = t[0];
        for (int c = 0; c < 3; ++c) xyz[3 * i + c] = to_double(t[1 + c]) * conversion;
      });
      k += (size_t)num_sites;
      if (num_sites > 0) line = L[k - 1];
      potentials.reserve(potentials.size() + (size_t)num_sites);
      for (int i = 0; i < num_sites; ++i)
        potentials.emplace_back(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], std::string(label[i]), i);
    }
    if (has(line, "ORDER")) {
      std::vector<std::string_view> t;
      {
        Tok tk(line);
        std::string_view s;
        while (tk.next(s)) t.push_back(s);
      }
      if (t.size() == 2) {
        const int order = to_int(t[1]);
        getline(line);
        const int num_multipoles = to_int(Tok(line).need());
        const int nc             = multipole_components(order);
        if (num_multipoles < 0 || k + (size_t)num_multipoles > L.size())
          throw std::runtime_error("Unexpected end of potential file.");
        std::vector<int> site(num_multipoles);
        std::vector<double> vals((size_t)num_multipoles * nc);
        const size_t k0 = k;
        parallel_for((size_t)num_multipoles, [&](size_t m) {
          Tok tk(L[k0 + m]);
          site[m] = to_int(tk.need()) - 1;
          for (int v = 0; v < nc; ++v) vals[m * nc + v] = to_double(tk.need());
        });
        k += (size_t)num_multipoles;
        if (num_multipoles > 0) line = L[k - 1];
        auto zero = [&](int s) {
          Multipole mul(order);
          for (int v = 0; v < nc; ++v) mul.add_value(0.0);
          potentials.at(s).add_multipole(mul);
        };
        int site_before = -1;
        for (int m = 0; m < num_multipoles; ++m) {
          const int site_num = site[m];
          if (site_num != site_before + 1)
            for (int d = 1; d < site_num - site_before; ++d) zero(site_before + d);
          Multipole mul(order);
          for (int v = 0; v < nc; ++v) mul.add_value(vals[(size_t)m * nc + v]);
          mul.remove_trace();
          potentials.at(site_num).add_multipole(mul);
          site_before = site_num;
          if (m == num_multipoles - 1 && site_num != num_sites - 1)
            for (int d = 1; d < num_sites - site_num; ++d) zero(site_num + d);
        }
      } else if (t.size() == 3) {
        const int order1 = to_int(t[1]), order2 = to_int(t[2]);
        if (order1 != 1 || order2 != 1)
          throw std::runtime_error("Only dipole-dipole polarizabilities are currently supported.");
        getline(line);
        const int num_pol = to_int(Tok(line).need());
        const int nc      = multipole_components(order1 + order2);
        if (num_pol < 0 || k + (size_t)num_pol > L.size()) throw std::runtime_error("Unexpected end of potential file.");
        std::vector<int> site(num_pol);
        std::vector<double> vals((size_t)num_pol * nc);
        const size_t k0 = k;
        parallel_for((size_t)num_pol, [&](size_t p) {
          Tok tk(L[k0 + p]);
          site[p] = to_int(tk.need()) - 1;
          for (int v = 0; v < nc; ++v) vals[p * nc + v] = to_double(tk.need());
        });
        k += (size_t)num_pol;
        if (num_pol > 0) line = L[k - 1];
        for (int p = 0; p < num_pol; ++p) {
          Polarizability pol{std::vector<double>(vals.begin() + (size_t)p * nc, vals.begin() + (size_t)(p + 1) * nc)};
          if (potentials.at(site[p]).is_polarizable())
            throw std::runtime_error("Potential can only have one polarizability.");
          potentials[site[p]].set_polarizability(pol);
        }
      } else {
        throw std::runtime_error("Invalid number in potfile ORDER.");
      }
    }
    if (has(line, "EXCLISTS")) {
      getline(line);
      const int num_excl = to_int(Tok(line).need());
      if (num_excl < 0 || k + (size_t)num_excl > L.size()) throw std::runtime_error("Unexpected end of potential file.");
      std::vector<int> site(num_excl);
      std::vector<std::vector<int>> ex(num_excl);
      const size_t k0 = k;
      parallel_for((size_t)num_excl, [&](size_t q) {
        Tok tk(L[k0 + q]);
        site[q] = to_int(tk.need()) - 1;
        std::string_view s;
        while (tk.next(s)) {
          const int e = to_int(s) - 1;
          if (e >= 0) ex[q].push_back(e);
        }
      });
      k += (size_t)num_excl;
      if (num_excl > 0) line = L[k - 1];
      for (int q = 0; q < num_excl; ++q) {
        Potential& p = potentials.at(site[q]);
        for (int e : ex[q]) p.add_exclusion(e);
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
