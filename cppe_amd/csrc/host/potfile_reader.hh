// .pot reader (src/core/potfile_reader.{hh,cc}); host-side input path, not accelerated.
#pragma once
#include <string>
#include <vector>

#include "potential.hh"

namespace libcppe {

class PotfileReader {
 public:
  explicit PotfileReader(std::string potfile_name);
  std::vector<Potential> read();

 private:
  std::string m_potfile;
};

}  // namespace libcppe
