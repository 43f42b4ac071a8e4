// RAII handle onto the device library (include/cppe_hip.h) built from host Potential objects.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cppe_hip.h"
#include "pe_options.hh"
#include "potential.hh"

namespace libcppe {

// Throws std::runtime_error with the library's message when rc != CPPE_HIP_OK.
void hip_check(int rc);

class DeviceSystem {
 public:
  // sites: the particles of this system (their order = the system order); damping: Thole
  // factor of this system's fields (0 = undamped); with_alpha: upload polarizabilities.
  DeviceSystem(const std::vector<Potential>& sites, const PeOptions& opt, double damping, bool with_alpha);
  ~DeviceSystem();
  DeviceSystem(const DeviceSystem&) = delete;
  DeviceSystem& operator=(const DeviceSystem&) = delete;
  cppe_hip_system* get() const { return m_sys; }
  int n() const { return m_n; }

 private:
  cppe_hip_system* m_sys = nullptr;
  int m_n = 0;
};

cppe_hip_options make_options(const PeOptions& opt, double damping);

}  // namespace libcppe
