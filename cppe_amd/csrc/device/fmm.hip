// Fast multipole method on gfx950 (placeholder until the operator kernels land).
#include "system.hh"

namespace cppe_dev {

struct FmmState {};

void apply_b_fmm(System&, const double*, double*, bool) {
  throw std::runtime_error("fmm summation not available in this build");
}
void fmm_field(System&, int, const double*, double*, double, bool) {
  throw std::runtime_error("fmm summation not available in this build");
}
void fmm_static(System&, const double*, const double*, const double*, double*, double) {
  throw std::runtime_error("fmm summation not available in this build");
}
void fmm_destroy(System& S) {
  delete S.fmm;
  S.fmm = nullptr;
}

}  // namespace cppe_dev
