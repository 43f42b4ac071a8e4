// One expansion order of the FMM kernels; the build compiles this file once per order with
// -DCPPE_FMM_ORDER=P (2..8) so that the orders compile in parallel.
#ifndef CPPE_FMM_ORDER
#define CPPE_FMM_ORDER 5
#endif
#define CPPE_FMMOPS_ONLY_P CPPE_FMM_ORDER
#include "fmm_impl.hh"

namespace cppe_dev {
CPPE_INSTANTIATE_FMM(CPPE_FMM_ORDER)
}  // namespace cppe_dev
