// Allocation-free variant of cppe_hip_nuclear_multipole_energy, for callers that cannot allocate
// device memory (code under HIP graph capture, tight loops that re-evaluate the energy as the
// nuclei move). No counterpart in the reference.
//
// The energy kernels need one partial sum per 256 sites. Here those partials live in a file-scope
// device array of fixed size (kScratchPartials doubles, 128 KiB per device, held until the process
// ends), so the entry point itself calls no hipMalloc, hipFree or allocating library routine.
// Its device address is looked up once per device (hipGetSymbolAddress, with that device
// current) and kept; host code never passes the array's name to a kernel. Every call on a device
// uses that device's array: calls that overlap on one device (two streams, or two threads) race
// on it. Each call writes every partial it later reads, so nothing has to be cleared between calls.
#include <mutex>

#include "system.hh"

namespace cppe_dev {

__global__ __launch_bounds__(256) void k_nuc_mul_energy(int n, const double* __restrict__ pos,
                                                        const int* __restrict__ kmax,
                                                        const double* __restrict__ mul, int na,
                                                        const double* __restrict__ ch,
                                                        const double* __restrict__ apos,
                                                        double* __restrict__ part);
__global__ void k_sum_partials(int np, const double* __restrict__ part, double* __restrict__ out);
int capi_fail(int code, const std::string& msg);

constexpr int kScratchPartials = 16384;  // one partial per 256 sites: at most 4,194,304 sites
static __device__ __align__(256) double scratch[kScratchPartials];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first call there.
static double* scratch_here() {
  int dev = 0;
  CPPE_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw hip_error("device index beyond the allocation-free table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    CPPE_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace cppe_dev

using namespace cppe_dev;

extern "C" int cppe_hip_nuclear_multipole_energy_noalloc(int n, const double* pos, const int* kmax,
                                                         const double* mul, int natoms, const double* charges,
                                                         const double* atom_pos, double* energy, void* stream) {
  try {
    if (n < 0 || natoms < 0 || !energy) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
    if (n > 256LL * kScratchPartials)
      return capi_fail(CPPE_HIP_ERR_ARG, "allocation-free nuclear-multipole energy: " + std::to_string(n) +
                                             " sites exceed its maximum of " + std::to_string(256 * kScratchPartials));
    if (n > 0 && natoms > 0 && (!pos || !kmax || !mul || !charges || !atom_pos))
      return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* part   = scratch_here();
    if (n == 0 || natoms == 0) {
      // zero partials summed: writes 0 to *energy (a kernel, not a memset, so the call can be captured)
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, 0, part, energy);
    } else {
      const int nb = ceil_div(n, 256);
      hipLaunchKernelGGL(k_nuc_mul_energy, dim3(nb), dim3(256), 0, st, n, pos, kmax, mul, natoms, charges,
                         atom_pos, part);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, nb, part, energy);
    }
    CPPE_HIP_CHECK(hipGetLastError());
    return CPPE_HIP_OK;
  } catch (const std::exception& e) {
    return capi_fail(CPPE_HIP_ERR_DEVICE, e.what());
  }
}
