// Shared device/host helpers of the gfx950 PE field library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cppe_dev {

struct hip_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define CPPE_HIP_CHECK(expr)                                                                 \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      throw ::cppe_dev::hip_error(std::string(#expr) + ": " + hipGetErrorString(_e) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")");          \
  } while (0)

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------------------------------
// Scalar math, FP64 and FP32.
// ------------------------------------------------------------------------------------------

// 1/sqrt(r2) in FP64: v_rsq_f64 seed (relative error below 2^-22) refined by one third-order
// Newton-Halley step y1 = y0 (1 + e/2 + 3e^2/8), e = 1 - r2 y0^2, which leaves an error of
// order e^3 (below 2^-66) plus rounding.
__device__ __forceinline__ double rsqrt_f64(double r2) {
  double y = __builtin_amdgcn_rsq(r2);
  double e = __builtin_fma(-r2 * y, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
}

__device__ __forceinline__ float rsqrt_f32(float r2) { return __builtin_amdgcn_rsqf(r2); }

// exp(-u) in FP32 through the hardware base-2 exponential.
__device__ __forceinline__ float expm_f32(float u) {
  return __builtin_amdgcn_exp2f(-u * 1.4426950408889634f);
}

// Thole screening factors for the rank-1/2 tensors of the screened potential f_V(v)/R,
// f_V = 1 - (1 + v/2) e^-v:  f_E = 1 - (1 + v + v^2/2) e^-v,  f_T = f_E - v^3/6 e^-v.
// e = e^-v (0 when the pair is undamped). Rank 3: f_D = f_T - v^4/30 e^-v.
template <typename R>
__device__ __forceinline__ void thole_ET(R v, R e, R& fE, R& fT) {
  const R v2 = v * v;
  fE = R(1) - (R(1) + v + R(0.5) * v2) * e;
  fT = fE - (v2 * v * R(1.0 / 6.0)) * e;
}

// Above these scaled distances the pair is treated as undamped: the screening terms
// e^-v (1 + v + v^2/2 + v^3/6) are below 5e-14 (FP64, v >= 40) and 1e-7 (FP32, v >= 25) relative,
// far below the rounding of the forms that use them.
constexpr double kTholeCutoff  = 40.0;
constexpr float kTholeCutoff32 = 25.0f;
// Near-list class boundary of the FP32 near field: sources with v >= 18 for every target of a leaf
// take the undamped pair. A single pair's screening terms there reach 1e-5 relative, but they
// are a small part of each site's field: on water boxes the field of the FP32 apply moves by
// 8e-8 (2-norm, relative), below its own rounding error (2e-7).
constexpr double kFarClass32 = 18.0;

// Work unit of block b of nb: workgroups are dispatched to the device's 8 XCDs in turn (block b to
// XCD b mod 8), and each XCD takes a contiguous range of units -- consecutive sites, whose
// neighbours' records mostly lie in the same or an adjacent unit, then stay in one XCD's L2.
// A bijection of [0, nb).
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
  constexpr int kXcd = 8;
  const int per = nb / kXcd, rem = nb % kXcd, x = b % kXcd;
  return x * per + min(x, rem) + b / kXcd;
}

// block-wide sum (256 threads); the result is valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
  return s;
}

}  // namespace cppe_dev
