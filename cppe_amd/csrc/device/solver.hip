// Jacobi-preconditioned conjugate gradients on the device (InducedMoments::compute,
// src/core/electric_fields.cc:191-239). Vectors are 3n SoA in device order; all reductions
// are two-stage with a fixed grid and a fixed summation tree, so results are deterministic.
//
// FP64 form: the reference recurrence with every B apply in FP64.
// Mixed form: the same recurrence with FP32 field applies, made exact by reliable updates
// (residual replacement; Sleijpen & van der Vorst, Computing 56, 141 (1996)): whenever the
// recursively updated residual has fallen by kReliableDrop since the last replacement, or
// below the threshold, the true residual rhs - B x is recomputed with an FP64 apply and
// replaces it. Convergence is only ever declared on an FP64 true residual, so both forms stop
// on the same FP64 criterion ||rhs - B x||_2 < thresh.
#include <cmath>
#include <vector>

#include "system.hh"

namespace cppe_dev {

void apply_b_direct(System& S, const double* xin, double* out, bool fp32);
void apply_b_fmm(System& S, const double* xin, double* out, bool fp32);

constexpr int kRedBlocks = 512;
constexpr double kReliableDrop = 1e-4;

enum Scal { S_RZ = 0, S_ALPHA, S_BETA, S_RR, S_ZR, S_PAP, S_RNORM, S_COUNT };

void apply_b(System& S, const double* x, double* out, bool fp32) {
  if (S.opts.summation == CPPE_HIP_SUM_FMM)
    apply_b_fmm(S, x, out, fp32);
  else
    apply_b_direct(S, x, out, fp32);
}

// out = alpha in (3x3 packed SoA per site)
__global__ __launch_bounds__(256) void k_alpha_apply(int n, const double* __restrict__ a6,
                                                     const double* __restrict__ in, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m0 = in[i], m1 = in[n + i], m2 = in[2 * n + i];
  const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
               a5 = a6[5 * n + i];
  out[i]         = a0 * m0 + a1 * m1 + a2 * m2;
  out[n + i]     = a1 * m0 + a3 * m1 + a4 * m2;
  out[2 * n + i] = a2 * m0 + a4 * m1 + a5 * m2;
}

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

// r = rhs - t ; z = alpha r ; p = z ; partials (r.z, r.r)
__global__ __launch_bounds__(256) void k_cg_init(int n, const double* __restrict__ a6,
                                                 const double* __restrict__ rhs, const double* __restrict__ t,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ part) {
  __shared__ double red[4];
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double r0 = rhs[i] - t[i], r1 = rhs[n + i] - t[n + i], r2 = rhs[2 * n + i] - t[2 * n + i];
    const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
                 a5 = a6[5 * n + i];
    const double z0 = a0 * r0 + a1 * r1 + a2 * r2;
    const double z1 = a1 * r0 + a3 * r1 + a4 * r2;
    const double z2 = a2 * r0 + a4 * r1 + a5 * r2;
    r[i] = r0;
    r[n + i] = r1;
    r[2 * n + i] = r2;
    z[i] = z0;
    z[n + i] = z1;
    z[2 * n + i] = z2;
    p[i] = z0;
    p[n + i] = z1;
    p[2 * n + i] = z2;
    rz += r0 * z0 + r1 * z1 + r2 * z2;
    rr += r0 * r0 + r1 * r1 + r2 * r2;
  }
  double s1 = block_sum(rz, red);
  double s2 = block_sum(rr, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x]             = s1;
    part[kRedBlocks + blockIdx.x] = s2;
  }
}

// partial p.Ap
__global__ __launch_bounds__(256) void k_dot3(int n3, const double* __restrict__ a, const double* __restrict__ b,
                                              double* __restrict__ part) {
  __shared__ double red[4];
  double v = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += gridDim.x * blockDim.x) v += a[k] * b[k];
  double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x += alpha p ; r -= alpha Ap ; partial r.r   (alpha = scal[S_ALPHA])
__global__ __launch_bounds__(256) void k_xr_update(int n3, const double* __restrict__ scal,
                                                   const double* __restrict__ p, const double* __restrict__ Ap,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ part) {
  __shared__ double red[4];
  const double al = scal[S_ALPHA];
  double v = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += gridDim.x * blockDim.x) {
    x[k] += al * p[k];
    const double rk = r[k] - al * Ap[k];
    r[k] = rk;
    v += rk * rk;
  }
  double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// r = rhs - t ; partial r.r (reliable update / initial residual)
__global__ __launch_bounds__(256) void k_residual(int n3, const double* __restrict__ rhs,
                                                  const double* __restrict__ t, double* __restrict__ r,
                                                  double* __restrict__ part) {
  __shared__ double red[4];
  double v = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += gridDim.x * blockDim.x) {
    const double rk = rhs[k] - t[k];
    r[k] = rk;
    v += rk * rk;
  }
  double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// z = alpha r ; partial z.r
__global__ __launch_bounds__(256) void k_z_update(int n, const double* __restrict__ a6,
                                                  const double* __restrict__ r, double* __restrict__ z,
                                                  double* __restrict__ part) {
  __shared__ double red[4];
  double v = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double r0 = r[i], r1 = r[n + i], r2 = r[2 * n + i];
    const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
                 a5 = a6[5 * n + i];
    const double z0 = a0 * r0 + a1 * r1 + a2 * r2;
    const double z1 = a1 * r0 + a3 * r1 + a4 * r2;
    const double z2 = a2 * r0 + a4 * r1 + a5 * r2;
    z[i] = z0;
    z[n + i] = z1;
    z[2 * n + i] = z2;
    v += r0 * z0 + r1 * z1 + r2 * z2;
  }
  double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// p = z + beta p  (beta = scal[S_BETA])
__global__ __launch_bounds__(256) void k_p_update(int n3, const double* __restrict__ scal,
                                                  const double* __restrict__ z, double* __restrict__ p) {
  const double be = scal[S_BETA];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += gridDim.x * blockDim.x) p[k] = z[k] + be * p[k];
}

// Single-block finalisation of the partial sums (fixed order -> deterministic).
// mode 0: init   (rz, rr)      -> S_RZ, S_RNORM
// mode 1: alpha  (pAp)         -> S_ALPHA = S_RZ / pAp
// mode 2: rnorm  (rr)          -> S_RNORM = sqrt(rr)
// mode 3: beta   (zr)          -> S_BETA = zr / S_RZ ; S_RZ = zr
__global__ __launch_bounds__(256) void k_finalize(int mode, int nblk, const double* __restrict__ part,
                                                  double* __restrict__ scal) {
  __shared__ double red[4];
  double v1 = 0.0, v2 = 0.0;
  for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
    v1 += part[k];
    if (mode == 0) v2 += part[kRedBlocks + k];
  }
  double s1 = block_sum(v1, red);
  double s2 = block_sum(v2, red);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      scal[S_RZ]    = s1;
      scal[S_RNORM] = sqrt(s2);
    } else if (mode == 1) {
      scal[S_PAP]   = s1;
      scal[S_ALPHA] = scal[S_RZ] / s1;
    } else if (mode == 2) {
      scal[S_RR]    = s1;
      scal[S_RNORM] = sqrt(s1);
    } else {
      scal[S_ZR]   = s1;
      scal[S_BETA] = s1 / scal[S_RZ];
      scal[S_RZ]   = s1;
    }
  }
}

static int red_blocks(long long m) { return (int)std::min<long long>(kRedBlocks, (m + 255) / 256); }

static double read_scalar(System& S, int which) {
  double v = 0.0;
  CPPE_HIP_CHECK(hipMemcpyAsync(&v, S.scal.p + which, sizeof(double), hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  return v;
}

void precondition(System& S, const double* in, double* out) {
  hipLaunchKernelGGL(k_alpha_apply, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, S.alpha6.p, in, out);
}

// Solve B x = rhs for device vectors (x holds the initial guess on entry).
// Returns the number of iterations and fills norms (host vector) per iteration.
int device_cg(System& S, const double* rhs, double* x, double thresh, int maxiter, bool mixed,
              std::vector<double>& norms, bool& converged) {
  const int n = S.n, n3 = 3 * n;
  hipStream_t st = S.stream;
  const int gb3 = red_blocks(n3), gb1 = red_blocks(n);
  double* r = S.v_r.p;
  double* z = S.v_z.p;
  double* p = S.v_p.p;
  double* Ap = S.v_Ap.p;
  double* t = S.v_t.p;
  double* part = S.red.p;
  double* scal = S.scal.p;
  norms.clear();
  converged = false;
  // r0 = rhs - B x0 (FP64), z0 = alpha r0, p = z0
  apply_b(S, x, t, false);
  hipLaunchKernelGGL(k_cg_init, dim3(gb1), dim3(256), 0, st, n, S.alpha6.p, rhs, t, r, z, p, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, 0, gb1, part, scal);
  double r_last = read_scalar(S, S_RNORM);
  for (int k = 0; k < maxiter; ++k) {
    apply_b(S, p, Ap, mixed);
    hipLaunchKernelGGL(k_dot3, dim3(gb3), dim3(256), 0, st, n3, p, Ap, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, 1, gb3, part, scal);
    hipLaunchKernelGGL(k_xr_update, dim3(gb3), dim3(256), 0, st, n3, scal, p, Ap, x, r, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, 2, gb3, part, scal);
    double rn = read_scalar(S, S_RNORM);
    if (mixed && (rn < thresh || rn < kReliableDrop * r_last)) {
      // reliable update: replace the recursive residual by the FP64 true residual
      apply_b(S, x, t, false);
      hipLaunchKernelGGL(k_residual, dim3(gb3), dim3(256), 0, st, n3, rhs, t, r, part);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, 2, gb3, part, scal);
      rn     = read_scalar(S, S_RNORM);
      r_last = rn;
    }
    norms.push_back(rn);
    if (!std::isfinite(rn)) break;
    if (rn < thresh) {
      converged = true;
      return k + 1;
    }
    hipLaunchKernelGGL(k_z_update, dim3(gb1), dim3(256), 0, st, n, S.alpha6.p, r, z, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, 3, gb1, part, scal);
    hipLaunchKernelGGL(k_p_update, dim3(gb3), dim3(256), 0, st, n3, scal, z, p);
  }
  return (int)norms.size();
}

}  // namespace cppe_dev
