// Dense inverse of B on the device (BMatrix::direct_inverse, src/core/bmatrix.cc:94-97):
// Gauss-Jordan elimination with partial pivoting on [B | I], one pivot column per step, all
// steps enqueued on the stream without host round trips. Test-size matrices (3n <= ~10^4).
#include "system.hh"

namespace cppe_dev {

// pivot row = argmax_{r >= k} |A[r][k]| (first maximum), single block
__global__ __launch_bounds__(256) void k_gj_pivot(int N, int k, const double* __restrict__ A, int* piv) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  double best = -1.0;
  int bidx    = k;
  for (int r = k + threadIdx.x; r < N; r += blockDim.x) {
    double v = fabs(A[(size_t)r * N + k]);
    if (v > best) {
      best = v;
      bidx = r;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = bidx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = bv[threadIdx.x + s];
      int oi   = bi[threadIdx.x + s];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *piv = bi[0];
}

// swap rows k and piv of A and X, then scale row k by 1/A[k][k]; save column k factors
__global__ __launch_bounds__(256) void k_gj_swap_scale(int N, int k, double* A, double* X, const int* piv,
                                                       double* fac) {
  const int p = *piv;
  // one block does the swap + scale of row k (columns strided)
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    if (p != k) {
      double a = A[(size_t)k * N + c], b = A[(size_t)p * N + c];
      A[(size_t)k * N + c] = b;
      A[(size_t)p * N + c] = a;
      double x = X[(size_t)k * N + c], y = X[(size_t)p * N + c];
      X[(size_t)k * N + c] = y;
      X[(size_t)p * N + c] = x;
    }
  }
  __syncthreads();
  const double d = A[(size_t)k * N + k];
  __syncthreads();
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    A[(size_t)k * N + c] /= d;
    X[(size_t)k * N + c] /= d;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < N; r += blockDim.x) fac[r] = (r == k) ? 0.0 : A[(size_t)r * N + k];
}

__global__ __launch_bounds__(256) void k_gj_eliminate(int N, int k, double* A, double* X,
                                                      const double* __restrict__ fac) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * N) return;
  const int r = (int)(t / N), c = (int)(t % N);
  const double f = fac[r];
  if (f == 0.0) return;
  A[(size_t)r * N + c] -= f * A[(size_t)k * N + c];
  X[(size_t)r * N + c] -= f * X[(size_t)k * N + c];
}

__global__ void k_identity(int N, double* X) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (long long)N * N) X[t] = ((t / N) == (t % N)) ? 1.0 : 0.0;
}

// A (device, N x N row-major) is destroyed; X receives A^-1.
void dense_inverse(hipStream_t st, int N, double* A, double* X) {
  dbuf<int> piv;
  dbuf<double> fac;
  piv.alloc(1);
  fac.alloc(N);
  const long long NN = (long long)N * N;
  hipLaunchKernelGGL(k_identity, dim3(ceil_div(NN, 256)), dim3(256), 0, st, N, X);
  for (int k = 0; k < N; ++k) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, st, N, k, A, piv.p);
    hipLaunchKernelGGL(k_gj_swap_scale, dim3(1), dim3(256), 0, st, N, k, A, X, piv.p, fac.p);
    hipLaunchKernelGGL(k_gj_eliminate, dim3(ceil_div(NN, 256)), dim3(256), 0, st, N, k, A, X, fac.p);
  }
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace cppe_dev
