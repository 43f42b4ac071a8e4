// C ABI of the learned CG warm start (declarations: include/cppe_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "ml.hh"

namespace cppe_dev {
void precondition(System& S, const double* in, double* out);
int capi_fail(int code, const std::string& msg);

// Samples from device-order features/bases/moments: V (9 floats), T (3 floats), weight 1/|v1|^2,
// so the loss sum_s w_s |sum_a y_sa V_sa - T_s|^2 is the squared dipole error relative to alpha F.
// With the alpha^-1 objective (ainv6 given) the bases and the target are mapped by L^T, where
// L L^T = alpha^-1 is the site's Cholesky factor: |L^T e|^2 = e^T alpha^-1 e, the site's diagonal
// block of the B-norm e^T B e (B = alpha^-1 - T) that CG minimises, relative to |v1| in that norm.
__global__ void k_ml_samples(int n, const double* __restrict__ Vsoa, const double* __restrict__ mu,
                             const double* __restrict__ ainv6, float* __restrict__ V, float* __restrict__ T,
                             float* __restrict__ W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[3][3], t[3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) v[a][c] = Vsoa[(size_t)(3 * a + c) * n + i];
  for (int c = 0; c < 3; ++c) t[c] = mu ? mu[(size_t)c * n + i] : 0.0;
  if (ainv6) {
    const double a11 = ainv6[i], a12 = ainv6[n + i], a13 = ainv6[2 * (size_t)n + i];
    const double a22 = ainv6[3 * (size_t)n + i], a23 = ainv6[4 * (size_t)n + i], a33 = ainv6[5 * (size_t)n + i];
    const double l11 = sqrt(a11), l21 = a12 / l11, l31 = a13 / l11;
    const double l22 = sqrt(a22 - l21 * l21), l32 = (a23 - l31 * l21) / l22;
    const double l33 = sqrt(a33 - l31 * l31 - l32 * l32);
    auto lt = [&](double* u) {  // u <- L^T u
      const double u0 = l11 * u[0] + l21 * u[1] + l31 * u[2];
      const double u1 = l22 * u[1] + l32 * u[2];
      const double u2 = l33 * u[2];
      u[0] = u0;
      u[1] = u1;
      u[2] = u2;
    };
    for (int a = 0; a < 3; ++a) lt(v[a]);
    lt(t);
  }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) V[(size_t)i * 9 + 3 * a + c] = (float)v[a][c];
  for (int c = 0; c < 3; ++c) T[(size_t)i * 3 + c] = (float)t[c];
  W[i] = (float)(1.0 / (v[0][0] * v[0][0] + v[0][1] * v[0][1] + v[0][2] * v[0][2] + 1e-30));
}

// guess mu = sum_a y_a v_a into 3n SoA (device order)
__global__ void k_ml_combine(int n, const float* __restrict__ y, const double* __restrict__ V,
                             double* __restrict__ mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c0 = y[3 * (size_t)i], c1 = y[3 * (size_t)i + 1], c2 = y[3 * (size_t)i + 2];
  for (int c = 0; c < 3; ++c) {
    const double g = c0 * V[(size_t)c * n + i] + c1 * V[(size_t)(3 + c) * n + i] + c2 * V[(size_t)(6 + c) * n + i];
    mu[(size_t)c * n + i] = __builtin_isfinite(g) ? g : 0.0;
  }
}

__global__ void k_feature_moments(long long ns, const float* __restrict__ X, double* __restrict__ acc) {
  // acc[2*f], acc[2*f+1] += sum x, sum x^2 of feature f = blockIdx.y
  __shared__ double r1[256], r2[256];
  const int f = blockIdx.y;
  double s1 = 0.0, s2 = 0.0;
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (long long)gridDim.x * blockDim.x) {
    const double v = X[s * kMlpIn + f];
    s1 += v;
    s2 += v * v;
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      r1[threadIdx.x] += r1[threadIdx.x + k];
      r2[threadIdx.x] += r2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[((size_t)blockIdx.x * kMlpIn + f) * 2] = r1[0];
    acc[((size_t)blockIdx.x * kMlpIn + f) * 2 + 1] = r2[0];
  }
}

__global__ void k_soa_rows(int n, int nc, const int* __restrict__ perm, const float* __restrict__ rows_dev,
                           float* __restrict__ rows_caller) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  for (int c = 0; c < nc; ++c) rows_caller[(size_t)j * nc + c] = rows_dev[(size_t)i * nc + c];
}

}  // namespace cppe_dev

using namespace cppe_dev;

struct cppe_hip_mlp {
  Mlp M;
  dbuf<float> rawX;  // collected raw features (before prepare)
  bool prepared = false;
  int objective = 0;  // 0: dipole error, 1: alpha^-1-norm error (set before collecting)
};
struct cppe_hip_system {
  System S;
};

#define ML_BEGIN try {
#define ML_END                                                  \
  }                                                             \
  catch (const std::exception& e) {                             \
    return cppe_dev::capi_fail(CPPE_HIP_ERR_DEVICE, e.what());  \
  }

static void mlp_init(Mlp& M) {
  M.w.alloc(kMlpParams);
  M.m.alloc(kMlpParams);
  M.v.alloc(kMlpParams);
  M.mean.alloc(kMlpIn);
  M.inv_std.alloc(kMlpIn);
  M.m.zero(M.stream);
  M.v.zero(M.stream);
  // default: Glorot-uniform hidden layers (fixed seed), output layer zero with b3 = (1, 0, 0) so
  // the untrained guess is alpha F (the reference's first guess)
  std::vector<float> w(kMlpParams, 0.f);
  std::mt19937 rng(12345);
  auto fill = [&](int off, int fin, int fout) {
    const float lim = std::sqrt(6.f / (fin + fout));
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int k = 0; k < fin * fout; ++k) w[off + k] = U(rng);
  };
  fill(kOffW1, kMlpIn, kMlpHid);
  fill(kOffW2, kMlpHid, kMlpHid);
  w[kOffB3] = 1.f;
  std::vector<float> mean(kMlpIn, 0.f), istd(kMlpIn, 1.f);
  CPPE_HIP_CHECK(hipMemcpyAsync(M.w.p, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(M.mean.p, mean.data(), sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(M.inv_std.p, istd.data(), sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  mlp_refresh_bf16(M);
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
}

extern "C" {

int cppe_hip_mlp_create(cppe_hip_mlp** out) {
  ML_BEGIN
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return capi_fail(CPPE_HIP_ERR_DEVICE, "no HIP device available");
  auto* m = new cppe_hip_mlp();
  CPPE_HIP_CHECK(hipStreamCreateWithFlags(&m->M.stream, hipStreamNonBlocking));
  mlp_init(m->M);
  *out = m;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_destroy(cppe_hip_mlp* m) {
  ML_BEGIN
  if (m) {
    hipStream_t s = m->M.stream;
    delete m;
    if (s) (void)hipStreamDestroy(s);
  }
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_set(cppe_hip_mlp* m, const float* params, const float* mean, const float* inv_std) {
  ML_BEGIN
  if (!m || !params) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  CPPE_HIP_CHECK(hipMemcpyAsync(M.w.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, M.stream));
  if (mean) CPPE_HIP_CHECK(hipMemcpyAsync(M.mean.p, mean, sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  if (inv_std)
    CPPE_HIP_CHECK(hipMemcpyAsync(M.inv_std.p, inv_std, sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  M.m.zero(M.stream);
  M.v.zero(M.stream);
  M.step = 0;
  mlp_refresh_bf16(M);
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_get(cppe_hip_mlp* m, float* params, float* mean, float* inv_std) {
  ML_BEGIN
  if (!m) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  if (params) CPPE_HIP_CHECK(hipMemcpyAsync(params, M.w.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, M.stream));
  if (mean) CPPE_HIP_CHECK(hipMemcpyAsync(mean, M.mean.p, sizeof(float) * kMlpIn, hipMemcpyDeviceToHost, M.stream));
  if (inv_std)
    CPPE_HIP_CHECK(hipMemcpyAsync(inv_std, M.inv_std.p, sizeof(float) * kMlpIn, hipMemcpyDeviceToHost, M.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_ml_features(cppe_hip_system* sys, float* X, double* V) {
  ML_BEGIN
  if (sys && sys->S.nseg > 1)
    return capi_fail(CPPE_HIP_ERR_UNSUPPORTED, "descriptors of a snapshot batch are not supported");
  if (!sys || !X || !V) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S = sys->S;
  if (!S.has_alpha) return capi_fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  const int n = S.n;
  dbuf<float> dX, dXc;
  dbuf<double> dV;
  dX.alloc((size_t)n * kMlpIn);
  dXc.alloc((size_t)n * kMlpIn);
  dV.alloc(9 * (size_t)n);
  ml_features(S, dX.p, dV.p);
  hipLaunchKernelGGL(k_soa_rows, dim3(ceil_div(n, 256)), dim3(256), 0, S.stream, n, kMlpIn, S.perm.p, dX.p, dXc.p);
  // bases: SoA device order -> caller order rows of 9
  std::vector<double> hv(9 * (size_t)n);
  CPPE_HIP_CHECK(hipMemcpyAsync(hv.data(), dV.p, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(X, dXc.p, sizeof(float) * (size_t)n * kMlpIn, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  for (int i = 0; i < n; ++i) {
    const int j = S.h_perm[i];
    for (int k = 0; k < 9; ++k) V[9 * (size_t)j + k] = hv[(size_t)k * n + i];
  }
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_forward(cppe_hip_mlp* m, const float* X, long long ns, float* y) {
  ML_BEGIN
  if (!m || !X || !y || ns <= 0) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  dbuf<float> dX, dy;
  dX.alloc((size_t)ns * kMlpIn);
  dy.alloc((size_t)ns * 3);
  CPPE_HIP_CHECK(hipMemcpyAsync(dX.p, X, sizeof(float) * ns * kMlpIn, hipMemcpyHostToDevice, M.stream));
  mlp_forward(M, dX.p, (int)ns, dy.p, M.stream);
  CPPE_HIP_CHECK(hipMemcpyAsync(y, dy.p, sizeof(float) * ns * 3, hipMemcpyDeviceToHost, M.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_loss_grad(cppe_hip_mlp* m, const float* X, const float* V, const float* T, const float* w,
                           long long ns, float* grad, double* loss) {
  ML_BEGIN
  if (!m || !X || !V || !T || !w || !grad || ns <= 0) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  dbuf<float> dX, dV, dT, dW;
  dbuf<unsigned short> dXb;
  dX.alloc((size_t)ns * kMlpIn);
  dXb.alloc((size_t)ns * kMlpIn);
  dV.alloc((size_t)ns * 9);
  dT.alloc((size_t)ns * 3);
  dW.alloc((size_t)ns);
  CPPE_HIP_CHECK(hipMemcpyAsync(dX.p, X, sizeof(float) * ns * kMlpIn, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(dV.p, V, sizeof(float) * ns * 9, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(dT.p, T, sizeof(float) * ns * 3, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(dW.p, w, sizeof(float) * ns, hipMemcpyHostToDevice, M.stream));
  mlp_normalise(M, dX.p, ns, dXb.p, M.stream);
  double l = 0.0;
  mlp_loss_grad(M, dXb.p, dV.p, dT.p, dW.p, ns, nullptr, -1, 0.0f, &l, M.stream);
  CPPE_HIP_CHECK(hipMemcpyAsync(grad, M.grad.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, M.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  if (loss) *loss = l;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_reserve(cppe_hip_mlp* m, long long ns) {
  ML_BEGIN
  if (!m || ns <= 0) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  if (M.ns > 0) return capi_fail(CPPE_HIP_ERR_ARG, "reserve before collecting");
  m->rawX.alloc((size_t)ns * kMlpIn);
  M.dV.alloc((size_t)ns * 9);
  M.dT.alloc((size_t)ns * 3);
  M.dS.alloc((size_t)ns);
  M.cap = ns;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_collect(cppe_hip_mlp* m, cppe_hip_system* sys) {
  ML_BEGIN
  if (sys && sys->S.nseg > 1)
    return capi_fail(CPPE_HIP_ERR_UNSUPPORTED, "descriptors of a snapshot batch are not supported");
  if (!m || !sys) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M    = m->M;
  System& S = sys->S;
  if (!S.mu_valid) return capi_fail(CPPE_HIP_ERR_ARG, "system has no solved moments to learn from");
  if (M.ns + S.n > M.cap) return capi_fail(CPPE_HIP_ERR_ARG, "training set capacity exceeded");
  dbuf<double> dV;
  dV.alloc(9 * (size_t)S.n);
  ml_features(S, m->rawX.p + (size_t)M.ns * kMlpIn, dV.p);
  hipLaunchKernelGGL(k_ml_samples, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, dV.p, S.v_mu.p,
                     m->objective == 1 ? S.ainv6.p : nullptr, M.dV.p + (size_t)M.ns * 9, M.dT.p + (size_t)M.ns * 3,
                     M.dS.p + M.ns);
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  M.ns += S.n;
  m->prepared = false;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_set_objective(cppe_hip_mlp* m, int objective) {
  ML_BEGIN
  if (!m || objective < 0 || objective > 1) return capi_fail(CPPE_HIP_ERR_ARG, "objective must be 0 or 1");
  if (m->M.ns > 0) return capi_fail(CPPE_HIP_ERR_ARG, "set the objective before collecting");
  m->objective = objective;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_samples(cppe_hip_mlp* m, long long s0, long long count, float* V, float* T, float* w) {
  ML_BEGIN
  if (!m || s0 < 0 || count < 0 || s0 + count > m->M.ns || !V || !T || !w)
    return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  CPPE_HIP_CHECK(hipMemcpy(V, M.dV.p + s0 * 9, sizeof(float) * 9 * count, hipMemcpyDeviceToHost));
  CPPE_HIP_CHECK(hipMemcpy(T, M.dT.p + s0 * 3, sizeof(float) * 3 * count, hipMemcpyDeviceToHost));
  CPPE_HIP_CHECK(hipMemcpy(w, M.dS.p + s0, sizeof(float) * count, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_prepare(cppe_hip_mlp* m) {
  ML_BEGIN
  if (!m) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  Mlp& M = m->M;
  if (M.ns <= 0) return capi_fail(CPPE_HIP_ERR_ARG, "empty training set");
  const int nb = 256;
  dbuf<double> acc;
  acc.alloc((size_t)nb * kMlpIn * 2);
  hipLaunchKernelGGL(k_feature_moments, dim3(nb, kMlpIn), dim3(256), 0, M.stream, M.ns, m->rawX.p, acc.p);
  std::vector<double> h((size_t)nb * kMlpIn * 2);
  CPPE_HIP_CHECK(hipMemcpyAsync(h.data(), acc.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, M.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  std::vector<float> mean(kMlpIn), istd(kMlpIn);
  for (int f = 0; f < kMlpIn; ++f) {
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < nb; ++b) {
      s1 += h[((size_t)b * kMlpIn + f) * 2];
      s2 += h[((size_t)b * kMlpIn + f) * 2 + 1];
    }
    const double mu = s1 / M.ns, var = std::max(0.0, s2 / M.ns - mu * mu);
    mean[f] = (float)mu;
    istd[f] = (float)(1.0 / std::sqrt(var + 1e-12 + 1e-6 * mu * mu));
    if (var < 1e-20) istd[f] = 0.f;  // constant feature: carried by the biases
  }
  CPPE_HIP_CHECK(hipMemcpyAsync(M.mean.p, mean.data(), sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(M.inv_std.p, istd.data(), sizeof(float) * kMlpIn, hipMemcpyHostToDevice, M.stream));
  M.dX.alloc((size_t)M.ns * kMlpIn);
  mlp_normalise(M, m->rawX.p, M.ns, M.dX.p, M.stream);
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  m->prepared = true;
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_mlp_train(cppe_hip_mlp* m, int epochs, int batch, float lr, unsigned seed, double* losses) {
  ML_BEGIN
  if (!m || epochs < 0 || batch < 32) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  if (!m->prepared) return capi_fail(CPPE_HIP_ERR_ARG, "call cppe_hip_mlp_prepare first");
  Mlp& M           = m->M;
  const int ntiles = (int)((M.ns + 31) / 32);
  const int tpb    = std::max(1, batch / 32);
  std::vector<int> order(ntiles);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937 rng(seed);
  dbuf<int> dorder;
  dorder.alloc(ntiles);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    CPPE_HIP_CHECK(hipMemcpyAsync(dorder.p, order.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, M.stream));
    double esum = 0.0;
    int nb      = 0;
    for (int t0 = 0; t0 < ntiles; t0 += tpb) {
      const int nt = std::min(tpb, ntiles - t0);
      double l     = 0.0;
      mlp_loss_grad(M, M.dX.p, M.dV.p, M.dT.p, M.dS.p, M.ns, dorder.p + t0, nt, 1.0f / (32.0f * nt), &l, M.stream);
      mlp_adam(M, lr, M.stream);
      esum += l;
      ++nb;
    }
    if (losses) losses[ep] = esum / std::max(nb, 1);
  }
  CPPE_HIP_CHECK(hipStreamSynchronize(M.stream));
  return CPPE_HIP_OK;
  ML_END
}

int cppe_hip_ml_guess(cppe_hip_system* sys, cppe_hip_mlp* m, double* device_ms) {
  ML_BEGIN
  if (sys && sys->S.nseg > 1)
    return capi_fail(CPPE_HIP_ERR_UNSUPPORTED, "descriptors of a snapshot batch are not supported");
  if (!sys || !m) return capi_fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S = sys->S;
  if (!S.has_alpha) return capi_fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  hipStream_t st = S.stream;
  const int n = S.n;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (device_ms) {
    CPPE_HIP_CHECK(hipEventCreate(&e0));
    CPPE_HIP_CHECK(hipEventCreate(&e1));
  }
  MlGrid& G = ml_grid(S);  // neighbour grid and work buffers, kept by the system
  G.X.alloc((size_t)n * kMlpIn);
  G.y.alloc((size_t)n * 3);
  G.V.alloc(9 * (size_t)n);
  if (device_ms) CPPE_HIP_CHECK(hipEventRecord(e0, st));
  ml_features(S, G.X.p, G.V.p);
  CPPE_HIP_CHECK(hipStreamSynchronize(m->M.stream));
  mlp_forward(m->M, G.X.p, n, G.y.p, st);
  hipLaunchKernelGGL(k_ml_combine, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, G.y.p, G.V.p, S.v_mu.p);
  if (device_ms) {  // timed: wait for the guess; otherwise it stays queued before what follows on the stream
    CPPE_HIP_CHECK(hipEventRecord(e1, st));
    CPPE_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CPPE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *device_ms = ms;
  }
  S.mu_valid = true;
  return CPPE_HIP_OK;
  ML_END
}

}  // extern "C"
