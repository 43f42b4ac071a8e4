// C ABI of the gfx950 polarizable-embedding library (declarations: include/cppe_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "system.hh"

namespace cppe_dev {

void dense_inverse(hipStream_t st, int N, double* A, double* X);
void fmm_autotune(System& S, int reps, int* choice, double* ms);
void field1_targets(System& S, const double* mu_soa, int nt, const int* tdev, double* out);
void fmm_set_tune(System& S, const int* choice);
bool near_variant_allowed(System& S, int variant);
bool decode_overlap(int code, int& form, int& splits);
static bool overlap_code_valid(int code) {
  int form = 0, splits = 1;
  return decode_overlap(code, form, splits);
}

__global__ void k_scatter_xyz(int n, const double* __restrict__ aos, double* x, double* y, double* z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = aos[3 * i];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}

// per-site setup in caller order -> device order: alpha6, alpha^-1 (cofactors / determinant,
// as Eigen's 3x3 inverse does), s = alpha_iso^(-1/6) (0 for undamped sites)
__global__ void k_site_setup(int n, const int* __restrict__ perm, const double* __restrict__ a6_aos,
                             const double* __restrict__ iso, double* a6, double* ainv6, double* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  if (a6_aos) {
    const double* a = &a6_aos[6 * j];
    const double xx = a[0], xy = a[1], xz = a[2], yy = a[3], yz = a[4], zz = a[5];
    const double c00 = yy * zz - yz * yz, c01 = yz * xz - xy * zz, c02 = xy * yz - yy * xz;
    const double det = xx * c00 + xy * c01 + xz * c02;
    const double id  = 1.0 / det;
    const double i00 = c00 * id, i01 = c01 * id, i02 = c02 * id;
    const double i11 = (xx * zz - xz * xz) * id, i12 = (xz * xy - xx * yz) * id, i22 = (xx * yy - xy * xy) * id;
    const double v[6] = {xx, xy, xz, yy, yz, zz};
    const double w[6] = {i00, i01, i02, i11, i12, i22};
    for (int k = 0; k < 6; ++k) {
      a6[(size_t)k * n + i]    = v[k];
      ainv6[(size_t)k * n + i] = w[k];
    }
  }
  double sv = 0.0;
  if (iso) {
    const double a = iso[j];
    if (a > 0.0) sv = pow(a, -1.0 / 6.0);
  }
  s[i] = sv;
}

__global__ void k_excl_count(int n, const int* __restrict__ perm, const int* __restrict__ off, int* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  cnt[i]      = off[j + 1] - off[j];
}

__global__ void k_excl_fill(int n, const int* __restrict__ perm, const int* __restrict__ iperm,
                            const int* __restrict__ off, const int* __restrict__ idx,
                            const int* __restrict__ noff, int* nidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  int o = noff[i];
  for (int k = off[j]; k < off[j + 1]; ++k) {
    const int e = idx[k];
    nidx[o++]   = (e >= 0 && e < n) ? iperm[e] : -1;
  }
}

System::~System() {
  cg_graphs_release(*this);
  if (fmm) fmm_destroy(*this);
  if (ml_grid) ml_grid_destroy(*this);
  if (pc) precond_destroy(*this);
  if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace cppe_dev

using namespace cppe_dev;

struct cppe_hip_system {
  System S;
};

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace cppe_dev {
int capi_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace cppe_dev

#define CPPE_API_BEGIN try {
#define CPPE_API_END                                               \
  }                                                                \
  catch (const cppe_dev::hip_error& e) {                           \
    return fail(CPPE_HIP_ERR_DEVICE, e.what());                    \
  }                                                                \
  catch (const std::exception& e) {                                \
    return fail(CPPE_HIP_ERR_TREE, e.what());                      \
  }                                                                \
  catch (...) {                                                    \
    return fail(CPPE_HIP_ERR_DEVICE, "unknown error");             \
  }

template <typename T>
static void upload(dbuf<T>& b, const T* h, size_t count, hipStream_t s) {
  b.alloc(count);
  CPPE_HIP_CHECK(hipMemcpyAsync(b.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

extern "C" {

const char* cppe_hip_last_error(void) { return g_last_error.c_str(); }

int cppe_hip_device_count(int* count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return CPPE_HIP_OK;
}

}  // extern "C"

// Builds a system (one segment) or a batch of snapshots (several segments, FMM machinery with a
// per-segment ncrit). Positions and per-site data are in caller order.
static cppe_hip_system* create_system(int n, const double* pos, const double* alpha6, const double* alpha_iso,
                                      const int* excl_off, const int* excl_idx, const cppe_hip_options* opts,
                                      const std::vector<int>& seg_off, const std::vector<int>& seg_ncrit) {
  auto* H   = new cppe_hip_system();
  System& S = H->S;
  try {
    S.n    = n;
    S.opts = *opts;
    CPPE_HIP_CHECK(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    hipStream_t st = S.stream;
    setup_segments(S, seg_off, seg_ncrit);
    dbuf<double> aos;
    upload(aos, pos, 3 * (size_t)n, st);
    S.x.alloc(n);
    S.y.alloc(n);
    S.z.alloc(n);
    hipLaunchKernelGGL(k_scatter_xyz, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, aos.p, S.x.p, S.y.p, S.z.p);
    build_octree(S);
    // per-site data in device order
    dbuf<double> a6h, isoh;
    if (alpha6) upload(a6h, alpha6, 6 * (size_t)n, st);
    if (alpha_iso) upload(isoh, alpha_iso, (size_t)n, st);
    S.has_alpha = alpha6 != nullptr;
    S.all_damped = alpha_iso && std::all_of(alpha_iso, alpha_iso + n, [](double v) { return v > 0.0; });
    S.alpha6.alloc(6 * (size_t)n);
    S.ainv6.alloc(6 * (size_t)n);
    S.s.alloc(n);
    if (!alpha6) {
      S.alpha6.zero(st);
      S.ainv6.zero(st);
    }
    hipLaunchKernelGGL(k_site_setup, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.perm.p,
                       alpha6 ? a6h.p : nullptr, alpha_iso ? isoh.p : nullptr, S.alpha6.p, S.ainv6.p, S.s.p);
    if (excl_off && excl_idx) {
      const int nex = excl_off[n];
      S.h_excl_off.assign(excl_off, excl_off + n + 1);
      S.h_excl_idx.assign(excl_idx, excl_idx + nex);
      dbuf<int> off_h, idx_h, cnt;
      upload(off_h, excl_off, (size_t)n + 1, st);
      upload(idx_h, excl_idx, (size_t)std::max(nex, 1), st);
      cnt.alloc(n + 1);
      cnt.zero(st);
      hipLaunchKernelGGL(k_excl_count, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.perm.p, off_h.p, cnt.p);
      // exclusive scan on the host copy of counts (n + 1 ints): setup, once per system
      std::vector<int> hc(n + 1);
      CPPE_HIP_CHECK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
      CPPE_HIP_CHECK(hipStreamSynchronize(st));
      std::vector<int> ho(n + 1, 0);
      for (int i = 0; i < n; ++i) ho[i + 1] = ho[i] + hc[i];
      upload(S.excl_off, ho.data(), (size_t)n + 1, st);
      S.excl_idx.alloc(std::max(nex, 1));
      hipLaunchKernelGGL(k_excl_fill, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.perm.p, S.iperm.p, off_h.p,
                         idx_h.p, S.excl_off.p, S.excl_idx.p);
      CPPE_HIP_CHECK(hipStreamSynchronize(st));
    }
    build_lists(S);  // direct summation: one leaf whose near list is itself
    const size_t n3 = 3 * (size_t)n;
    S.v_mu.alloc(n3);
    S.v_rhs.alloc(n3);
    S.v_r.alloc(n3);
    S.v_z.alloc(n3);
    S.v_p.alloc(n3);
    S.v_Ap.alloc(n3);
    S.v_t.alloc(n3);
    S.v_d.alloc(n3 * 4);  // room for up to 10 source components (static fields)
    S.io_a.alloc(n3 * 4);
    S.io_b.alloc(n3);
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    delete H;
    throw;
  }
  return H;
}

static int check_create_args(int n, const double* pos, const cppe_hip_options* opts) {
  if (n <= 0 || !pos || !opts) return fail(CPPE_HIP_ERR_ARG, "invalid system arguments");
  if (opts->ncrit < 2) return fail(CPPE_HIP_ERR_ARG, "tree_ncrit must be >= 2");
  if (opts->summation == CPPE_HIP_SUM_FMM && (opts->order < 1 || opts->order > 8))
    return fail(CPPE_HIP_ERR_UNSUPPORTED, "tree_expansion_order must be in 1..8");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CPPE_HIP_ERR_DEVICE, "no HIP device available");
  return CPPE_HIP_OK;
}

extern "C" {

int cppe_hip_system_create(cppe_hip_system** out, int n, const double* pos, const double* alpha6,
                           const double* alpha_iso, const int* excl_off, const int* excl_idx,
                           const cppe_hip_options* opts) {
  CPPE_API_BEGIN
  if (out) *out = nullptr;
  if (!out) return fail(CPPE_HIP_ERR_ARG, "invalid system arguments");
  const int rc = check_create_args(n, pos, opts);
  if (rc != CPPE_HIP_OK) return rc;
  // direct summation is the tree path with a single leaf (ncrit above n): its near list is every
  // pair, and the far field is empty; the expansion order then only sizes unused buffers
  cppe_hip_options o = *opts;
  const bool direct  = o.summation != CPPE_HIP_SUM_FMM;
  if (direct && (o.order < 2 || o.order > 8)) o.order = 2;
  *out = create_system(n, pos, alpha6, alpha_iso, excl_off, excl_idx, &o, {0, n}, {direct ? n + 1 : o.ncrit});
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_batch_create(cppe_hip_system** out, int nseg, const int* seg_off, const int* seg_ncrit,
                          const double* pos, const double* alpha6, const double* alpha_iso, const int* excl_off,
                          const int* excl_idx, const cppe_hip_options* opts) {
  CPPE_API_BEGIN
  if (out) *out = nullptr;
  if (!out || nseg <= 0 || !seg_off || !seg_ncrit) return fail(CPPE_HIP_ERR_ARG, "invalid batch arguments");
  const int n = seg_off[nseg];
  if (seg_off[0] != 0) return fail(CPPE_HIP_ERR_ARG, "segment offsets must start at 0");
  std::vector<int> off(seg_off, seg_off + nseg + 1), nc(nseg);
  for (int g = 0; g < nseg; ++g) {
    const int sz = off[g + 1] - off[g];
    if (sz <= 0) return fail(CPPE_HIP_ERR_ARG, "every snapshot needs at least one site");
    if (seg_ncrit[g] != 0 && seg_ncrit[g] < 2) return fail(CPPE_HIP_ERR_ARG, "tree_ncrit must be >= 2");
    nc[g] = seg_ncrit[g] == 0 ? sz + 1 : seg_ncrit[g];  // 0: whole snapshot in one leaf (direct sum)
  }
  if (excl_off && excl_idx)
    for (int g = 0; g < nseg; ++g)
      for (int i = off[g]; i < off[g + 1]; ++i)
        for (int k = excl_off[i]; k < excl_off[i + 1]; ++k)
          if (excl_idx[k] >= 0 && (excl_idx[k] < off[g] || excl_idx[k] >= off[g + 1]))
            return fail(CPPE_HIP_ERR_ARG, "an exclusion joins two snapshots");
  cppe_hip_options o = *opts;
  o.summation        = CPPE_HIP_SUM_FMM;
  const int rc       = check_create_args(n, pos, &o);
  if (rc != CPPE_HIP_OK) return rc;
  *out = create_system(n, pos, alpha6, alpha_iso, excl_off, excl_idx, &o, off, nc);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_system_destroy(cppe_hip_system* sys) {
  CPPE_API_BEGIN
  delete sys;
  return CPPE_HIP_OK;
  CPPE_API_END
}

static void host_to_soa(System& S, const double* h, int nc, double* d_soa) {
  CPPE_HIP_CHECK(hipMemcpyAsync(S.io_a.p, h, sizeof(double) * nc * (size_t)S.n, hipMemcpyHostToDevice, S.stream));
  aos_to_soa(S, S.io_a.p, d_soa, nc);
}

static void soa_to_host(System& S, const double* d_soa, double* h) {
  soa_to_aos(S, d_soa, S.io_b.p, 3);
  CPPE_HIP_CHECK(hipMemcpyAsync(h, S.io_b.p, sizeof(double) * 3 * (size_t)S.n, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
}

int cppe_hip_field(cppe_hip_system* sys, int m, const double* src, double* F) {
  CPPE_API_BEGIN
  if (!sys || m < 0 || m > 2 || !src || !F) return fail(CPPE_HIP_ERR_ARG, "invalid field arguments");
  System& S    = sys->S;
  if (S.nseg > 1 && m != 1) return fail(CPPE_HIP_ERR_UNSUPPORTED, "order-0/2 fields of a snapshot batch: use multipole_fields");
  const int nc = (m + 1) * (m + 2) / 2;
  host_to_soa(S, src, nc, S.v_d.p);
  if (S.opts.summation == CPPE_HIP_SUM_FMM)
    fmm_field(S, m, S.v_d.p, S.v_t.p, S.opts.damping, false);
  else
    direct_field(S, m, S.v_d.p, S.v_t.p, S.opts.damping);
  soa_to_host(S, S.v_t.p, F);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_multipole_fields(cppe_hip_system* sys, const double* q, const double* mu, const double* theta6,
                              double* F) {
  CPPE_API_BEGIN
  if (!sys || !q || !F) return fail(CPPE_HIP_ERR_ARG, "invalid multipole field arguments");
  System& S = sys->S;
  const int n = S.n;
  // pack q | mu | half-diagonal quadrupole (electric_fields.cc:134-150) into 10 components
  std::vector<double> h(10 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) {
    double* v = &h[10 * (size_t)i];
    v[0]      = q[i];
    if (mu)
      for (int c = 0; c < 3; ++c) v[1 + c] = mu[3 * i + c];
    if (theta6) {
      const double* t = &theta6[6 * (size_t)i];
      v[4] = 0.5 * t[0];
      v[5] = t[1];
      v[6] = t[2];
      v[7] = 0.5 * t[3];
      v[8] = t[4];
      v[9] = 0.5 * t[5];
    }
  }
  S.io_a.alloc(10 * (size_t)n);
  S.v_d.alloc(10 * (size_t)n);
  host_to_soa(S, h.data(), 10, S.v_d.p);
  const double* dq  = S.v_d.p;
  const double* dmu = S.v_d.p + n;
  const double* dth = S.v_d.p + 4 * (size_t)n;
  if (S.opts.summation == CPPE_HIP_SUM_FMM)
    fmm_static(S, dq, dmu, dth, S.v_t.p, S.opts.damping);
  else
    direct_static(S, dq, dmu, dth, S.v_t.p, S.opts.damping);
  soa_to_host(S, S.v_t.p, F);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_apply_b(cppe_hip_system* sys, const double* x, double* out) {
  CPPE_API_BEGIN
  if (!sys || !x || !out) return fail(CPPE_HIP_ERR_ARG, "invalid apply arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  host_to_soa(S, x, 3, S.v_d.p);
  apply_b(S, S.v_d.p, S.v_t.p, S.opts.precision == CPPE_HIP_PREC_MIXED);
  soa_to_host(S, S.v_t.p, out);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_apply_b_multi(cppe_hip_system* sys, int nvec, const double* x, double* out) {
  CPPE_API_BEGIN
  if (!sys || nvec < 1 || !x || !out) return fail(CPPE_HIP_ERR_ARG, "invalid apply arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  if (S.nseg > 1) return fail(CPPE_HIP_ERR_UNSUPPORTED, "several right-hand sides need a one-snapshot system");
  setup_multi_rhs(S, nvec);
  fmm_reserve_multi(S);
  CgState& C      = S.cgm;
  const size_t n3 = 3 * (size_t)S.n;
  for (int k = 0; k < nvec; ++k) host_to_soa(S, x + k * n3, 3, C.p + k * n3);
  apply_b_multi(S, C.p, C.Ap, nvec, S.opts.precision == CPPE_HIP_PREC_MIXED);
  for (int k = 0; k < nvec; ++k) soa_to_host(S, C.Ap + k * n3, out + k * n3);
  return CPPE_HIP_OK;
  CPPE_API_END
}

__global__ void k_alpha_inv_apply(int n, const double* __restrict__ a6, const double* __restrict__ in,
                                  double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m0 = in[i], m1 = in[n + i], m2 = in[2 * n + i];
  const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
               a5 = a6[5 * n + i];
  out[i]         = a0 * m0 + a1 * m1 + a2 * m2;
  out[n + i]     = a1 * m0 + a3 * m1 + a4 * m2;
  out[2 * n + i] = a2 * m0 + a4 * m1 + a5 * m2;
}

int cppe_hip_apply_alpha(cppe_hip_system* sys, const double* x, double* out) {
  CPPE_API_BEGIN
  if (!sys || !x || !out) return fail(CPPE_HIP_ERR_ARG, "invalid apply arguments");
  System& S = sys->S;
  host_to_soa(S, x, 3, S.v_d.p);
  precondition(S, S.v_d.p, S.v_t.p);
  soa_to_host(S, S.v_t.p, out);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_apply_alpha_inv(cppe_hip_system* sys, const double* x, double* out) {
  CPPE_API_BEGIN
  if (!sys || !x || !out) return fail(CPPE_HIP_ERR_ARG, "invalid apply arguments");
  System& S = sys->S;
  host_to_soa(S, x, 3, S.v_d.p);
  hipLaunchKernelGGL(k_alpha_inv_apply, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, S.ainv6.p,
                     S.v_d.p, S.v_t.p);
  soa_to_host(S, S.v_t.p, out);
  return CPPE_HIP_OK;
  CPPE_API_END
}

}  // extern "C"

namespace cppe_dev {
__global__ __launch_bounds__(256) void k_dense_b(int n, const double* __restrict__ pos,
                                                 const double* __restrict__ iso,
                                                 const double* __restrict__ ainv_aos, double damping,
                                                 const int* __restrict__ eoff, const int* __restrict__ eidx,
                                                 const int* __restrict__ index, double* __restrict__ out);
__global__ __launch_bounds__(256) void k_nuclear(int n, const double* __restrict__ pos, int na,
                                                 const double* __restrict__ ch, const double* __restrict__ apos,
                                                 double* __restrict__ F);
__global__ __launch_bounds__(256) void k_nuc_mul_energy(int n, const double* __restrict__ pos,
                                                        const int* __restrict__ kmax,
                                                        const double* __restrict__ mul, int na,
                                                        const double* __restrict__ ch,
                                                        const double* __restrict__ apos,
                                                        double* __restrict__ part);
__global__ void k_sum_partials(int np, const double* __restrict__ part, double* __restrict__ out);
__global__ __launch_bounds__(256) void k_nuc_field_grad(int n, const double* __restrict__ pos, int na,
                                                        const double* __restrict__ ch,
                                                        const double* __restrict__ apos, double* __restrict__ G);
__global__ __launch_bounds__(256) void k_nuc_mul_grad(int n, const double* __restrict__ pos,
                                                      const int* __restrict__ kmax, const double* __restrict__ mul,
                                                      const double* __restrict__ ch,
                                                      const double* __restrict__ apos, double* __restrict__ part);
__global__ __launch_bounds__(256) void k_grad_sum(int nb, const double* __restrict__ part, double* __restrict__ out);
__global__ __launch_bounds__(256) void k_dot_part(long long n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ part);
__global__ void k_soa6_to_aos9(int n, const int* __restrict__ perm, const double* __restrict__ a6,
                               double* __restrict__ out9) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  const double v[6] = {a6[i], a6[n + i], a6[2 * n + i], a6[3 * n + i], a6[4 * n + i], a6[5 * n + i]};
  const int map[9]  = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  for (int k = 0; k < 9; ++k) out9[9 * (size_t)j + k] = v[map[k]];
}
}  // namespace cppe_dev

extern "C" {

static void dense_b_device(System& S, const int* index, dbuf<double>& dout) {
  const int n    = S.n;
  hipStream_t st = S.stream;
  // caller-order positions and isotropic alpha; exclusions with their original values
  dbuf<double> pos, iso, ainv9;
  dbuf<int> didx, eoff, eidx;
  std::vector<double> hx(n), hy(n), hz(n);
  const std::vector<int>& perm = S.h_perm;
  CPPE_HIP_CHECK(hipMemcpyAsync(hx.data(), S.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(hy.data(), S.y.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(hz.data(), S.z.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  std::vector<double> a6(6 * (size_t)n);
  CPPE_HIP_CHECK(hipMemcpyAsync(a6.data(), S.alpha6.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<double> hp(3 * (size_t)n), hiso(n);
  for (int i = 0; i < n; ++i) {
    const int j = perm[i];
    hp[3 * j] = hx[i];
    hp[3 * j + 1] = hy[i];
    hp[3 * j + 2] = hz[i];
    hiso[j] = (a6[i] + a6[3 * (size_t)n + i] + a6[5 * (size_t)n + i]) / 3.0;
  }
  std::vector<int> co = S.h_excl_off, ci = S.h_excl_idx;
  if (co.empty()) co.assign(n + 1, 0);
  if (ci.empty()) ci.push_back(-1);
  upload(pos, hp.data(), hp.size(), st);
  upload(iso, hiso.data(), hiso.size(), st);
  upload(eoff, co.data(), co.size(), st);
  upload(eidx, ci.data(), ci.size(), st);
  upload(didx, index, (size_t)n, st);
  ainv9.alloc(9 * (size_t)n);
  hipLaunchKernelGGL(k_soa6_to_aos9, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.perm.p, S.ainv6.p, ainv9.p);
  const size_t N = 3 * (size_t)n;
  dout.alloc(N * N);
  const long long np = (long long)n * n;
  hipLaunchKernelGGL(k_dense_b, dim3(ceil_div(np, 256)), dim3(256), 0, st, n, pos.p, iso.p, ainv9.p,
                     S.opts.damping, eoff.p, eidx.p, didx.p, dout.p);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
}

int cppe_hip_dense_b(cppe_hip_system* sys, const int* index, double* out) {
  CPPE_API_BEGIN
  if (!sys || !index || !out) return fail(CPPE_HIP_ERR_ARG, "invalid dense arguments");
  if (sys->S.nseg > 1) return fail(CPPE_HIP_ERR_UNSUPPORTED, "dense B is not available for snapshot batches");
  dbuf<double> d;
  dense_b_device(sys->S, index, d);
  const size_t N = 3 * (size_t)sys->S.n;
  CPPE_HIP_CHECK(hipMemcpy(out, d.p, sizeof(double) * N * N, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_dense_b_inverse(cppe_hip_system* sys, const int* index, double* out) {
  CPPE_API_BEGIN
  if (!sys || !index || !out) return fail(CPPE_HIP_ERR_ARG, "invalid dense arguments");
  if (sys->S.nseg > 1) return fail(CPPE_HIP_ERR_UNSUPPORTED, "dense B is not available for snapshot batches");
  if (sys->S.n > 4000) return fail(CPPE_HIP_ERR_UNSUPPORTED, "dense inverse limited to 4000 sites");
  dbuf<double> d, inv;
  dense_b_device(sys->S, index, d);
  const int N = 3 * sys->S.n;
  inv.alloc((size_t)N * N);
  dense_inverse(sys->S.stream, N, d.p, inv.p);
  CPPE_HIP_CHECK(hipMemcpy(out, inv.p, sizeof(double) * (size_t)N * N, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

static int run_solve(System& S, int guess_mode, const double* guess_host, double thresh, int maxiter,
                     int* iters, double* norms, double* device_ms) {
  hipStream_t st = S.stream;
  const size_t n3 = 3 * (size_t)S.n;
  hipEvent_t e0, e1;
  CPPE_HIP_CHECK(hipEventCreate(&e0));
  CPPE_HIP_CHECK(hipEventCreate(&e1));
  if (guess_mode == 3 && guess_host) host_to_soa(S, guess_host, 3, S.v_mu.p);
  CPPE_HIP_CHECK(hipEventRecord(e0, st));
  if (guess_mode == 0)
    precondition(S, S.v_rhs.p, S.v_mu.p);
  else if (guess_mode == 2 || (guess_mode == 1 && !S.mu_valid))
    fill_f64(S, S.v_mu.p, n3, 0.0);
  CgResult& res = S.last;
  device_cg(S, S.cg, thresh, maxiter, S.opts.precision == CPPE_HIP_PREC_MIXED, res);
  CPPE_HIP_CHECK(hipEventRecord(e1, st));
  CPPE_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CPPE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (device_ms) *device_ms = ms;
  if (iters) *iters = S.nseg == 1 ? res.iters[0] : res.loop_iters;
  if (norms)
    for (size_t k = 0; k < res.norms0.size() && (int)k < maxiter; ++k) norms[k] = res.norms0[k];
  S.mu_valid = true;
  if (!res.all_converged) return fail(CPPE_HIP_ERR_NOCONV, "Failed to converge induced dipole moments.");
  return CPPE_HIP_OK;
}

int cppe_hip_solve(cppe_hip_system* sys, const double* rhs, const double* guess, int make_guess,
                   double thresh, int maxiter, double* mu, int* iters, double* norms) {
  CPPE_API_BEGIN
  if (!sys || !rhs || !mu || maxiter < 0) return fail(CPPE_HIP_ERR_ARG, "invalid solve arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  host_to_soa(S, rhs, 3, S.v_rhs.p);
  int mode = make_guess ? 0 : (guess ? 3 : 2);
  int rc   = run_solve(S, mode, guess, thresh, maxiter, iters, norms, nullptr);
  soa_to_host(S, S.v_mu.p, mu);
  return rc;
  CPPE_API_END
}

int cppe_hip_solve_multi(cppe_hip_system* sys, int nrhs, const double* rhs, const double* guess, int make_guess,
                         double thresh, int maxiter, double* mu, int* iters, double* norms) {
  CPPE_API_BEGIN
  if (!sys || nrhs < 1 || !rhs || !mu || maxiter < 0) return fail(CPPE_HIP_ERR_ARG, "invalid solve arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  if (S.nseg > 1) return fail(CPPE_HIP_ERR_UNSUPPORTED, "several right-hand sides need a one-snapshot system");
  setup_multi_rhs(S, nrhs);
  CgState& C      = S.cgm;
  const size_t n3 = 3 * (size_t)S.n;
  for (int k = 0; k < nrhs; ++k) {
    host_to_soa(S, rhs + k * n3, 3, C.rhs + k * n3);
    if (make_guess)
      precondition(S, C.rhs + k * n3, C.x + k * n3);
    else if (guess)
      host_to_soa(S, guess + k * n3, 3, C.x + k * n3);
    else
      fill_f64(S, C.x + k * n3, n3, 0.0);
  }
  CgResult res;
  device_cg(S, C, thresh, maxiter, S.opts.precision == CPPE_HIP_PREC_MIXED, res);
  for (int k = 0; k < nrhs; ++k) {
    soa_to_host(S, C.x + k * n3, mu + k * n3);
    if (iters) iters[k] = res.iters[k];
    if (norms)
      for (size_t it = 0; it < res.norms[k].size() && (int)it < maxiter; ++it)
        norms[(size_t)k * maxiter + it] = res.norms[k][it];
  }
  if (!res.all_converged) return fail(CPPE_HIP_ERR_NOCONV, "Failed to converge induced dipole moments.");
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_upload_rhs(cppe_hip_system* sys, const double* rhs) {
  CPPE_API_BEGIN
  if (!sys || !rhs) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  host_to_soa(sys->S, rhs, 3, sys->S.v_rhs.p);
  CPPE_HIP_CHECK(hipStreamSynchronize(sys->S.stream));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_solve_resident(cppe_hip_system* sys, int guess_mode, double thresh, int maxiter, int* iters,
                            double* norms, double* device_ms) {
  CPPE_API_BEGIN
  if (!sys || guess_mode < 0 || guess_mode > 2) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  return run_solve(sys->S, guess_mode, nullptr, thresh, maxiter, iters, norms, device_ms);
  CPPE_API_END
}

int cppe_hip_field_targets(cppe_hip_system* sys, const double* mu, int nt, const int* targets, double* F) {
  CPPE_API_BEGIN
  if (!sys || nt < 0 || (nt > 0 && (!targets || !F))) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S = sys->S;
  if (nt == 0) return CPPE_HIP_OK;
  std::vector<int> td(nt);
  for (int t = 0; t < nt; ++t) {
    if (targets[t] < 0 || targets[t] >= S.n) return fail(CPPE_HIP_ERR_ARG, "target index out of range");
    td[t] = S.iperm_host(targets[t]);
  }
  const double* src = S.v_mu.p;
  if (mu) {
    host_to_soa(S, mu, 3, S.v_d.p);
    src = S.v_d.p;
  } else if (!S.mu_valid) {
    return fail(CPPE_HIP_ERR_ARG, "system has no resident moments");
  }
  dbuf<int> dt;
  dbuf<double> dout;
  upload(dt, td.data(), (size_t)nt, S.stream);
  dout.alloc(3 * (size_t)nt);
  field1_targets(S, src, nt, dt.p, dout.p);
  CPPE_HIP_CHECK(hipMemcpyAsync(F, dout.p, sizeof(double) * 3 * (size_t)nt, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_segment_results(cppe_hip_system* sys, int* iters, double* rnorm, int* converged) {
  CPPE_API_BEGIN
  if (!sys) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  const CgResult& r = sys->S.last;
  if ((int)r.iters.size() != sys->S.nseg) return fail(CPPE_HIP_ERR_ARG, "no solve has run on this system");
  for (int g = 0; g < sys->S.nseg; ++g) {
    if (iters) iters[g] = r.iters[g];
    if (rnorm) rnorm[g] = r.rnorm[g];
    if (converged) converged[g] = r.converged[g];
  }
  return CPPE_HIP_OK;
  CPPE_API_END
}

__global__ void k_fill_pattern(int n3, double* v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) v[k] = 0.1 * sin(0.37 * (double)k + 1.0);  // fixed pseudo-random dipoles
}

int cppe_hip_autotune(cppe_hip_system* sys, int reps, int* choice, double* ms) {
  CPPE_API_BEGIN
  if (!sys || reps <= 0 || !choice || !ms) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  // time on dipoles of realistic bit activity (the clock under load depends on the data)
  hipLaunchKernelGGL(k_fill_pattern, dim3(ceil_div(3 * S.n, 256)), dim3(256), 0, S.stream, 3 * S.n, S.v_d.p);
  fmm_autotune(S, reps, choice, ms);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_tune(cppe_hip_system* sys, const int* choice) {
  CPPE_API_BEGIN
  if (!sys || !choice) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  for (int prec = 0; prec < 2; ++prec) {
    const int v = choice[4 * prec], b = choice[4 * prec + 1], g = choice[4 * prec + 2], o = choice[4 * prec + 3];
    if (v < 0 || v > 11 || (b != 64 && b != 128 && b != 256) || (v == 2 && b != 64) ||
        (g != 16 && g != 32 && g != 64) || !overlap_code_valid(o))
      return fail(CPPE_HIP_ERR_ARG, "invalid launch shape");
    if (!near_variant_allowed(sys->S, v))
      return fail(CPPE_HIP_ERR_UNSUPPORTED, "near-field variant 2 needs leaves of at most 1024 sites");
  }
  fmm_set_tune(sys->S, choice);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_reliable_drop(cppe_hip_system* sys, double drop) {
  CPPE_API_BEGIN
  if (!sys || !(drop > 0.0 && drop < 1.0)) return fail(CPPE_HIP_ERR_ARG, "reliable drop must be in (0, 1)");
  sys->S.reliable_drop = drop;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_refinement(cppe_hip_system* sys, double drop, int stall, double min_progress) {
  CPPE_API_BEGIN
  if (!sys || !(drop > 0.0 && drop < 1.0) || stall < 1 || !(min_progress >= 0.0))
    return fail(CPPE_HIP_ERR_ARG, "refinement: drop in (0, 1), stall >= 1, min_progress >= 0");
  sys->S.reliable_drop       = drop;
  sys->S.refine_stall        = stall;
  sys->S.refine_min_progress = min_progress;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_confirmation(cppe_hip_system* sys, double eta, int forms) {
  CPPE_API_BEGIN
  if (!sys || !(eta >= 0.0) || !std::isfinite(eta) || forms < 0 || forms > 3)
    return fail(CPPE_HIP_ERR_ARG, "confirmation: eta >= 0, forms in 0..3");
  sys->S.confirm_eta   = eta;
  sys->S.confirm_forms = forms;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_last_solve_confirmations(cppe_hip_system* sys, int* count) {
  CPPE_API_BEGIN
  if (!sys || !count) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  count[0] = sys->S.last.confirmations;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_preconditioner(cppe_hip_system* sys, int kind, double radius, double weight) {
  CPPE_API_BEGIN
  if (!sys || (kind != CPPE_HIP_PRECOND_JACOBI && kind != CPPE_HIP_PRECOND_LOCAL) || !(radius > 0.0) ||
      !(weight >= 0.0))
    return fail(CPPE_HIP_ERR_ARG, "preconditioner: kind 0 or 1, radius > 0, weight >= 0");
  System& S = sys->S;
  if (kind == CPPE_HIP_PRECOND_LOCAL && !S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  S.precond        = kind;
  S.precond_rc     = radius;
  S.precond_weight = weight;
  cg_graphs_release(S);  // the captured iterations hold the preconditioner's buffers and weight
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_last_solve_info(cppe_hip_system* sys, int* info) {
  CPPE_API_BEGIN
  if (!sys || !info) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  info[0] = sys->S.last.loop_iters;
  info[1] = sys->S.last.reliable_updates;
  info[2] = sys->S.last.fp64_fallback;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_graphs(cppe_hip_system* sys, int enable) {
  CPPE_API_BEGIN
  if (!sys) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  sys->S.use_graphs = enable != 0;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_set_graph_steps(cppe_hip_system* sys, int steps) {
  CPPE_API_BEGIN
  if (!sys || steps < 1 || steps > CgState::kMaxGraphSteps)
    return fail(CPPE_HIP_ERR_ARG, "steps per launch must be 1 to 8");
  sys->S.graph_steps = steps;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_last_solve_launches(cppe_hip_system* sys, int* launches) {
  CPPE_API_BEGIN
  if (!sys || !launches) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  launches[0] = sys->S.last.step_launches;
  launches[1] = sys->S.last.steps_launched;
  launches[2] = sys->S.last.host_reads;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_download_mu(cppe_hip_system* sys, double* mu) {
  CPPE_API_BEGIN
  if (!sys || !mu) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  soa_to_host(sys->S, sys->S.v_mu.p, mu);
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_profile_apply(cppe_hip_system* sys, int fp32, int reps, double* phase_ms) {
  CPPE_API_BEGIN
  if (!sys || reps <= 0 || !phase_ms) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S = sys->S;
  if (!S.has_alpha) return fail(CPPE_HIP_ERR_ARG, "system has no polarizabilities");
  hipStream_t st = S.stream;
  const int n3   = 3 * S.n;
  hipLaunchKernelGGL(k_fill_pattern, dim3(ceil_div(n3, 256)), dim3(256), 0, st, n3, S.v_d.p);
  apply_b(S, S.v_d.p, S.v_t.p, fp32 != 0);  // warm (allocations)
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  hipEvent_t ev[8];
  for (auto& e : ev) CPPE_HIP_CHECK(hipEventCreate(&e));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int var = fmm_near_variant(S, fp32 != 0);
  const bool fmm = true;  // both summations apply B through the tree path
  for (int r = 0; r < reps; ++r) {
    if (fmm) S.prof = ev;
    CPPE_HIP_CHECK(hipEventRecord(ev[5], st));
    if (!fmm) CPPE_HIP_CHECK(hipEventRecord(ev[3], st));
    apply_b(S, S.v_d.p, S.v_t.p, fp32 != 0);
    if (!fmm) CPPE_HIP_CHECK(hipEventRecord(ev[4], st));
    S.prof = nullptr;
    CPPE_HIP_CHECK(hipEventSynchronize(ev[4]));
    float t = 0.f;
    if (fmm) {
      CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
      acc[0] += t;
      CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[1], ev[2]));
      acc[1] += t;
      CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[2], ev[3]));
      acc[2] += t;
    }
    CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[3], ev[4]));
    acc[3] += t;
    CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[5], ev[4]));
    acc[4] += t;
    if (var >= 5) {
      CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[6], ev[7]));
      acc[5] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 5; ++k) phase_ms[k] = acc[k] / reps;
  phase_ms[5] = var >= 5 ? acc[5] / reps : -1.0;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_nuclear_fields(int n, const double* pos, int natoms, const double* charges, const double* atom_pos,
                            double* F) {
  CPPE_API_BEGIN
  if (n < 0 || natoms < 0) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  if (n == 0) return CPPE_HIP_OK;
  if (natoms == 0) {
    std::fill(F, F + 3 * (size_t)n, 0.0);
    return CPPE_HIP_OK;
  }
  hipStream_t st = nullptr;
  dbuf<double> dp, dc, da, dF;
  upload(dp, pos, 3 * (size_t)n, st);
  upload(dc, charges, (size_t)natoms, st);
  upload(da, atom_pos, 3 * (size_t)natoms, st);
  dF.alloc(3 * (size_t)n);
  hipLaunchKernelGGL(k_nuclear, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, dp.p, natoms, dc.p, da.p, dF.p);
  CPPE_HIP_CHECK(hipMemcpy(F, dF.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_nuclear_multipole_energy(int n, const double* pos, const int* kmax, const double* mul, int natoms,
                                      const double* charges, const double* atom_pos, double* energy) {
  CPPE_API_BEGIN
  if (n < 0 || natoms < 0 || !energy) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  *energy = 0.0;
  if (n == 0 || natoms == 0) return CPPE_HIP_OK;
  hipStream_t st = nullptr;
  dbuf<double> dp, dm, dc, da, part, res;
  dbuf<int> dk;
  upload(dp, pos, 3 * (size_t)n, st);
  upload(dm, mul, 10 * (size_t)n, st);
  upload(dk, kmax, (size_t)n, st);
  upload(dc, charges, (size_t)natoms, st);
  upload(da, atom_pos, 3 * (size_t)natoms, st);
  const int nb = ceil_div(n, 256);
  part.alloc(nb);
  res.alloc(1);
  hipLaunchKernelGGL(k_nuc_mul_energy, dim3(nb), dim3(256), 0, st, n, dp.p, dk.p, dm.p, natoms, dc.p, da.p, part.p);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, nb, part.p, res.p);
  CPPE_HIP_CHECK(hipMemcpy(energy, res.p, sizeof(double), hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_nuclear_field_gradient(int n, const double* pos, int natoms, const double* charges,
                                    const double* atom_pos, double* grad) {
  CPPE_API_BEGIN
  if (n < 0 || natoms < 0 || (n > 0 && natoms > 0 && (!pos || !charges || !atom_pos || !grad)))
    return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  if (n == 0 || natoms == 0) return CPPE_HIP_OK;
  hipStream_t st = nullptr;
  dbuf<double> dp, dc, da, dG;
  upload(dp, pos, 3 * (size_t)n, st);
  upload(dc, charges, (size_t)natoms, st);
  upload(da, atom_pos, 3 * (size_t)natoms, st);
  const size_t total = 9 * (size_t)natoms * (size_t)n;
  dG.alloc(total);
  hipLaunchKernelGGL(k_nuc_field_grad, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, dp.p, natoms, dc.p, da.p, dG.p);
  CPPE_HIP_CHECK(hipMemcpy(grad, dG.p, sizeof(double) * total, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_nuclear_multipole_gradient(int n, const double* pos, const int* kmax, const double* mul, int natoms,
                                        const double* charges, const double* atom_pos, double* grad) {
  CPPE_API_BEGIN
  if (n < 0 || natoms < 0 || (natoms > 0 && !grad)) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  if (natoms == 0) return CPPE_HIP_OK;
  std::fill(grad, grad + 3 * (size_t)natoms, 0.0);
  if (n == 0) return CPPE_HIP_OK;
  if (!pos || !kmax || !mul || !charges || !atom_pos) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  hipStream_t st = nullptr;
  dbuf<double> dp, dm, dc, da, part, res;
  dbuf<int> dk;
  upload(dp, pos, 3 * (size_t)n, st);
  upload(dm, mul, 10 * (size_t)n, st);
  upload(dk, kmax, (size_t)n, st);
  upload(dc, charges, (size_t)natoms, st);
  upload(da, atom_pos, 3 * (size_t)natoms, st);
  const int nb = ceil_div(n, 256);
  part.alloc(3 * (size_t)nb * natoms);
  res.alloc(3 * (size_t)natoms);
  hipLaunchKernelGGL(k_nuc_mul_grad, dim3(nb, natoms), dim3(256), 0, st, n, dp.p, dk.p, dm.p, dc.p, da.p, part.p);
  hipLaunchKernelGGL(k_grad_sum, dim3(natoms), dim3(256), 0, st, nb, part.p, res.p);
  CPPE_HIP_CHECK(hipMemcpy(grad, res.p, sizeof(double) * 3 * natoms, hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_dot(long long n, const double* a, const double* b, double* out) {
  CPPE_API_BEGIN
  if (n < 0 || !out) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  *out = 0.0;
  if (n == 0) return CPPE_HIP_OK;
  hipStream_t st = nullptr;
  dbuf<double> da, db, part, res;
  upload(da, a, (size_t)n, st);
  upload(db, b, (size_t)n, st);
  const int nb = (int)std::min<long long>(512, (n + 255) / 256);
  part.alloc(nb);
  res.alloc(1);
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(256), 0, st, n, da.p, db.p, part.p);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, nb, part.p, res.p);
  CPPE_HIP_CHECK(hipMemcpy(out, res.p, sizeof(double), hipMemcpyDeviceToHost));
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_tree_stats(cppe_hip_system* sys, long long* stats) {
  CPPE_API_BEGIN
  if (!sys || !stats) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  const Octree& T = sys->S.tree;
  stats[0] = T.ncells;
  stats[1] = T.nleaves;
  stats[2] = T.n_m2l;
  stats[3] = T.n_p2p;
  stats[4] = T.max_level;
  stats[5] = T.p2p_site_pairs;
  return CPPE_HIP_OK;
  CPPE_API_END
}

int cppe_hip_tree_export(cppe_hip_system* sys, double* geo, long long* meta, long long* cell_start, int* perm,
                         long long* m2l, long long* p2p) {
  CPPE_API_BEGIN
  if (!sys) return fail(CPPE_HIP_ERR_ARG, "invalid arguments");
  System& S      = sys->S;
  Octree& T      = S.tree;
  hipStream_t st = S.stream;
  const int nc   = T.ncells;
  auto dl = [&](const auto& b, size_t count) {
    using T_ = std::remove_pointer_t<decltype(b.p)>;
    std::vector<T_> h(count);
    if (count) CPPE_HIP_CHECK(hipMemcpyAsync(h.data(), b.p, sizeof(T_) * count, hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
    return h;
  };
  if (geo) {
    auto x = dl(T.cx, nc), y = dl(T.cy, nc), z = dl(T.cz, nc), r = dl(T.cr, nc), rm = dl(T.crmax, nc);
    for (int c = 0; c < nc; ++c) {
      geo[5 * c] = x[c];
      geo[5 * c + 1] = y[c];
      geo[5 * c + 2] = z[c];
      geo[5 * c + 3] = r[c];
      geo[5 * c + 4] = rm[c];
    }
  }
  if (meta) {
    auto pa = dl(T.parent, nc), lv = dl(T.level, nc), om = dl(T.octmask, nc), ct = dl(T.count, nc);
    for (int c = 0; c < nc; ++c) {
      meta[4 * c] = pa[c];
      meta[4 * c + 1] = lv[c];
      meta[4 * c + 2] = om[c];
      meta[4 * c + 3] = ct[c];
    }
  }
  if (cell_start) {
    auto s0 = dl(T.start, nc);
    for (int c = 0; c < nc; ++c) cell_start[c] = s0[c];
  }
  if (perm)
    for (int i = 0; i < S.n; ++i) perm[i] = S.h_perm[i];
  if (m2l && T.n_m2l) {
    auto off = dl(T.m2l_off, nc + 1), src = dl(T.m2l_src, T.n_m2l);
    for (int c = 0; c < nc; ++c)
      for (int k = off[c]; k < off[c + 1]; ++k) {
        m2l[2 * k]     = src[k];
        m2l[2 * k + 1] = c;
      }
  }
  if (p2p && T.n_p2p) {
    auto off = dl(T.p2p_off, nc + 1), src = dl(T.p2p_src, T.n_p2p);
    for (int c = 0; c < nc; ++c)
      for (int k = off[c]; k < off[c + 1]; ++k) {
        p2p[2 * k]     = c;
        p2p[2 * k + 1] = src[k];
      }
  }
  return CPPE_HIP_OK;
  CPPE_API_END
}

}  // extern "C"
