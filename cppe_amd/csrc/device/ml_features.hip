// Per-site inputs of the learned CG warm start (no counterpart in the reference).
//
// For every polarizable site i (device order) with permanent field F_i (the solver's rhs):
//   basis vectors (rotation-equivariant):
//     v1 = alpha_i F_i                                   (the reference's first guess)
//     v2 = alpha_i sum_{j in N(i)} fc(r) T2d_ij alpha_j F_j     (local 2nd Neumann term)
//     v3 = alpha_i sum_{j in N(i)} fc(r) T2d_ij v2_j             (local 3rd Neumann term)
//   with N(i) = sites within rc = 6 A (11.338 bohr), excluded pairs skipped, T2d the
//   Thole-damped dipole tensor of the system and fc(r) = (cos(pi r/rc) + 1)/2;
//   64 invariant descriptors (feature k):
//     0-7   sum_j g_k(r)                        radial Gaussians, g_k = exp(-eta (r - r_k)^2) fc(r)
//     8-15  sum_j g_k(r) alpha_iso_j / 5        polarizability-weighted radial density
//     16-23 sum_j g_k(r) (rhat . Fhat_i)        angular, first moment about the local field
//     24-31 sum_j g_k(r) P2(rhat . Fhat_i)      angular, second moment
//     32-39 sum_j g_k(r) (Fhat_j . Fhat_i)      field alignment of the neighbours
//     40-47 sum_j g_k(r) (rhat . Fhat_j)        neighbour fields along the bond
//     48-50 F_i / |F_i| in the site's molecular frame (from its first two exclusion partners)
//     51    log |F_i|,  52 alpha_iso_i / 5,  53 (Fhat alpha Fhat) / alpha_iso
//     54    sum_j fc(r),  55 sum_j fc(r) / r^3 (bohr^-3 x 100)
//     56    (v2 . v1)/|v1|^2,  57 |v2|/|v1|,  58 (v3 . v1)/|v1|^2,  59 |v3|/|v1|
//     60    (v3 . v2)/(|v2||v3|),  61 log|v1|,  62 (v1 . F)/(|v1||F|),  63 1
// Neighbours come from a uniform grid of cell size rc built on the device (sort by cell).
#include <hipcub/hipcub.hpp>

#include "ml.hh"
#include "pairs.hh"

namespace cppe_dev {

constexpr double kRc    = 11.338356747600001;  // 6 A in bohr (6 * 1.8897261246)
constexpr double kRk0   = 1.5;                 // first radial centre (bohr)
constexpr double kRkStep = 1.4;                // centre spacing (bohr)
constexpr double kEta   = 0.5;                 // Gaussian width parameter (bohr^-2)

struct GridInfo {
  double x0, y0, z0;  // lower corner
  int nx, ny, nz;
};

__global__ void k_grid_cell(int n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, GridInfo g, unsigned* __restrict__ cell,
                            int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx = min(g.nx - 1, max(0, (int)floor((x[i] - g.x0) / kRc)));
  int cy = min(g.ny - 1, max(0, (int)floor((y[i] - g.y0) / kRc)));
  int cz = min(g.nz - 1, max(0, (int)floor((z[i] - g.z0) / kRc)));
  cell[i] = (unsigned)((cz * g.ny + cy) * g.nx + cx);
  idx[i]  = i;
}

__global__ void k_grid_ranges(int n, int ncell, const unsigned* __restrict__ cell, int* __restrict__ cstart) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ncell) return;
  // cstart[k] = first sorted position with cell >= k
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cell[mid] < (unsigned)k)
      lo = mid + 1;
    else
      hi = mid;
  }
  cstart[k] = lo;
}

__device__ __forceinline__ void site_alpha(const double* a6, int n, int i, double a[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) a[k] = a6[(size_t)k * n + i];
}

__device__ __forceinline__ void amul(const double a[6], double vx, double vy, double vz, double& ox, double& oy,
                                     double& oz) {
  ox = a[0] * vx + a[1] * vy + a[2] * vz;
  oy = a[1] * vx + a[3] * vy + a[4] * vz;
  oz = a[2] * vx + a[4] * vy + a[5] * vz;
}

// T2d(d) w with Thole damping (same screening as the response operator): 3 fT d (d.w)/r^5 - fE w/r^3
__device__ __forceinline__ void t2_apply(double dx, double dy, double dz, double r, double a, double wx,
                                         double wy, double wz, double& ox, double& oy, double& oz) {
  const double ri  = 1.0 / r;
  const double ri3 = ri * ri * ri;
  double fE = 1.0, fT = 1.0;
  if (a > 0.0) {
    const double u = a * r;
    if (u < kTholeCutoff) thole_ET(u, exp(-u), fE, fT);
  }
  const double dw = dx * wx + dy * wy + dz * wz;
  const double c5 = 3.0 * fT * dw * ri3 * ri * ri;
  ox              = c5 * dx - fE * ri3 * wx;
  oy              = c5 * dy - fE * ri3 * wy;
  oz              = c5 * dz - fE * ri3 * wz;
}

struct FeatArgs {
  int n;
  const double *x, *y, *z, *s, *a6, *F;  // F: 3n SoA
  const int* eoff;
  const int* eidx;
  double damping;
  GridInfo g;
  const int* cstart;
  const int* sorted;   // site indices sorted by grid cell
  const unsigned* cellof;
  float* X;            // n x 64 raw features
  double* V;           // 9n: v1 (3n SoA), v2, v3
};

// pass 1: descriptors 0-55, v1, v2.  pass 2: v3, descriptors 56-63.
template <int PASS>
__global__ __launch_bounds__(128) void k_ml_features(FeatArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const double Fx = a.F[i], Fy = a.F[n + i], Fz = a.F[2 * (size_t)n + i];
  const double Fn = sqrt(Fx * Fx + Fy * Fy + Fz * Fz) + 1e-30;
  const double fx = Fx / Fn, fy = Fy / Fn, fz = Fz / Fn;
  double ai[6];
  site_alpha(a.a6, n, i, ai);
  const double iso_i = (ai[0] + ai[3] + ai[5]) / 3.0;
  const double di    = a.damping * a.s[i];
  float feat[56];
  if (PASS == 1)
    for (int k = 0; k < 56; ++k) feat[k] = 0.f;
  double sx = 0.0, sy = 0.0, sz = 0.0;  // sum fc T2d w_j
  const double* W = PASS == 1 ? nullptr : a.V + 3 * (size_t)n;  // v2 for pass 2
  const unsigned c = a.cellof[i];
  const int cx = (int)(c % a.g.nx), cy = (int)((c / a.g.nx) % a.g.ny), cz = (int)(c / ((unsigned)a.g.nx * a.g.ny));
  const int e0 = a.eoff ? a.eoff[i] : 0, e1 = a.eoff ? a.eoff[i + 1] : 0;
  for (int dzc = -1; dzc <= 1; ++dzc)
    for (int dyc = -1; dyc <= 1; ++dyc)
      for (int dxc = -1; dxc <= 1; ++dxc) {
        const int qx = cx + dxc, qy = cy + dyc, qz = cz + dzc;
        if (qx < 0 || qy < 0 || qz < 0 || qx >= a.g.nx || qy >= a.g.ny || qz >= a.g.nz) continue;
        const int q = (qz * a.g.ny + qy) * a.g.nx + qx;
        for (int t = a.cstart[q]; t < a.cstart[q + 1]; ++t) {
          const int j = a.sorted[t];
          if (j == i) continue;
          const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;  // i -> j
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 >= kRc * kRc) continue;
          const double r  = sqrt(r2);
          const double fc = 0.5 * (cos(M_PI * r / kRc) + 1.0);
          bool ex = false;
          for (int k = e0; k < e1; ++k) ex |= (a.eidx[k] == j);
          if (PASS == 1) {
            const double Fjx = a.F[j], Fjy = a.F[n + j], Fjz = a.F[2 * (size_t)n + j];
            const double Fjn = sqrt(Fjx * Fjx + Fjy * Fjy + Fjz * Fjz) + 1e-30;
            double aj[6];
            site_alpha(a.a6, n, j, aj);
            const double iso_j = (aj[0] + aj[3] + aj[5]) / 3.0;
            const double cth   = (dx * fx + dy * fy + dz * fz) / r;
            const double p2    = 1.5 * cth * cth - 0.5;
            const double al    = (Fjx * fx + Fjy * fy + Fjz * fz) / Fjn;
            const double bj    = (dx * Fjx + dy * Fjy + dz * Fjz) / (r * Fjn);
            for (int k = 0; k < 8; ++k) {
              const double rk = kRk0 + kRkStep * k;
              const double gk = exp(-kEta * (r - rk) * (r - rk)) * fc;
              feat[k] += (float)gk;
              feat[8 + k] += (float)(gk * iso_j * 0.2);
              feat[16 + k] += (float)(gk * cth);
              feat[24 + k] += (float)(gk * p2);
              feat[32 + k] += (float)(gk * al);
              feat[40 + k] += (float)(gk * bj);
            }
            feat[54] += (float)fc;
            feat[55] += (float)(fc / (r2 * r) * 100.0);
            if (!ex) {
              double wx, wy, wz, tx, ty, tz;
              amul(aj, Fjx, Fjy, Fjz, wx, wy, wz);
              t2_apply(dx, dy, dz, r, di * a.s[j], wx, wy, wz, tx, ty, tz);
              sx += fc * tx;
              sy += fc * ty;
              sz += fc * tz;
            }
          } else if (!ex) {
            double tx, ty, tz;
            t2_apply(dx, dy, dz, r, di * a.s[j], W[j], W[n + j], W[2 * (size_t)n + j], tx, ty, tz);
            sx += fc * tx;
            sy += fc * ty;
            sz += fc * tz;
          }
        }
      }
  double v1x, v1y, v1z;
  amul(ai, Fx, Fy, Fz, v1x, v1y, v1z);
  double ox, oy, oz;
  amul(ai, sx, sy, sz, ox, oy, oz);
  float* Xi = a.X + (size_t)i * kMlpIn;
  if (PASS == 1) {
    a.V[i] = v1x;
    a.V[n + i] = v1y;
    a.V[2 * (size_t)n + i] = v1z;
    a.V[3 * (size_t)n + i] = ox;
    a.V[4 * (size_t)n + i] = oy;
    a.V[5 * (size_t)n + i] = oz;
    // molecular frame from the first two exclusion partners (fallback: field frame)
    double e1x = fx, e1y = fy, e1z = fz, e2x = 0, e2y = 0, e2z = 0;
    if (e1 - e0 >= 2) {
      const int p = a.eidx[e0], q = a.eidx[e0 + 1];
      if (p >= 0 && q >= 0) {
        double ux = a.x[p] - xi, uy = a.y[p] - yi, uz = a.z[p] - zi;
        double un = sqrt(ux * ux + uy * uy + uz * uz) + 1e-30;
        e1x = ux / un;
        e1y = uy / un;
        e1z = uz / un;
        double wx = a.x[q] - xi, wy = a.y[q] - yi, wz = a.z[q] - zi;
        double pr = wx * e1x + wy * e1y + wz * e1z;
        wx -= pr * e1x;
        wy -= pr * e1y;
        wz -= pr * e1z;
        double wn = sqrt(wx * wx + wy * wy + wz * wz) + 1e-30;
        e2x = wx / wn;
        e2y = wy / wn;
        e2z = wz / wn;
      }
    }
    const double e3x = e1y * e2z - e1z * e2y, e3y = e1z * e2x - e1x * e2z, e3z = e1x * e2y - e1y * e2x;
    for (int k = 0; k < 48; ++k) Xi[k] = feat[k];
    Xi[48] = (float)(fx * e1x + fy * e1y + fz * e1z);
    Xi[49] = (float)(fx * e2x + fy * e2y + fz * e2z);
    Xi[50] = (float)(fx * e3x + fy * e3y + fz * e3z);
    Xi[51] = (float)log(Fn);
    Xi[52] = (float)(iso_i * 0.2);
    double afx, afy, afz;
    amul(ai, fx, fy, fz, afx, afy, afz);
    Xi[53] = (float)((afx * fx + afy * fy + afz * fz) / iso_i);
    Xi[54] = feat[54];
    Xi[55] = feat[55];
  } else {
    a.V[6 * (size_t)n + i] = ox;
    a.V[7 * (size_t)n + i] = oy;
    a.V[8 * (size_t)n + i] = oz;
    const double v2x = a.V[3 * (size_t)n + i], v2y = a.V[4 * (size_t)n + i], v2z = a.V[5 * (size_t)n + i];
    const double n1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z) + 1e-30;
    const double n2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z) + 1e-30;
    const double n3 = sqrt(ox * ox + oy * oy + oz * oz) + 1e-30;
    Xi[56] = (float)((v2x * v1x + v2y * v1y + v2z * v1z) / (n1 * n1));
    Xi[57] = (float)(n2 / n1);
    Xi[58] = (float)((ox * v1x + oy * v1y + oz * v1z) / (n1 * n1));
    Xi[59] = (float)(n3 / n1);
    Xi[60] = (float)((ox * v2x + oy * v2y + oz * v2z) / (n2 * n3));
    Xi[61] = (float)log(n1);
    Xi[62] = (float)((v1x * Fx + v1y * Fy + v1z * Fz) / (n1 * Fn));
    Xi[63] = 1.0f;
  }
}

__global__ __launch_bounds__(256) void k_bbox(int n, const double* __restrict__ x, const double* __restrict__ y,
                                              const double* __restrict__ z, double* __restrict__ out) {
  __shared__ double r[6][256];
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int c = 0; c < 3; ++c) {
      mn[c] = fmin(mn[c], v[c]);
      mx[c] = fmax(mx[c], v[c]);
    }
  }
  for (int c = 0; c < 3; ++c) {
    r[c][threadIdx.x]     = mn[c];
    r[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; ++c) {
        r[c][threadIdx.x]     = fmin(r[c][threadIdx.x], r[c][threadIdx.x + s]);
        r[3 + c][threadIdx.x] = fmax(r[3 + c][threadIdx.x], r[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = r[threadIdx.x][0];
}

// Features X (n x 64, raw) and bases V (9n) for the system's resident rhs.
void ml_features(System& S, float* X, double* V) {
  const int n    = S.n;
  hipStream_t st = S.stream;
  dbuf<double> bb;
  bb.alloc(6);
  hipLaunchKernelGGL(k_bbox, dim3(1), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, bb.p);
  double h[6];
  CPPE_HIP_CHECK(hipMemcpyAsync(h, bb.p, sizeof(h), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  GridInfo g;
  g.x0 = h[0];
  g.y0 = h[1];
  g.z0 = h[2];
  g.nx = std::max(1, (int)std::ceil((h[3] - h[0]) / kRc + 1e-9));
  g.ny = std::max(1, (int)std::ceil((h[4] - h[1]) / kRc + 1e-9));
  g.nz = std::max(1, (int)std::ceil((h[5] - h[2]) / kRc + 1e-9));
  const long long ncell = (long long)g.nx * g.ny * g.nz;
  if (ncell > (1ll << 26)) throw std::runtime_error("ml features: neighbour grid too large");
  dbuf<unsigned> cell, cell_s;
  dbuf<int> idx, idx_s, cstart;
  cell.alloc(n);
  cell_s.alloc(n);
  idx.alloc(n);
  idx_s.alloc(n);
  cstart.alloc(ncell + 1);
  hipLaunchKernelGGL(k_grid_cell, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, g, cell.p,
                     idx.p);
  dbuf<char> tmp;
  size_t bytes = 0;
  int bits = 1;
  while ((1ll << bits) <= ncell) ++bits;
  CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, cell.p, cell_s.p, idx.p, idx_s.p, n, 0, bits, st));
  tmp.alloc(bytes + 16);
  CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, cell.p, cell_s.p, idx.p, idx_s.p, n, 0, bits, st));
  hipLaunchKernelGGL(k_grid_ranges, dim3(ceil_div(ncell + 1, 256)), dim3(256), 0, st, n, (int)ncell, cell_s.p,
                     cstart.p);
  FeatArgs a{n, S.x.p, S.y.p, S.z.p, S.s.p, S.alpha6.p, S.v_rhs.p, S.excl_off.p, S.excl_idx.p,
             S.opts.damping, g, cstart.p, idx_s.p, cell.p, X, V};
  hipLaunchKernelGGL(k_ml_features<1>, dim3(ceil_div(n, 128)), dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_ml_features<2>, dim3(ceil_div(n, 128)), dim3(128), 0, st, a);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace cppe_dev
