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
// Neighbours come from a uniform grid of cell size rc/2 built on the device once per system
// (sort by cell; a neighbourhood is 25 contiguous rows of 5 cells), stored as a neighbour list of
// indices (the exclusion flag in bit 31). Each pass recomputes the pair geometry from the
// neighbour's site record: the displacement formed in FP64 and rounded, the rest in FP32 --
// reading a cached FP32 tensor block (24 B per pair, ~650 MB at 300k sites) cost a pass more than
// the ~30 instructions that form it. The descriptors: the eight radial Gaussians by the recurrence
// g_{k+1} = g_k A B_k, A = exp(2 eta dr (r - r_0)), B_k = exp(-eta dr^2 (2k + 1)), two exponentials
// per neighbour. The bases: Thole-damped dipole tensor as in the response operator (FP32),
// neighbour vectors gathered as FP32 records, sums in FP64 (a guess needs no more than FP32
// accuracy).
#include <hipcub/hipcub.hpp>

#include "ml.hh"
#include "pairs.hh"

namespace cppe_dev {

constexpr double kRc     = 11.338356747600001;  // 6 A in bohr (6 * 1.8897261246)
constexpr double kRk0    = 1.5;                 // first radial centre (bohr)
constexpr double kRkStep = 1.4;                 // centre spacing (bohr)
constexpr double kEta    = 0.5;                 // Gaussian width parameter (bohr^-2)
constexpr double kCell   = kRc / 2;             // grid cell edge
constexpr int kReach     = 2;                   // cells scanned on each side

void ml_grid_destroy(System& S) {
  delete S.ml_grid;
  S.ml_grid = nullptr;
}

__global__ void k_grid_cell(int n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, GridInfo g, unsigned* __restrict__ cell,
                            int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx = min(g.nx - 1, max(0, (int)floor((x[i] - g.x0) / kCell)));
  int cy = min(g.ny - 1, max(0, (int)floor((y[i] - g.y0) / kCell)));
  int cz = min(g.nz - 1, max(0, (int)floor((z[i] - g.z0) / kCell)));
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

// Visits every site j != i within kRc of site i (rows of contiguous cells of the grid);
// f(j, dx, dy, dz, r2) with d = r_j - r_i.
template <typename Fn>
__device__ __forceinline__ void for_neighbours(const FeatArgs& a, int i, double xi, double yi, double zi, Fn&& f) {
  const unsigned c = a.cellof[i];
  const int cx = (int)(c % a.g.nx), cy = (int)((c / a.g.nx) % a.g.ny), cz = (int)(c / ((unsigned)a.g.nx * a.g.ny));
  const int x0 = max(0, cx - kReach), x1 = min(a.g.nx - 1, cx + kReach);
  for (int qz = max(0, cz - kReach); qz <= min(a.g.nz - 1, cz + kReach); ++qz)
    for (int qy = max(0, cy - kReach); qy <= min(a.g.ny - 1, cy + kReach); ++qy) {
      const int row = (qz * a.g.ny + qy) * a.g.nx;
      const int t1  = a.cstart[row + x1 + 1];
      for (int t = a.cstart[row + x0]; t < t1; ++t) {
        const int j = a.sorted[t];
        if (j == i) continue;
        const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= kRc * kRc) continue;
        f(j, dx, dy, dz, r2);
      }
    }
}

// Neighbour list: count, then fill (pairs within rc, j != i; excluded pairs kept with a zero
// tensor block so that the densities still see them).
__global__ __launch_bounds__(128) void k_ml_ncount(FeatArgs a, long long* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  long long c = 0;
  for_neighbours(a, i, a.x[i], a.y[i], a.z[i], [&](int, double, double, double, double) { ++c; });
  cnt[i] = c;
}

// Per-site records of the list passes: position and alpha_iso^(-1/6) (0: undamped site).
__global__ void k_ml_pos(int n, const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, const double* __restrict__ s, double4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double4(x[i], y[i], z[i], s ? s[i] : 0.0);
}

// Neighbour indices (pairs within rc, j != i; excluded pairs flagged in bit 31: the densities
// still see them, the bases take no tensor block for them).
__global__ __launch_bounds__(128) void k_ml_nfill(FeatArgs a, const long long* __restrict__ noff, int* __restrict__ nj,
                                                  float* __restrict__ iso) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  iso[i] = (float)((a.a6[i] + a.a6[3 * (size_t)n + i] + a.a6[5 * (size_t)n + i]) / 3.0);
  const int e0 = a.eoff ? a.eoff[i] : 0, e1 = a.eoff ? a.eoff[i + 1] : 0;
  long long p = noff[i];
  for_neighbours(a, i, a.x[i], a.y[i], a.z[i], [&](int j, double, double, double, double) {
    bool ex = false;
    for (int k = e0; k < e1; ++k) ex |= (a.eidx[k] == j);
    nj[p] = ex ? (int)((unsigned)j | 0x80000000u) : j;
    ++p;
  });
}

// Pair geometry of the list passes: d = r_j - r_i formed in FP64 and rounded, r, 1/r in FP32.
struct MlPair {
  float dx, dy, dz, r2, ri;
};
__device__ __forceinline__ MlPair ml_pair(const double4& pi, const double4& pj) {
  MlPair q;
  q.dx = (float)(pj.x - pi.x);
  q.dy = (float)(pj.y - pi.y);
  q.dz = (float)(pj.z - pi.z);
  q.r2 = q.dx * q.dx + q.dy * q.dy + q.dz * q.dz;
  q.ri = __builtin_amdgcn_rsqf(q.r2);
  return q;
}

// cutoff function fc(r) = (cos(pi r / rc) + 1) / 2 through the hardware cosine (argument in turns)
__device__ __forceinline__ float ml_fc(float r) {
  return 0.5f * (__builtin_amdgcn_cosf(r * (float)(0.5 / kRc)) + 1.0f);
}

struct ListArgs {
  int n;
  double damping;  // Thole factor of the bases
  const long long* noff;
  const int* nj;
  const double4* pos;
  const float* iso;
  float4* frec;  // per site: F / |F| and alpha_iso (written by k_ml_v1)
  float4* w1;    // per site: v1 in FP32 (k_ml_v1)
  float4* w2;    // per site: v2 in FP32 (k_ml_desc)
  const double *x, *y, *z, *a6, *F;
  const int* eoff;
  const int* eidx;
  float* X;
  double* V;
};

// Lanes per site of the list kernels: consecutive lanes take consecutive pairs of a site's
// neighbour list (coalesced), partial sums are combined by a fixed xor-shuffle tree.
constexpr int kDescLanes = 8, kBasisLanes = 16;

template <int G>
__device__ __forceinline__ float lane_sum(float v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
template <int G>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One pair of the response bases: acc += K_ij w_j, K_ij = fc(r) T2d_ij, the Thole-damped dipole
// tensor (3 fT d d^T / r^5 - fE / r^3) of the response operator times the cutoff function, formed
// in FP32 (sj: Thole scale of site j; di: damping times that of site i); the product summed in FP64.
__device__ __forceinline__ void basis_acc(const MlPair& q, float r, float fc, float di, double sj, const float4& w,
                                          double& sx, double& sy, double& sz) {
  constexpr float kLog2e = 1.44269504088896340736f;
  const float ri2 = q.ri * q.ri, ri3 = ri2 * q.ri;
  const float u   = di * (float)sj * r;  // Thole argument; 0: a pair with an undamped site
  const float e   = u > 0.0f ? __builtin_amdgcn_exp2f(-u * kLog2e) : 0.0f;  // e = 0: fE = fT = 1
  const float u2  = u * u;
  const float fE  = 1.0f - (1.0f + u + 0.5f * u2) * e;
  const float fT  = fE - (u2 * u * (1.0f / 6.0f)) * e;
  const float c3  = fE * ri3 * fc, c5 = 3.0f * fT * ri3 * ri2 * fc;
  const float dw  = q.dx * w.x + q.dy * w.y + q.dz * w.z;
  // K w = c5 (d . w) d - c3 w
  sx += (double)(c5 * dw * q.dx - c3 * w.x);
  sy += (double)(c5 * dw * q.dy - c3 * w.y);
  sz += (double)(c5 * dw * q.dz - c3 * w.z);
}

// Descriptors 0-47, 54 and 55 (FP32 arithmetic) over the neighbour list, kDescLanes lanes per site,
// and in the same pass over the list the first response basis v2 = alpha_i sum_j K_ij v1_j (pairs
// that are not excluded).
__global__ __launch_bounds__(256) void k_ml_desc(ListArgs a) {
  const int gid  = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0   = gid / kDescLanes;
  const int sub  = gid % kDescLanes;
  const bool ok  = i0 < a.n;
  const int i    = ok ? i0 : a.n - 1;  // trailing lanes shadow the last site and store nothing
  const int n = a.n;
  const double4 pi = a.pos[i];
  const double Fx = a.F[i], Fy = a.F[n + i], Fz = a.F[2 * (size_t)n + i];
  const double Fn = sqrt(Fx * Fx + Fy * Fy + Fz * Fz) + 1e-30;
  const float ffx = (float)(Fx / Fn), ffy = (float)(Fy / Fn), ffz = (float)(Fz / Fn);
  const float di   = (float)(a.damping * pi.w);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  float feat[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) feat[k] = 0.f;
  float f54 = 0.f, f55 = 0.f;
  constexpr float kA    = (float)(2.0 * kEta * kRkStep);
  const long long p1 = ok ? a.noff[i + 1] : 0;
  // the lane's pairs in order, two at a time (both indices, then both records, requested first)
  for (long long pb = a.noff[i] + sub; pb < p1; pb += 2 * kDescLanes) {
    const bool has2 = pb + kDescLanes < p1;
    const int n0    = a.nj[pb];  // negative: an excluded pair (no basis term)
    const int n1    = has2 ? a.nj[pb + kDescLanes] : n0;
    const int j0 = n0 & 0x7fffffff, j1 = n1 & 0x7fffffff;
    const double4 pq[2] = {a.pos[j0], a.pos[j1]};
    const float4 fq[2]  = {a.frec[j0], a.frec[j1]};
    const float4 wq[2]  = {a.w1[j0], a.w1[j1]};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
    if (b == 1 && !has2) break;
    const MlPair q  = ml_pair(pi, pq[b]);
    const float ri  = q.ri;
    const float r   = q.r2 * ri;
    const float ux = q.dx * ri, uy = q.dy * ri, uz = q.dz * ri;
    const float fc  = ml_fc(r);
    if ((b == 0 ? n0 : n1) >= 0) basis_acc(q, r, fc, di, pq[b].w, wq[b], sx, sy, sz);
    const float4 fj = fq[b];  // F_j / |F_j|, alpha_iso_j
    const float iso = fj.w;
    const float cth = ux * ffx + uy * ffy + uz * ffz;
    const float p2  = 1.5f * cth * cth - 0.5f;
    const float al  = fj.x * ffx + fj.y * ffy + fj.z * ffz;
    const float bj  = ux * fj.x + uy * fj.y + uz * fj.z;
    const float u   = r - (float)kRk0;
    float g         = __expf(-(float)kEta * u * u) * fc;
    const float A   = __expf(kA * u);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      feat[k] += g;
      feat[8 + k] += g * iso * 0.2f;
      feat[16 + k] += g * cth;
      feat[24 + k] += g * p2;
      feat[32 + k] += g * al;
      feat[40 + k] += g * bj;
      g *= A * (float)exp(-kEta * kRkStep * kRkStep * (2 * k + 1));
    }
    f54 += fc;
    f55 += fc * ri * ri * ri * 100.0f;
    }
  }
#pragma unroll
  for (int k = 0; k < 48; ++k) feat[k] = lane_sum<kDescLanes>(feat[k]);
  f54 = lane_sum<kDescLanes>(f54);
  f55 = lane_sum<kDescLanes>(f55);
  sx  = lane_sum<kDescLanes>(sx);
  sy  = lane_sum<kDescLanes>(sy);
  sz  = lane_sum<kDescLanes>(sz);
  if (!ok || sub != 0) return;
  float* Xi = a.X + (size_t)i * kMlpIn;
#pragma unroll
  for (int k = 0; k < 48; ++k) Xi[k] = feat[k];
  Xi[54] = f54;
  Xi[55] = f55;
  double ai[6], ox, oy, oz;
  site_alpha(a.a6, n, i, ai);
  amul(ai, sx, sy, sz, ox, oy, oz);
  a.V[3 * (size_t)n + i] = ox;
  a.V[4 * (size_t)n + i] = oy;
  a.V[5 * (size_t)n + i] = oz;
  a.w2[i]                = make_float4((float)ox, (float)oy, (float)oz, 0.0f);
}

// v1 = alpha F, the per-site FP32 records the list kernels gather, and the per-site descriptors
__global__ void k_ml_v1(ListArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  const double Fx = a.F[i], Fy = a.F[n + i], Fz = a.F[2 * (size_t)n + i];
  double ai[6], ox, oy, oz;
  site_alpha(a.a6, n, i, ai);
  amul(ai, Fx, Fy, Fz, ox, oy, oz);
  a.V[i]                 = ox;
  a.V[n + i]             = oy;
  a.V[2 * (size_t)n + i] = oz;
  a.w1[i]                = make_float4((float)ox, (float)oy, (float)oz, 0.0f);
  const double Fi = 1.0 / (sqrt(Fx * Fx + Fy * Fy + Fz * Fz) + 1e-30);
  a.frec[i]       = make_float4((float)(Fx * Fi), (float)(Fy * Fi), (float)(Fz * Fi), a.iso[i]);
  // descriptors 48-53: the field direction in the molecular frame of the first two exclusion
  // partners (fallback: field frame), log |F|, alpha_iso / 5, F^ alpha F^ / alpha_iso
  const double Fn = sqrt(Fx * Fx + Fy * Fy + Fz * Fz) + 1e-30;
  const double fx = Fx / Fn, fy = Fy / Fn, fz = Fz / Fn;
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  float* Xi = a.X + (size_t)i * kMlpIn;
  const double iso_i = (ai[0] + ai[3] + ai[5]) / 3.0;
  const int e0 = a.eoff ? a.eoff[i] : 0, e1 = a.eoff ? a.eoff[i + 1] : 0;
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
  Xi[48] = (float)(fx * e1x + fy * e1y + fz * e1z);
  Xi[49] = (float)(fx * e2x + fy * e2y + fz * e2z);
  Xi[50] = (float)(fx * e3x + fy * e3y + fz * e3z);
  Xi[51] = (float)log(Fn);
  Xi[52] = (float)(iso_i * 0.2);
  double afx, afy, afz;
  amul(ai, fx, fy, fz, afx, afy, afz);
  Xi[53] = (float)((afx * fx + afy * fy + afz * fz) / iso_i);
}

// The second response basis v3 = alpha_i sum_j K_ij v2_j (K as in basis_acc; excluded pairs: none)
// and descriptors 56-63, kBasisLanes lanes per site.
__global__ __launch_bounds__(256) void k_ml_basis(ListArgs a) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0  = gid / kBasisLanes;
  const int sub = gid % kBasisLanes;
  const bool ok = i0 < a.n;
  const int i   = ok ? i0 : a.n - 1;  // trailing lanes shadow the last site and store nothing
  const int n = a.n;
  const double4 pi = a.pos[i];
  const float di   = (float)(a.damping * pi.w);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const long long p1 = ok ? a.noff[i + 1] : 0;
  // the lane's pairs p, p + G, p + 2G, ... in order, four at a time: every index, then every
  // record, is requested before the first pair is evaluated
  constexpr int kU = 4;
  for (long long pb = a.noff[i] + sub; pb < p1; pb += kU * kBasisLanes) {
    int jb[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const long long p = pb + (long long)u * kBasisLanes;
      jb[u]             = p < p1 ? a.nj[p] : -1;  // negative: excluded pair or none
    }
    double4 pb4[kU];
    float4 wb[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int jj = jb[u] < 0 ? i : jb[u];
      pb4[u]       = a.pos[jj];
      wb[u]        = a.w2[jj];
    }
#pragma unroll
    for (int b = 0; b < kU; ++b) {
      if (jb[b] < 0) continue;
      const MlPair q = ml_pair(pi, pb4[b]);
      const float r  = q.r2 * q.ri;
      basis_acc(q, r, ml_fc(r), di, pb4[b].w, wb[b], sx, sy, sz);
    }
  }
  sx = lane_sum<kBasisLanes>(sx);
  sy = lane_sum<kBasisLanes>(sy);
  sz = lane_sum<kBasisLanes>(sz);
  if (!ok || sub != 0) return;
  double ai[6], ox, oy, oz;
  site_alpha(a.a6, n, i, ai);
  amul(ai, sx, sy, sz, ox, oy, oz);
  a.V[6 * (size_t)n + i] = ox;
  a.V[7 * (size_t)n + i] = oy;
  a.V[8 * (size_t)n + i] = oz;
  const double Fx = a.F[i], Fy = a.F[n + i], Fz = a.F[2 * (size_t)n + i];
  const double Fn = sqrt(Fx * Fx + Fy * Fy + Fz * Fz) + 1e-30;
  const double v1x = a.V[i], v1y = a.V[n + i], v1z = a.V[2 * (size_t)n + i];
  const double v2x = a.V[3 * (size_t)n + i], v2y = a.V[4 * (size_t)n + i], v2z = a.V[5 * (size_t)n + i];
  const double n1 = sqrt(v1x * v1x + v1y * v1y + v1z * v1z) + 1e-30;
  const double n2 = sqrt(v2x * v2x + v2y * v2y + v2z * v2z) + 1e-30;
  const double n3 = sqrt(ox * ox + oy * oy + oz * oz) + 1e-30;
  float* Xi = a.X + (size_t)i * kMlpIn;
  Xi[56] = (float)((v2x * v1x + v2y * v1y + v2z * v1z) / (n1 * n1));
  Xi[57] = (float)(n2 / n1);
  Xi[58] = (float)((ox * v1x + oy * v1y + oz * v1z) / (n1 * n1));
  Xi[59] = (float)(n3 / n1);
  Xi[60] = (float)((ox * v2x + oy * v2y + oz * v2z) / (n2 * n3));
  Xi[61] = (float)log(n1);
  Xi[62] = (float)((v1x * Fx + v1y * Fy + v1z * Fz) / (n1 * Fn));
  Xi[63] = 1.0f;
}

// Bounding box of the sites: per-block partial min/max, then one block over the partials.
__global__ __launch_bounds__(256) void k_bbox_part(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, double* __restrict__ part) {
  __shared__ double r[6][256];
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
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
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = r[threadIdx.x][0];
}

MlGrid& ml_grid(System& S) {
  if (S.ml_grid) return *S.ml_grid;
  const int n    = S.n;
  hipStream_t st = S.stream;
  constexpr int kParts = 256;
  dbuf<double> part;
  part.alloc(6 * kParts);
  hipLaunchKernelGGL(k_bbox_part, dim3(kParts), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, part.p);
  std::vector<double> hp(6 * kParts);
  CPPE_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  double h[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int b = 0; b < kParts; ++b)
    for (int c = 0; c < 3; ++c) {
      h[c]     = std::min(h[c], hp[6 * b + c]);
      h[3 + c] = std::max(h[3 + c], hp[6 * b + 3 + c]);
    }
  auto* G = new MlGrid();
  try {
    GridInfo& g = G->g;
    g.x0 = h[0];
    g.y0 = h[1];
    g.z0 = h[2];
    g.nx = std::max(1, (int)std::ceil((h[3] - h[0]) / kCell + 1e-9));
    g.ny = std::max(1, (int)std::ceil((h[4] - h[1]) / kCell + 1e-9));
    g.nz = std::max(1, (int)std::ceil((h[5] - h[2]) / kCell + 1e-9));
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    if (ncell > (1ll << 26)) throw std::runtime_error("ml features: neighbour grid too large");
    dbuf<unsigned> cell;
    dbuf<int> idx;
    cell.alloc(n);
    idx.alloc(n);
    G->cellof.alloc(n);
    G->sorted.alloc(n);
    G->cstart.alloc(ncell + 1);
    hipLaunchKernelGGL(k_grid_cell, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, g, cell.p,
                       idx.p);
    dbuf<char> tmp;
    size_t bytes = 0;
    int bits     = 1;
    while ((1ll << bits) <= ncell) ++bits;
    dbuf<unsigned> cell_s;
    cell_s.alloc(n);
    CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, cell.p, cell_s.p, idx.p, G->sorted.p, n, 0,
                                                      bits, st));
    tmp.alloc(bytes + 16);
    CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, cell.p, cell_s.p, idx.p, G->sorted.p, n, 0,
                                                      bits, st));
    hipLaunchKernelGGL(k_grid_ranges, dim3(ceil_div(ncell + 1, 256)), dim3(256), 0, st, n, (int)ncell, cell_s.p,
                       G->cstart.p);
    CPPE_HIP_CHECK(hipMemcpyAsync(G->cellof.p, cell.p, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, st));
    // neighbour list (geometry only)
    FeatArgs fa{n, S.x.p, S.y.p, S.z.p, S.s.p, S.alpha6.p, nullptr, S.excl_off.p, S.excl_idx.p,
                S.opts.damping, g, G->cstart.p, G->sorted.p, G->cellof.p, nullptr, nullptr};
    dbuf<long long> cnt;
    cnt.alloc(n + 1);
    hipLaunchKernelGGL(k_ml_ncount, dim3(ceil_div(n, 128)), dim3(128), 0, st, fa, cnt.p);
    long long zero = 0;
    CPPE_HIP_CHECK(hipMemcpyAsync(cnt.p + n, &zero, sizeof(long long), hipMemcpyHostToDevice, st));
    G->noff.alloc(n + 1);
    size_t sb = 0;
    CPPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, cnt.p, G->noff.p, n + 1, st));
    tmp.alloc(sb + 16);
    CPPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, sb, cnt.p, G->noff.p, n + 1, st));
    CPPE_HIP_CHECK(hipMemcpyAsync(&G->npairs, G->noff.p + n, sizeof(long long), hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
    const size_t np = (size_t)std::max<long long>(G->npairs, 1);
    if (G->npairs >= (1ll << 31)) throw std::runtime_error("ml features: more neighbour pairs than 32-bit indices");
    G->nj.alloc(np);
    G->iso.alloc(n);
    G->pos.alloc(n);
    hipLaunchKernelGGL(k_ml_nfill, dim3(ceil_div(n, 128)), dim3(128), 0, st, fa, G->noff.p, G->nj.p, G->iso.p);
    hipLaunchKernelGGL(k_ml_pos, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, S.s.p,
                       G->pos.p);
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    delete G;
    throw;
  }
  S.ml_grid = G;
  return *G;
}

// Features X (n x 64, raw) and bases V (9n) for the system's resident rhs.
void ml_features(System& S, float* X, double* V) {
  const int n    = S.n;
  hipStream_t st = S.stream;
  MlGrid& G      = ml_grid(S);
  G.frec.alloc(n);
  G.w1.alloc(n);
  G.w2.alloc(n);
  ListArgs a{n, S.opts.damping, G.noff.p, G.nj.p, G.pos.p, G.iso.p, G.frec.p, G.w1.p, G.w2.p, S.x.p, S.y.p, S.z.p,
             S.alpha6.p, S.v_rhs.p, S.excl_off.p, S.excl_idx.p, X, V};
  hipLaunchKernelGGL(k_ml_v1, dim3(ceil_div(n, 256)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_ml_desc, dim3(ceil_div((long long)n * kDescLanes, 256)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_ml_basis, dim3(ceil_div((long long)n * kBasisLanes, 256)), dim3(256), 0, st, a);
}

}  // namespace cppe_dev
