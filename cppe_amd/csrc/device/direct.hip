// Direct-summation field kernels on gfx950 (the reference's compute_field_exact path,
// src/core/fmm/calculate.cc:262-300, and the nuclear / energy / dense-B helpers).
//
// Layout: sites in device (Morton) order, every per-site quantity SoA. One thread owns one
// target site; the block (256 threads = 4 wavefronts) streams source tiles of 256 sites
// through LDS, so each source value is read once per block and broadcast from LDS.
// Exclusions: a tile is checked per lane against its own index and CSR exclusion list; only
// wavefronts with a hit in the tile run the masked inner loop (rare with Morton order).
// Thole damping (pair parameter a = damping * s_i * s_j, s = alpha_iso^(-1/6), s = 0 for
// undamped sites): e^-aR is evaluated only below the screening cutoff (common.hh).
// The response operator B of a direct-summation system does not use these kernels: it runs on the
// tree path with a single leaf (every pair in the near field), see solver.hip's apply_b.
#include "pairs.hh"
#include "system.hh"

namespace cppe_dev {

constexpr int kTile = 256;

// ---------------------------------------------------------------------------------------------
// caller AoS (caller order) <-> device SoA (device order)
// ---------------------------------------------------------------------------------------------
__global__ void k_aos_to_soa(int n, int nc, const int* __restrict__ perm, const double* __restrict__ aos,
                             double* __restrict__ soa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  for (int c = 0; c < nc; ++c) soa[(size_t)c * n + i] = aos[(size_t)nc * j + c];
}

__global__ void k_soa_to_aos(int n, int nc, const int* __restrict__ perm, const double* __restrict__ soa,
                             double* __restrict__ aos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  for (int c = 0; c < nc; ++c) aos[(size_t)nc * j + c] = soa[(size_t)c * n + i];
}

void aos_to_soa(System& S, const double* aos, double* soa, int nc) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, nc, S.perm.p, aos,
                     soa);
}

void soa_to_aos(System& S, const double* soa, double* aos, int nc) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, nc, S.perm.p, soa,
                     aos);
}

// ---------------------------------------------------------------------------------------------
// exclusion helpers
// ---------------------------------------------------------------------------------------------
// Field of order-m sources (general, FP64): used for cppe_hip_field (m = 0, 1, 2).
__device__ __forceinline__ void pair_m(int m, double dx, double dy, double dz, const double* Sq, double a,
                                       bool damp, double& fx, double& fy, double& fz) {
  if (m == 1) {
    if (damp)
      pair_m1_f64<true>(dx, dy, dz, Sq[0], Sq[1], Sq[2], a, fx, fy, fz);
    else
      pair_m1_f64<false>(dx, dy, dz, Sq[0], Sq[1], Sq[2], a, fx, fy, fz);
    return;
  }
  double zero6[6] = {0, 0, 0, 0, 0, 0};
  if (m == 0)
    pair_static(dx, dy, dz, Sq[0], 0.0, 0.0, 0.0, zero6, a, damp, fx, fy, fz);
  else
    pair_static(dx, dy, dz, 0.0, 0.0, 0.0, 0.0, Sq, a, damp, fx, fy, fz);
}

// ---------------------------------------------------------------------------------------------
// Generic direct field: m in {0,1,2} (nsrc = 1,3,6 components) or static (nsrc = 10: q, mu,
// S6, combined with the sign convention of MultipoleFields). FP64, one-shot paths.
// ---------------------------------------------------------------------------------------------
template <int KIND>  // 0,1,2 = order m ; 3 = static combined
__global__ __launch_bounds__(256) void k_direct_generic(int n, const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const double* __restrict__ src,
                                                        const double* __restrict__ s, double damping,
                                                        const int* __restrict__ eoff,
                                                        const int* __restrict__ eidx,
                                                        double* __restrict__ F) {
  constexpr int NS = KIND == 0 ? 1 : (KIND == 1 ? 3 : (KIND == 2 ? 6 : 10));
  __shared__ double sx[kTile], sy[kTile], sz[kTile], ss[kTile];
  __shared__ double sS[NS][kTile];
  const int t  = threadIdx.x;
  const int i  = blockIdx.x * kTile + t;
  const bool act = i < n;
  const double xi = act ? x[i] : 0.0, yi = act ? y[i] : 0.0, zi = act ? z[i] : 0.0;
  const double di = act && s ? damping * s[i] : 0.0;
  const bool damp = damping > 0.0 && s != nullptr;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j0 = 0; j0 < n; j0 += kTile) {
    const int tn = min(kTile, n - j0);
    __syncthreads();
    if (t < tn) {
      sx[t] = x[j0 + t];
      sy[t] = y[j0 + t];
      sz[t] = z[j0 + t];
      ss[t] = s ? s[j0 + t] : 0.0;
      for (int c = 0; c < NS; ++c) sS[c][t] = src[(size_t)c * n + j0 + t];
    }
    __syncthreads();
    const bool hit = act && tile_hits(i, j0, j0 + tn, eoff, eidx);
    const bool masked = __ballot(hit) != 0;
    for (int jj = 0; jj < tn; ++jj) {
      if (masked && (!act || excluded(i, j0 + jj, eoff, eidx))) continue;
      const double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
      const double a  = di * ss[jj];
      double Sq[NS];
      for (int c = 0; c < NS; ++c) Sq[c] = sS[c][jj];
      if (KIND == 3)
        pair_static(dx, dy, dz, Sq[0], Sq[1], Sq[2], Sq[3], &Sq[NS == 10 ? 4 : 0], a, damp, fx, fy, fz);
      else
        pair_m(KIND, dx, dy, dz, Sq, a, damp, fx, fy, fz);
    }
  }
  if (act) {
    F[i]         = fx;
    F[n + i]     = fy;
    F[2 * n + i] = fz;
  }
}

void direct_field(System& S, int m, const double* src, double* F, double damping) {
  const dim3 g(ceil_div(S.n, kTile)), b(kTile);
  const double* s = damping > 0.0 ? S.s.p : nullptr;
  const int* eo = S.excl_off.p;
  const int* ei = S.excl_idx.p;
  if (m == 0)
    hipLaunchKernelGGL(k_direct_generic<0>, g, b, 0, S.stream, S.n, S.x.p, S.y.p, S.z.p, src, s, damping, eo, ei, F);
  else if (m == 1)
    hipLaunchKernelGGL(k_direct_generic<1>, g, b, 0, S.stream, S.n, S.x.p, S.y.p, S.z.p, src, s, damping, eo, ei, F);
  else
    hipLaunchKernelGGL(k_direct_generic<2>, g, b, 0, S.stream, S.n, S.x.p, S.y.p, S.z.p, src, s, damping, eo, ei, F);
}

// src layout for the static field: q (n), mu (3n), S6 (6n) consecutive SoA blocks.
void direct_static(System& S, const double* q, const double* mu, const double* th, double* F,
                   double damping) {
  (void)mu;
  (void)th;
  const double* s = damping > 0.0 ? S.s.p : nullptr;
  hipLaunchKernelGGL(k_direct_generic<3>, dim3(ceil_div(S.n, kTile)), dim3(kTile), 0, S.stream, S.n, S.x.p,
                     S.y.p, S.z.p, q, s, damping, S.excl_off.p, S.excl_idx.p, F);
}

// ---------------------------------------------------------------------------------------------
// Dipole field (P2P<1,3> convention, Thole-damped with the system's factor, exclusions as in the
// system) at selected targets by FP64 direct summation over every site of the target's snapshot:
// one block per target, sources strided over the threads, fixed-order block reduction. Used to
// measure the tree path's error against direct summation at sampled sites.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_field1_targets(int nt, const int* __restrict__ tdev, int n,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const double* __restrict__ s,
                                                        double damping, const int* __restrict__ eoff,
                                                        const int* __restrict__ eidx,
                                                        const int* __restrict__ seg_of,
                                                        const int* __restrict__ seg_off,
                                                        const double* __restrict__ mu, double* __restrict__ out) {
  __shared__ double red[3][256];
  const int t = blockIdx.x;
  if (t >= nt) return;
  const int i  = tdev[t];
  const int g  = seg_of[i];
  const int j0 = seg_off[g], j1 = seg_off[g + 1];
  const double xi = x[i], yi = y[i], zi = z[i];
  const double di = damping * s[i];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    if (excluded(i, j, eoff, eidx)) continue;
    if (damping > 0.0)
      pair_m1_f64<true>(xi - x[j], yi - y[j], zi - z[j], mu[j], mu[n + j], mu[2 * (size_t)n + j], di * s[j], fx, fy,
                        fz);
    else
      pair_m1_f64<false>(xi - x[j], yi - y[j], zi - z[j], mu[j], mu[n + j], mu[2 * (size_t)n + j], 0.0, fx, fy,
                         fz);
  }
  red[0][threadIdx.x] = fx;
  red[1][threadIdx.x] = fy;
  red[2][threadIdx.x] = fz;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * (size_t)t + threadIdx.x] = red[threadIdx.x][0];
}

void field1_targets(System& S, const double* mu_soa, int nt, const int* tdev, double* out) {
  hipLaunchKernelGGL(k_field1_targets, dim3(nt), dim3(256), 0, S.stream, nt, tdev, S.n, S.x.p, S.y.p, S.z.p, S.s.p,
                     S.opts.damping, S.excl_off.p, S.excl_idx.p, S.seg_of.p, S.seg_off.p, mu_soa, out);
}

// ---------------------------------------------------------------------------------------------
// Nuclear fields (electric_fields.cc:58-75) and the nuclear-multipole energy
// (multipole_expansion.cc:10-33), caller order, atoms staged in LDS.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nuclear(int n, const double* __restrict__ pos, int na,
                                                 const double* __restrict__ ch, const double* __restrict__ apos,
                                                 double* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int a = 0; a < na; ++a) {
    const double dx = px - apos[3 * a], dy = py - apos[3 * a + 1], dz = pz - apos[3 * a + 2];
    const double r  = sqrt(dx * dx + dy * dy + dz * dz);
    const double c  = ch[a] / (r * r * r);
    fx += c * dx;
    fy += c * dy;
    fz += c * dz;
  }
  F[3 * i]     = fx;
  F[3 * i + 1] = fy;
  F[3 * i + 2] = fz;
}

// E_i = sum_a Z_a sum_k (-1)^k/k! sum_(abc) k!/(a!b!c!) M_k(abc) d^(abc) (1/R), d = R_a - r_i
__global__ __launch_bounds__(256) void k_nuc_mul_energy(int n, const double* __restrict__ pos,
                                                        const int* __restrict__ kmax,
                                                        const double* __restrict__ mul, int na,
                                                        const double* __restrict__ ch,
                                                        const double* __restrict__ apos,
                                                        double* __restrict__ part) {
  __shared__ double red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e    = 0.0;
  if (i < n) {
    const int km = kmax[i];
    const double* M = &mul[10 * i];
    for (int a = 0; a < na; ++a) {
      const double dx = apos[3 * a] - pos[3 * i], dy = apos[3 * a + 1] - pos[3 * i + 1],
                   dz = apos[3 * a + 2] - pos[3 * i + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r  = sqrt(r2);
      const double r1 = 1.0 / r, r3 = r1 / r2, r5 = r3 / r2;
      double ea = 0.0;
      if (km >= 0) ea += M[0] * r1;
      if (km >= 1) ea += -(M[1] * (-dx * r3) + M[2] * (-dy * r3) + M[3] * (-dz * r3));
      if (km >= 2) {
        // T2 components xx,xy,xz,yy,yz,zz; symmetry factors 1,2,2,1,2,1; prefactor +1/2
        const double txx = 3 * dx * dx * r5 - r3, txy = 3 * dx * dy * r5, txz = 3 * dx * dz * r5;
        const double tyy = 3 * dy * dy * r5 - r3, tyz = 3 * dy * dz * r5, tzz = 3 * dz * dz * r5 - r3;
        ea += 0.5 * (M[4] * txx + 2 * M[5] * txy + 2 * M[6] * txz + M[7] * tyy + 2 * M[8] * tyz + M[9] * tzz);
      }
      e += ea * ch[a];
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Nuclear derivative of the nuclear fields (electric_fields.cc:77-104):
// G[(3a + p) * 3n + 3i + q] = Z_a T2(r_i - R_a)_pq,  T2_pq(d) = (3 d_p d_q - |d|^2 delta_pq) / |d|^5
__global__ __launch_bounds__(256) void k_nuc_field_grad(int n, const double* __restrict__ pos, int na,
                                                        const double* __restrict__ ch,
                                                        const double* __restrict__ apos, double* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t row = 3 * (size_t)n;
  for (int a = 0; a < na; ++a) {
    const double d[3] = {pos[3 * i] - apos[3 * a], pos[3 * i + 1] - apos[3 * a + 1], pos[3 * i + 2] - apos[3 * a + 2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r  = sqrt(r2);
    const double r3 = 1.0 / (r * r2), r5 = r3 / r2;
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q)
        G[(3 * (size_t)a + p) * row + 3 * (size_t)i + q] = ch[a] * (3.0 * d[p] * d[q] * r5 - (p == q ? r3 : 0.0));
  }
}

// Nuclear gradient of the nuclear-multipole energy (multipole_expansion.cc:35-62). With
// d = R_a - r_i, E_ai = Z_a sum_k (-1)^k/k! M_k . T_k(d), so dE/dR_a = Z_a sum_k (-1)^k/k! M_k . T_(k+1)(d):
//   k = 0: q T1 = -q d/r^3
//   k = 1: -(T2 mu) = mu/r^3 - 3 (d.mu) d/r^5
//   k = 2: 1/2 Theta : T3 = 1/2 (-15 (d.Theta.d) d/r^7 + 3 tr(Theta) d/r^5 + 6 Theta d/r^5)
// Grid (site blocks, atoms); per-block partials, summed in fixed order by k_grad_sum.
__global__ __launch_bounds__(256) void k_nuc_mul_grad(int n, const double* __restrict__ pos,
                                                      const int* __restrict__ kmax, const double* __restrict__ mul,
                                                      const double* __restrict__ ch,
                                                      const double* __restrict__ apos, double* __restrict__ part) {
  __shared__ double red[3][256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  double g[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const int km    = kmax[i];
    const double* M = &mul[10 * (size_t)i];
    const double d[3] = {apos[3 * a] - pos[3 * i], apos[3 * a + 1] - pos[3 * i + 1], apos[3 * a + 2] - pos[3 * i + 2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r  = sqrt(r2);
    const double r3 = 1.0 / (r * r2), r5 = r3 / r2, r7 = r5 / r2;
    if (km >= 0)
      for (int p = 0; p < 3; ++p) g[p] -= M[0] * d[p] * r3;
    if (km >= 1) {
      const double dm = d[0] * M[1] + d[1] * M[2] + d[2] * M[3];
      for (int p = 0; p < 3; ++p) g[p] += M[1 + p] * r3 - 3.0 * dm * d[p] * r5;
    }
    if (km >= 2) {
      const double* S = M + 4;  // xx xy xz yy yz zz
      const double td[3] = {S[0] * d[0] + S[1] * d[1] + S[2] * d[2], S[1] * d[0] + S[3] * d[1] + S[4] * d[2],
                            S[2] * d[0] + S[4] * d[1] + S[5] * d[2]};
      const double dtd = d[0] * td[0] + d[1] * td[1] + d[2] * td[2];
      const double tr  = S[0] + S[3] + S[5];
      for (int p = 0; p < 3; ++p) g[p] += 0.5 * (-15.0 * dtd * d[p] * r7 + 3.0 * tr * d[p] * r5 + 6.0 * td[p] * r5);
    }
    for (int p = 0; p < 3; ++p) g[p] *= ch[a];
  }
  for (int p = 0; p < 3; ++p) red[p][threadIdx.x] = g[p];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int p = 0; p < 3; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[((size_t)a * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// out[a * 3 + p] = sum_b part[(a * nb + b) * 3 + p], b in a fixed order; one block per atom.
__global__ __launch_bounds__(256) void k_grad_sum(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[3][256];
  const int a = blockIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int p = 0; p < 3; ++p) v[p] += part[((size_t)a * nb + b) * 3 + p];
  for (int p = 0; p < 3; ++p) red[p][threadIdx.x] = v[p];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int p = 0; p < 3; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * a + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_sum_partials(int np, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[256];
  double v = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) v += part[k];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ __launch_bounds__(256) void k_dot_part(long long n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double red[256];
  double v = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    v += a[k] * b[k];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Dense B (bmatrix.cc:64-92), caller order; one thread per (i, j) block of 3x3.
__global__ __launch_bounds__(256) void k_dense_b(int n, const double* __restrict__ pos,
                                                 const double* __restrict__ iso,
                                                 const double* __restrict__ ainv_aos, double damping,
                                                 const int* __restrict__ eoff, const int* __restrict__ eidx,
                                                 const int* __restrict__ index, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  const size_t N = 3 * (size_t)n;
  if (i == j) {
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) out[(3 * i + a) * N + 3 * j + c] = ainv_aos[9 * i + 3 * a + c];
    return;
  }
  bool ex = false;
  if (eoff)
    for (int k = eoff[i]; k < eoff[i + 1]; ++k) ex |= (eidx[k] == index[j]);
  double T[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (!ex) {
    const double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
    const double r    = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double fE = 1.0, fT = 1.0;
    if (damping > 0.0) {
      const double v = damping / pow(iso[i] * iso[j], 1.0 / 6.0);
      const double u = v * r;
      thole_ET(u, exp(-u), fE, fT);
    }
    const double r3 = 1.0 / (r * r * r), r5 = r3 / (r * r);
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) T[a][c] = -(3.0 * fT * d[a] * d[c] * r5 - (a == c ? fE * r3 : 0.0));
  }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) out[(3 * i + a) * N + 3 * j + c] = T[a][c];
}

}  // namespace cppe_dev
