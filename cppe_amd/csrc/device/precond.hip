// Local Neumann preconditioner of the device CG (no counterpart in the reference, whose
// InducedMoments::compute preconditions with alpha r, electric_fields.cc:191-239).
//
// B = alpha^-1 - T, so B^-1 = alpha + alpha T alpha + ...; the preconditioner keeps the first
// correction over close pairs only:
//   z = w + omega alpha (T_loc w),  w = alpha r,
// with T_loc_ij = T2d(r_i - r_j) (the Thole-damped dipole tensor of the system, i.e. -B_ij) for
// pairs closer than a radius (3 A by default: about six neighbours of a water site), excluded
// pairs and the site itself left out. T_loc is symmetric, so z = M r with a fixed symmetric M and
// CG keeps its recurrence; convergence is still tested on the unpreconditioned ||r||_2.
//
// Pairs are taken from the octree's near lists (every pair of sites in P2P partner leaves closer
// than the radius; partner lists are symmetric, so T_loc is). Only the partner index of each pair
// is stored: the solver's k_sg_prec forms each block where it applies it (a handful of pairs per
// site, so the arithmetic costs far less than reading stored blocks), from the site records
// (x, y, z, s) kept here -- see pc_pair in solver.hip for the arithmetic and its symmetry.
#include <algorithm>

#include "pairs.hh"
#include "system.hh"

namespace cppe_dev {

struct PcArgs {
  int n;
  const int* leaf_of;
  const int* p2p_off;
  const int* p2p_src;
  const int* start;
  const int* count;
  const double *x, *y, *z, *s;
  const int* eoff;
  const int* eidx;
  double rc2, damping;
};

template <typename Fn>
__device__ __forceinline__ void pc_for_pairs(const PcArgs& a, int i, Fn&& f) {
  const int A     = a.leaf_of[i];
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
  const int e0 = a.eoff ? a.eoff[i] : 0, e1 = a.eoff ? a.eoff[i + 1] : 0;
  for (int q = a.p2p_off[A]; q < a.p2p_off[A + 1]; ++q) {
    const int B  = a.p2p_src[q];
    const int j0 = a.start[B], j1 = j0 + a.count[B];
    for (int j = j0; j < j1; ++j) {
      if (j == i) continue;
      const double dx = xi - a.x[j], dy = yi - a.y[j], dz = zi - a.z[j];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= a.rc2) continue;
      bool ex = false;
      for (int k = e0; k < e1; ++k) ex |= a.eidx[k] == j;
      if (!ex) f(j, dx, dy, dz, r2);
    }
  }
}

__global__ __launch_bounds__(256) void k_pc_count(PcArgs a, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int c = 0;
  pc_for_pairs(a, i, [&](int, double, double, double, double) { ++c; });
  cnt[i] = c;
}

// Partner index of every pair of site i, in the order of its near list; and the site record.
__global__ __launch_bounds__(256) void k_pc_fill(PcArgs a, const int* __restrict__ off, int* __restrict__ nj,
                                                 double4* __restrict__ site) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int p = off[i];
  pc_for_pairs(a, i, [&](int j, double, double, double, double) { nj[p++] = j; });
  site[i] = make_double4(a.x[i], a.y[i], a.z[i], a.s[i]);
}

void precond_destroy(System& S) {
  delete S.pc;
  S.pc = nullptr;
}

// The pair list of the system's local preconditioner (built on first use; positions and the
// radius fix it).
void precond_build(System& S) {
  if (S.pc && S.pc->rc == S.precond_rc && S.pc->damping == S.opts.damping) return;
  precond_destroy(S);
  auto* P    = new Precond();
  S.pc       = P;
  P->rc      = S.precond_rc;
  P->damping = S.opts.damping;
  const Octree& T = S.tree;
  PcArgs a{S.n, T.leaf_of.p, T.p2p_off.p, T.p2p_src.p, T.start.p, T.count.p, S.x.p, S.y.p, S.z.p, S.s.p,
           S.excl_off.p, S.excl_idx.p, S.precond_rc * S.precond_rc, S.opts.damping};
  hipStream_t st = S.stream;
  P->off.alloc((size_t)S.n + 1);
  hipLaunchKernelGGL(k_pc_count, dim3(ceil_div(S.n, 256)), dim3(256), 0, st, a, P->off.p);
  std::vector<int> cnt(S.n + 1, 0);
  CPPE_HIP_CHECK(hipMemcpyAsync(cnt.data(), P->off.p, sizeof(int) * S.n, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  long long tot = 0;
  for (int i = 0; i < S.n; ++i) {
    const long long c = cnt[i];
    cnt[i]            = (int)tot;
    tot += c;
  }
  if (tot > 0x7fffffffLL) throw hip_error("local preconditioner: too many pairs within the radius");
  cnt[S.n]    = (int)tot;
  P->npairs   = tot;
  CPPE_HIP_CHECK(hipMemcpyAsync(P->off.p, cnt.data(), sizeof(int) * (S.n + 1), hipMemcpyHostToDevice, st));
  P->nj.alloc((size_t)std::max(1LL, tot));
  P->site.alloc((size_t)S.n);
  hipLaunchKernelGGL(k_pc_fill, dim3(ceil_div(S.n, 256)), dim3(256), 0, st, a, P->off.p, P->nj.p, P->site.p);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace cppe_dev
