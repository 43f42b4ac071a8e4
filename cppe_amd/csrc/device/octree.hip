// Morton-key octree and interaction lists on gfx950.
//
// The reference builds its adaptive octree by inserting sites one at a time and splitting a
// cell when it reaches ncrit sites (src/core/fmm/tree.cc:134-237). The resulting set of cells
// does not depend on insertion order: a cell exists iff it is a non-empty octant of a split
// cell, and a cell is split iff its region holds >= ncrit sites. This file builds that same
// set in parallel:
//   1. root centre = mean of the positions summed in index order, half-width = 1.001 * max
//      deviation (tree.cc:157-189) -- one block, the sum sequential to keep its rounding;
//   2. per site, the octant path of 21 levels, replaying the reference's FP64 cell-centre
//      arithmetic child = parent + (r/2) * (+-1) and the strict '>' octant tests
//      (tree.cc:99-102,119-120,196-198), packed 3 bits per level into a 63-bit Morton key;
//   3. stable radix sort of (key, index): leaf contents keep increasing site index, as in
//      the reference (its leaf lists are filled in insertion order);
//   4. leaf level of each site = first level whose prefix range holds < ncrit sites;
//   5. cells per level (level order, Morton order inside a level), centres re-derived from
//      the prefix path with the same arithmetic, rmax = sqrt(3 r r) (tree.cc:17);
//   6. dual-tree traversal (fmm/calculate.cc:74-127) level-synchronously over a frontier of
//      (target, source) cell pairs with the reference's MAC R*theta > rmaxA + rmaxB and split
//      rule; lists are then sorted by (target, source) into per-target CSR (owner computes).
// Cell identity (level, centre) and list membership are bit-identical to the reference; only
// the cell numbering differs (level order here, creation order there).
//
// Forests: a system may hold several independent snapshots ("segments", contiguous site ranges
// with their own ncrit). Each gets its own root (steps 1-2 per segment), the sort is segmented,
// cell ranges never cross a segment boundary, and the traversal starts from every (root, root)
// pair, so no list entry joins two snapshots. A segment whose ncrit exceeds its size is a single
// leaf whose near list is itself: its far field is empty and its near field is the exact direct
// sum. One segment with the system's ncrit is the reference's tree.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

#include "system.hh"

namespace cppe_dev {

constexpr int kMaxLevel = 21;

// ---------------------------------------------------------------------------------------------
// One block per snapshot (segment) of sites [off[g], off[g+1]).
__global__ __launch_bounds__(256) void k_root(const int* __restrict__ seg_off, const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z, double* __restrict__ out) {
  constexpr int CH = 2048;
  __shared__ double avg[3];
  __shared__ double red[3][256];
  __shared__ double buf[3][CH];
  const int t  = threadIdx.x;
  const int g  = blockIdx.x;
  const int i0 = seg_off[g], n = seg_off[g + 1] - i0;
  // sums in index order (one thread per axis) over LDS-staged chunks: the rounding of each
  // addition is that of the reference's sequential loop
  double s = 0.0;
  for (int base = 0; base < n; base += CH) {
    const int m = min(CH, n - base);
    __syncthreads();
    for (int k = t; k < m; k += blockDim.x) {
      buf[0][k] = x[i0 + base + k];
      buf[1][k] = y[i0 + base + k];
      buf[2][k] = z[i0 + base + k];
    }
    __syncthreads();
    if (t < 3)
      for (int k = 0; k < m; ++k) s = s + buf[t][k];
  }
  if (t < 3) avg[t] = s / (double)n;
  __syncthreads();
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int k = t; k < n; k += blockDim.x) {
    const int i = i0 + k;
    double dx = fabs(x[i] - avg[0]);
    double dy = fabs(y[i] - avg[1]);
    double dz = fabs(z[i] - avg[2]);
    if (dx > m0) m0 = dx;
    if (dy > m1) m1 = dy;
    if (dz > m2) m2 = dz;
  }
  red[0][t] = m0;
  red[1][t] = m1;
  red[2][t] = m2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int c = 0; c < 3; ++c) red[c][t] = fmax(red[c][t], red[c][t + s]);
    __syncthreads();
  }
  if (t == 0) {
    double xmax = red[0][0], ymax = red[1][0], zmax = red[2][0];
    double r = (xmax > ymax ? (xmax > zmax ? xmax : zmax) : (ymax > zmax ? ymax : zmax)) * 1.001;
    out[4 * g + 0] = avg[0];
    out[4 * g + 1] = avg[1];
    out[4 * g + 2] = avg[2];
    out[4 * g + 3] = r;
  }
}

// child centre update of add_child (tree.cc:99-102); r is the parent's half-width
__device__ __forceinline__ void child_centre(int oct, double& cx, double& cy, double& cz, double& r) {
#pragma clang fp contract(off)
  double rc = r / 2.0;
  cx = cx + rc * (double)((oct & 1) * 2 - 1);
  cy = cy + rc * (double)((oct & 2) - 1);
  cz = cz + rc * (double)((oct & 4) / 2 - 1);
  r  = rc;
}

__global__ __launch_bounds__(256) void k_keys(int n, const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z,
                                              const double* __restrict__ root, const int* __restrict__ seg_of,
                                              uint64_t* __restrict__ keys, int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* rt = root + 4 * seg_of[i];
  double cx = rt[0], cy = rt[1], cz = rt[2], r = rt[3];
  const double px = x[i], py = y[i], pz = z[i];
  uint64_t key = 0;
  for (int l = 0; l < kMaxLevel; ++l) {
    int oct = (px > cx) + ((py > cy) << 1) + ((pz > cz) << 2);
    key     = (key << 3) | (uint64_t)oct;
    child_centre(oct, cx, cy, cz, r);
  }
  keys[i] = key;
  idx[i]  = i;
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t* a, int lo, int hi, uint64_t v) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t prefix_of(uint64_t key, int L) {
  return L == 0 ? 0ull : (key >> (3 * (kMaxLevel - L)));
}

// range end (exclusive) of the level-L cell holding sorted position i, searched in [i, hi)
__device__ __forceinline__ int cell_end(const uint64_t* keys, int i, int hi, int L) {
  if (L == 0) return hi;
  uint64_t p    = prefix_of(keys[i], L);
  uint64_t next = (p + 1) << (3 * (kMaxLevel - L));
  if (L == 1 && p == 7) return hi;
  return lower_bound_u64(keys, i, hi, next);
}

// Leaf level of each sorted site: the first level whose cell holds < ncrit sites.
__global__ __launch_bounds__(256) void k_leaf_level(int n, const int* __restrict__ seg_of,
                                                    const int* __restrict__ seg_off,
                                                    const int* __restrict__ seg_ncrit,
                                                    const uint64_t* __restrict__ keys,
                                                    int* __restrict__ leaf_level, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g     = seg_of[i];
  const int ncrit = seg_ncrit[g];
  int lo = seg_off[g], hi = seg_off[g + 1];
  for (int L = 0; L <= kMaxLevel; ++L) {
    if (L > 0) {
      uint64_t p = prefix_of(keys[i], L);
      uint64_t a = p << (3 * (kMaxLevel - L));
      lo         = lower_bound_u64(keys, lo, hi, a);
      hi         = (L == 1 && p == 7) ? hi : lower_bound_u64(keys, lo, hi, (p + 1) << (3 * (kMaxLevel - L)));
    }
    if (hi - lo < ncrit) {
      leaf_level[i] = L;
      return;
    }
  }
  leaf_level[i] = kMaxLevel;
  atomicExch(err, 1);  // >= ncrit sites share one finest cell (coincident sites)
}

__global__ __launch_bounds__(256) void k_heads(int n, int L, const uint64_t* __restrict__ keys,
                                               const int* __restrict__ seg_of, const int* __restrict__ seg_off,
                                               const int* __restrict__ leaf_level, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int f = 0;
  if (leaf_level[i] >= L)
    f = (i == seg_off[seg_of[i]]) || (prefix_of(keys[i], L) != prefix_of(keys[i - 1], L));
  flag[i] = f;
}

__global__ __launch_bounds__(256) void k_fill_cells(
      int n, int L, int cell0, int prev0, int prev_n, const uint64_t* __restrict__ keys,
      const int* __restrict__ seg_of, const int* __restrict__ seg_off,
      const int* __restrict__ leaf_level, const int* __restrict__ flag, const int* __restrict__ rank,
      const double* __restrict__ root, double* cx, double* cy, double* cz, double* cr, double* crmax,
      int* start, int* count, int* parent, int* level) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int c    = cell0 + rank[i];
  const int g    = seg_of[i];
  uint64_t key   = keys[i];
  const double* rt = root + 4 * g;
  double x = rt[0], y = rt[1], z = rt[2], r = rt[3];
  for (int l = 1; l <= L; ++l) {
    int oct = (int)((key >> (3 * (kMaxLevel - l))) & 7ull);
    child_centre(oct, x, y, z, r);
  }
  cx[c]    = x;
  cy[c]    = y;
  cz[c]    = z;
  cr[c]    = r;
  crmax[c] = sqrt(3 * r * r);
  start[c] = i;
  count[c] = cell_end(keys, i, seg_off[g + 1], L) - i;
  level[c] = L;
  if (L == 0) {
    parent[c] = c;  // a root (one per snapshot) is its own parent
  } else {
    // parent = last cell of the previous level whose start <= i
    int lo = 0, hi = prev_n;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (start[prev0 + mid] <= i)
        lo = mid + 1;
      else
        hi = mid;
    }
    parent[c] = prev0 + lo - 1;
  }
}

__global__ __launch_bounds__(256) void k_link(int ncells, const int* __restrict__ parent,
                                              const int* __restrict__ level, const int* __restrict__ start,
                                              const double* __restrict__ cx, const double* __restrict__ cy,
                                              const double* __restrict__ cz, int* child_first, int* nchild,
                                              int* octmask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells || level[c] == 0) return;
  const int p = parent[c];
  atomicMin(&child_first[p], c);
  atomicAdd(&nchild[p], 1);
  int oct = (cx[c] > cx[p]) + ((cy[c] > cy[p]) << 1) + ((cz[c] > cz[p]) << 2);
  atomicOr(&octmask[p], 1 << oct);
}

__global__ void k_fill_int(int* a, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_gather_sorted(int n, const int* __restrict__ idx,
                                                       const double* __restrict__ xin,
                                                       const double* __restrict__ yin,
                                                       const double* __restrict__ zin, double* xo,
                                                       double* yo, double* zo, int* iperm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = idx[i];
  xo[i]       = xin[j];
  yo[i]       = yin[j];
  zo[i]       = zin[j];
  iperm[j]    = i;
}

__global__ void k_leaf_flags(int ncells, const int* nchild, int* flag) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) flag[c] = nchild[c] == 0;
}

__global__ void k_leaf_of(int nleaves, const int* __restrict__ leaf_ids, const int* __restrict__ start,
                          const int* __restrict__ count, int* __restrict__ leaf_of) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int c = leaf_ids[t];
  for (int i = start[c]; i < start[c] + count[c]; ++i) leaf_of[i] = c;
}

__global__ void k_scatter_ids(int n, const int* flag, const int* rank, int* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && flag[c]) out[rank[c]] = c;
}

template <typename T>
static void exclusive_scan(const T* in, T* out, int n, hipStream_t s, dbuf<char>& tmp) {
  size_t bytes = 0;
  CPPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
  tmp.alloc(bytes + 16);
  CPPE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, s));
}

// Builds the octree over the positions held in S.x/y/z (caller order on entry); on return
// S.x/y/z, S.perm and S.iperm are in device (sorted) order.
void build_octree(System& S) {
  const int n       = S.n;
  hipStream_t st    = S.stream;
  const int nseg    = S.nseg;
  Octree& T         = S.tree;
  dbuf<double> root;
  root.alloc(4 * (size_t)nseg);
  hipLaunchKernelGGL(k_root, dim3(nseg), dim3(256), 0, st, S.seg_off.p, S.x.p, S.y.p, S.z.p, root.p);
  dbuf<uint64_t> keys_in, keys;
  dbuf<int> idx_in, idx;
  keys_in.alloc(n);
  keys.alloc(n);
  idx_in.alloc(n);
  idx.alloc(n);
  hipLaunchKernelGGL(k_keys, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.x.p, S.y.p, S.z.p, root.p,
                     S.seg_of.p, keys_in.p, idx_in.p);
  dbuf<char> tmp;
  size_t bytes = 0;
  if (nseg == 1) {
    CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in.p, keys.p, idx_in.p, idx.p, n,
                                                      0, 3 * kMaxLevel, st));
    tmp.alloc(bytes + 16);
    CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys_in.p, keys.p, idx_in.p, idx.p, n, 0,
                                                      3 * kMaxLevel, st));
  } else {
    // snapshots stay where they are; each is sorted by key on its own (stable)
    CPPE_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, bytes, keys_in.p, keys.p, idx_in.p,
                                                               idx.p, n, nseg, S.seg_off.p, S.seg_off.p + 1, 0,
                                                               3 * kMaxLevel, st));
    tmp.alloc(bytes + 16);
    CPPE_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, bytes, keys_in.p, keys.p, idx_in.p, idx.p,
                                                               n, nseg, S.seg_off.p, S.seg_off.p + 1, 0,
                                                               3 * kMaxLevel, st));
  }
  // sorted positions, permutation
  dbuf<double> xs, ys, zs;
  xs.alloc(n);
  ys.alloc(n);
  zs.alloc(n);
  S.perm.alloc(n);
  S.iperm.alloc(n);
  hipLaunchKernelGGL(k_gather_sorted, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, idx.p, S.x.p, S.y.p,
                     S.z.p, xs.p, ys.p, zs.p, S.iperm.p);
  CPPE_HIP_CHECK(hipMemcpyAsync(S.perm.p, idx.p, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
  std::swap(S.x.p, xs.p);
  std::swap(S.y.p, ys.p);
  std::swap(S.z.p, zs.p);

  dbuf<int> leaf_level, err;
  leaf_level.alloc(n);
  err.alloc(1);
  err.zero(st);
  hipLaunchKernelGGL(k_leaf_level, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, S.seg_of.p, S.seg_off.p,
                     S.seg_ncrit.p, keys.p, leaf_level.p, err.p);
  int h_err = 0;
  CPPE_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  if (h_err) throw std::runtime_error("octree: more than ncrit sites share one cell of the finest level");

  // cells level by level; capacity grows as needed
  dbuf<int> flag, rank;
  flag.alloc(n);
  rank.alloc(n);
  std::vector<int> level_off{0};
  size_t cap = std::max(64, n / 4);
  auto ensure = [&](size_t need) {
    if (need <= T.cx.n && T.cx.p) return;
    size_t nc = std::max(need, (size_t)(T.cx.n * 2));
    auto grow_d = [&](dbuf<double>& b) {
      dbuf<double> nb;
      nb.alloc(nc);
      if (b.p) CPPE_HIP_CHECK(hipMemcpyAsync(nb.p, b.p, sizeof(double) * b.n, hipMemcpyDeviceToDevice, st));
      CPPE_HIP_CHECK(hipStreamSynchronize(st));
      std::swap(b.p, nb.p);
      std::swap(b.n, nb.n);
    };
    auto grow_i = [&](dbuf<int>& b) {
      dbuf<int> nb;
      nb.alloc(nc);
      if (b.p) CPPE_HIP_CHECK(hipMemcpyAsync(nb.p, b.p, sizeof(int) * b.n, hipMemcpyDeviceToDevice, st));
      CPPE_HIP_CHECK(hipStreamSynchronize(st));
      std::swap(b.p, nb.p);
      std::swap(b.n, nb.n);
    };
    grow_d(T.cx);
    grow_d(T.cy);
    grow_d(T.cz);
    grow_d(T.cr);
    grow_d(T.crmax);
    grow_i(T.start);
    grow_i(T.count);
    grow_i(T.parent);
    grow_i(T.level);
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
  };
  ensure(cap);
  int ncells = 0;
  int L      = 0;
  for (; L <= kMaxLevel; ++L) {
    hipLaunchKernelGGL(k_heads, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, L, keys.p, S.seg_of.p,
                       S.seg_off.p, leaf_level.p, flag.p);
    exclusive_scan(flag.p, rank.p, n, st, tmp);
    int last_rank = 0, last_flag = 0;
    CPPE_HIP_CHECK(hipMemcpyAsync(&last_rank, rank.p + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipMemcpyAsync(&last_flag, flag.p + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
    const int nl = last_rank + last_flag;
    if (nl == 0) break;
    ensure((size_t)ncells + nl);
    const int prev0 = L > 0 ? level_off[L - 1] : 0;
    const int prevn = L > 0 ? level_off[L] - level_off[L - 1] : 0;
    hipLaunchKernelGGL(k_fill_cells, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, L, ncells, prev0, prevn,
                       keys.p, S.seg_of.p, S.seg_off.p, leaf_level.p, flag.p, rank.p, root.p, T.cx.p, T.cy.p, T.cz.p, T.cr.p,
                       T.crmax.p, T.start.p, T.count.p, T.parent.p, T.level.p);
    ncells += nl;
    level_off.push_back(ncells);
  }
  T.ncells       = ncells;
  T.max_level    = (int)level_off.size() - 2;
  T.level_offset = level_off;
  T.child_first.alloc(ncells);
  T.nchild.alloc(ncells);
  T.octmask.alloc(ncells);
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(ncells, 256)), dim3(256), 0, st, T.child_first.p, ncells,
                     INT_MAX);
  T.nchild.zero(st);
  T.octmask.zero(st);
  hipLaunchKernelGGL(k_link, dim3(ceil_div(ncells, 256)), dim3(256), 0, st, ncells, T.parent.p, T.level.p,
                     T.start.p, T.cx.p, T.cy.p, T.cz.p, T.child_first.p, T.nchild.p, T.octmask.p);
  // leaf ids
  dbuf<int> lf, lr;
  lf.alloc(ncells);
  lr.alloc(ncells);
  hipLaunchKernelGGL(k_leaf_flags, dim3(ceil_div(ncells, 256)), dim3(256), 0, st, ncells, T.nchild.p, lf.p);
  exclusive_scan(lf.p, lr.p, ncells, st, tmp);
  int lrank = 0, lflag = 0;
  CPPE_HIP_CHECK(hipMemcpyAsync(&lrank, lr.p + ncells - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(&lflag, lf.p + ncells - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  T.nleaves = lrank + lflag;
  T.leaf_ids.alloc(T.nleaves);
  hipLaunchKernelGGL(k_scatter_ids, dim3(ceil_div(ncells, 256)), dim3(256), 0, st, ncells, lf.p, lr.p,
                     T.leaf_ids.p);
  T.leaf_of.alloc(n);
  hipLaunchKernelGGL(k_leaf_of, dim3(ceil_div(T.nleaves, 256)), dim3(256), 0, st, T.nleaves, T.leaf_ids.p,
                     T.start.p, T.count.p, T.leaf_of.p);
  S.h_perm.resize(n);
  CPPE_HIP_CHECK(hipMemcpyAsync(S.h_perm.data(), S.perm.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  T.built = true;
}

// ---------------------------------------------------------------------------------------------
// Dual-tree traversal (fmm/calculate.cc:74-127), one generation of the frontier per launch.
// ---------------------------------------------------------------------------------------------
struct TravOut {
  int2* next;
  unsigned long long* m2l;  // (target << 32) | source
  unsigned long long* p2p;  // (target << 32) | source
  int* counters;            // [0] next, [1] m2l, [2] p2p, [3] overflow
  int cap_next, cap_m2l, cap_p2p;
};

__global__ __launch_bounds__(256) void k_traverse(int nfront, const int2* __restrict__ front, TravOut o,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const double* __restrict__ cz,
                                                  const double* __restrict__ rmax,
                                                  const int* __restrict__ nchild,
                                                  const int* __restrict__ child_first, double theta) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfront) return;
  const int A = front[t].x, B = front[t].y;
  const double dx = cx[A] - cx[B];
  const double dy = cy[A] - cy[B];
  const double dz = cz[A] - cz[B];
  const double R  = sqrt(dx * dx + dy * dy + dz * dz);
  const bool leafA = nchild[A] == 0, leafB = nchild[B] == 0;
  if (R * theta > (rmax[A] + rmax[B])) {
    int slot = atomicAdd(&o.counters[1], 1);
    if (slot < o.cap_m2l)
      o.m2l[slot] = ((unsigned long long)A << 32) | (unsigned)B;
    else
      atomicExch(&o.counters[3], 1);
  } else if (leafA && leafB) {
    int slot = atomicAdd(&o.counters[2], 1);
    if (slot < o.cap_p2p)
      o.p2p[slot] = ((unsigned long long)A << 32) | (unsigned)B;
    else
      atomicExch(&o.counters[3], 1);
  } else if (leafB || (rmax[A] >= rmax[B] && !leafA)) {
    const int k = nchild[A], c0 = child_first[A];
    int base    = atomicAdd(&o.counters[0], k);
    if (base + k <= o.cap_next) {
      for (int c = 0; c < k; ++c) o.next[base + c] = make_int2(c0 + c, B);
    } else {
      atomicExch(&o.counters[3], 1);
    }
  } else {
    const int k = nchild[B], c0 = child_first[B];
    int base    = atomicAdd(&o.counters[0], k);
    if (base + k <= o.cap_next) {
      for (int c = 0; c < k; ++c) o.next[base + c] = make_int2(A, c0 + c);
    } else {
      atomicExch(&o.counters[3], 1);
    }
  }
}

// CSR offsets of target-sorted pair keys: off[c] = first entry with target >= c
__global__ void k_csr_offsets(int ncells, long long npairs, const unsigned long long* __restrict__ keys,
                              int* __restrict__ off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  unsigned long long v = (unsigned long long)c << 32;
  long long lo = 0, hi = npairs;
  while (lo < hi) {
    long long mid = (lo + hi) >> 1;
    if (keys[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  off[c] = (int)lo;
}

__global__ void k_low_words(long long n, const unsigned long long* __restrict__ keys, int* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)(keys[i] & 0xffffffffull);
}

__global__ void k_pair_sites(long long n, const unsigned long long* __restrict__ keys,
                             const int* __restrict__ count, unsigned long long* __restrict__ acc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (i < n) {
    int A = (int)(keys[i] >> 32), B = (int)(keys[i] & 0xffffffffull);
    v     = (unsigned long long)count[A] * (unsigned long long)count[B];
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(acc, v);
}

static void sort_keys(dbuf<unsigned long long>& keys, long long n, int ncells, hipStream_t st,
                      dbuf<char>& tmp) {
  if (n == 0) return;
  dbuf<unsigned long long> out;
  out.alloc(n);
  int bits = 32;
  while ((1ll << (bits - 32)) <= ncells && bits < 64) ++bits;
  size_t bytes = 0;
  CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.p, out.p, (int)n, 0, bits, st));
  tmp.alloc(bytes + 16);
  CPPE_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.p, out.p, (int)n, 0, bits, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  std::swap(keys.p, out.p);
  std::swap(keys.n, out.n);
}

// Expanded near lists (one entry per near source site of each target leaf) for the slotted
// near-field kernel.
__global__ void k_xcount(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                         const int* __restrict__ count, long long* __restrict__ xcnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  long long t = 0;
  for (int k = off[c]; k < off[c + 1]; ++k) t += count[src[k]];
  xcnt[c] = t;
}

// Class of every entry of a leaf's expanded near list (near-list order), one wavefront per cell,
// lanes = sources: the smallest Thole argument v = damping s_i s_j R_ij over the leaf's own sites i
// (exact, FP64), then 0: v < 18, 1: 18 <= v < 40, 2: v >= 40 (a small margin above each bound).
// Undamped sites (s = 0) and the leaf's own sites (the site itself is among the targets, R = 0)
// give class 0 without the search, so one-leaf (direct-summation) systems cost O(n) here.
__global__ __launch_bounds__(64) void k_xclass(const int* __restrict__ off, const int* __restrict__ src,
                                               const int* __restrict__ start, const int* __restrict__ count,
                                               const double* __restrict__ x, const double* __restrict__ y,
                                               const double* __restrict__ z, const double* __restrict__ s,
                                               double damping, const long long* __restrict__ xoff,
                                               unsigned char* __restrict__ xcls) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const int i0 = start[c], ni = count[c];
  long long q = xoff[c];
  for (int k = off[c]; k < off[c + 1]; ++k) {
    const int B = src[k], j0 = start[B], nb = count[B];
    for (int u0 = 0; u0 < nb; u0 += 64) {
      const int u = u0 + lane;
      if (u < nb && (B == c || s[j0 + u] == 0.0)) {
        xcls[q + u] = 0;
      } else if (u < nb) {
        const int j = j0 + u;
        const double xj = x[j], yj = y[j], zj = z[j];
        double m = 1e300;
        for (int i = i0; i < i0 + ni; ++i) {  // uniform index: the leaf's sites are read by every lane
          const double dx = xj - x[i], dy = yj - y[i], dz = zj - z[i];
          m = fmin(m, (s[i] * s[i]) * (dx * dx + dy * dy + dz * dz));
        }
        const double v = damping * s[j] * sqrt(m);
        xcls[q + u] = v < kFarClass32 * (1.0 + 1e-9) ? 0 : (v < kTholeCutoff * (1.0 + 1e-9) ? 1 : 2);
      }
    }
    q += nb;
  }
}

// One wavefront per cell: the cell's near sources written class by class (k_xclass), each class
// in the order of the near list (stable compaction with a ballot per 64 sources); xcls = null puts
// every source in class 0.
__global__ __launch_bounds__(64) void k_xfill(const int* __restrict__ off, const int* __restrict__ src,
                                              const int* __restrict__ start, const int* __restrict__ count,
                                              const unsigned char* __restrict__ xcls,
                                              const long long* __restrict__ xoff, int* __restrict__ xsrc,
                                              long long* __restrict__ xs64, long long* __restrict__ xs32) {
  const int c = blockIdx.x, lane = threadIdx.x;
  long long pos = xoff[c];
  for (int pass = 0; pass < 3; ++pass) {
    long long q = xoff[c];
    for (int k = off[c]; k < off[c + 1]; ++k) {
      const int B = src[k], j0 = start[B], nb = count[B];
      for (int u0 = 0; u0 < nb; u0 += 64) {
        const int u  = u0 + lane;
        const int cl = u < nb ? (xcls ? (int)xcls[q + u] : 0) : -1;
        const unsigned long long m = __ballot(cl == pass);
        if (cl == pass) xsrc[pos + __popcll(m & ((1ull << lane) - 1ull))] = j0 + u;
        pos += __popcll(m);
      }
      q += nb;
    }
    if (lane == 0 && pass == 0) xs32[c] = pos;
    if (lane == 0 && pass == 1) xs64[c] = pos;
  }
}

void build_near_expanded(System& S) {
  Octree& T = S.tree;
  if (T.nx >= 0) return;
  hipStream_t st = S.stream;
  const int nc   = T.ncells;
  dbuf<long long> cnt;
  cnt.alloc(nc + 1);
  hipLaunchKernelGGL(k_xcount, dim3(ceil_div(nc, 256)), dim3(256), 0, st, nc, T.p2p_off.p, T.p2p_src.p, T.count.p,
                     cnt.p);
  long long zero = 0;
  CPPE_HIP_CHECK(hipMemcpyAsync(cnt.p + nc, &zero, sizeof(long long), hipMemcpyHostToDevice, st));
  T.xoff.alloc(nc + 1);
  dbuf<char> tmp;
  exclusive_scan(cnt.p, T.xoff.p, nc + 1, st, tmp);
  long long total = 0;
  CPPE_HIP_CHECK(hipMemcpyAsync(&total, T.xoff.p + nc, sizeof(long long), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  T.xsrc.alloc(std::max<long long>(total, 1));
  T.xsplit[0].alloc(nc);
  T.xsplit[1].alloc(nc);
  dbuf<unsigned char> cls;
  const bool damped = S.s.p && S.opts.damping > 0.0;
  if (damped) {
    cls.alloc(std::max<long long>(total, 1));
    hipLaunchKernelGGL(k_xclass, dim3(nc), dim3(64), 0, st, T.p2p_off.p, T.p2p_src.p, T.start.p, T.count.p, S.x.p,
                       S.y.p, S.z.p, S.s.p, S.opts.damping, T.xoff.p, cls.p);
  }
  hipLaunchKernelGGL(k_xfill, dim3(nc), dim3(64), 0, st, T.p2p_off.p, T.p2p_src.p, T.start.p, T.count.p,
                     damped ? cls.p : nullptr, T.xoff.p, T.xsrc.p, T.xsplit[0].p, T.xsplit[1].p);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  T.nx    = total;
  T.xdamp = S.opts.damping;
}

// ---- lists of the symmetric near field (near_sym.hh) -------------------------------------------
// Per leaf B (thread per cell): over its near list (sorted by source leaf), the number of chunks
// of the leaves A < B before each entry (block index of A's first chunk among B's incoming
// reverse blocks) and in total.
__global__ void k_sym_pre(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                          const int* __restrict__ nq, int* __restrict__ pre, int* __restrict__ nin) {
  const int B = blockIdx.x * blockDim.x + threadIdx.x;
  if (B >= ncells) return;
  int run = 0;
  for (int k = off[B]; k < off[B + 1]; ++k) {
    const int A = src[k];
    pre[k]      = run;
    if (A < B) run += nq[A];
  }
  nin[B] = run;
}

// Per leaf A: upper-list length (own sites + sites of near leaves B > A) and reverse slots it
// receives (incoming blocks x 3 |A|).
__global__ void k_sym_count(int ncells, const int* __restrict__ off, const int* __restrict__ src,
                            const int* __restrict__ count, const int* __restrict__ nin,
                            long long* __restrict__ ucnt, long long* __restrict__ rsz) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A > ncells) return;
  if (A == ncells || off[A + 1] == off[A]) {  // the scans' last entry; cells without a near list
    ucnt[A] = 0;
    rsz[A]  = 0;
    return;
  }
  long long c = count[A];
  for (int k = off[A]; k < off[A + 1]; ++k)
    if (src[k] > A) c += count[src[k]];
  ucnt[A] = c;
  rsz[A]  = (long long)nin[A] * 3 * count[A];
}

// One wavefront per leaf A: its own sites, then the entries of its expanded near list whose leaf
// is above A (order kept, compaction by ballot), each with its reverse slot; usplit: position of
// the first kept entry at or after the FP32 far split of the expanded list.
__global__ __launch_bounds__(64) void k_sym_fill(const long long* __restrict__ xoff, const int* __restrict__ xsrc,
                                                 const long long* __restrict__ xs32, const int* __restrict__ leaf_of,
                                                 const int* __restrict__ start, const int* __restrict__ count,
                                                 const int* __restrict__ p2p_off, const int* __restrict__ p2p_src,
                                                 const int* __restrict__ pre, const long long* __restrict__ rev_off,
                                                 const long long* __restrict__ uoff, int* __restrict__ usrc,
                                                 int2* __restrict__ udst, long long* __restrict__ usplit,
                                                 int* __restrict__ bad) {
  const int A = blockIdx.x, lane = threadIdx.x;
  const long long u0 = uoff[A], u1 = uoff[A + 1];
  if (u1 == u0) {
    if (lane == 0) usplit[A] = u1;
    return;
  }
  const int na = count[A];
  for (int u = lane; u < na; u += 64) {
    usrc[u0 + u] = start[A] + u;
    udst[u0 + u] = make_int2(0, 0);
  }
  long long pos   = u0 + na;
  const long long x0 = xoff[A], x1 = xoff[A + 1];
  const long long split = xs32 ? xs32[A] : x1;
  long long usp   = -1;
  for (long long e0 = x0; e0 < x1; e0 += 64) {
    const long long e = e0 + lane;
    const bool valid  = e < x1;
    const int j       = valid ? xsrc[e] : 0;
    const int B       = valid ? leaf_of[j] : -1;
    const bool keep   = valid && B > A;
    const unsigned long long m = __ballot(keep);
    if (usp < 0 && split < e0 + 64) usp = pos + __popcll(__ballot(keep && e < split));
    if (keep) {
      const long long idx = pos + __popcll(m & ((1ull << lane) - 1ull));
      int lo = p2p_off[B], hi = p2p_off[B + 1];  // A in B's near list (sorted by source leaf)
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (p2p_src[mid] < A)
          lo = mid + 1;
        else
          hi = mid;
      }
      usrc[idx] = j;
      if (lo < p2p_off[B + 1] && p2p_src[lo] == A) {
        const long long slot = rev_off[B] + (long long)pre[lo] * 3 * count[B] + (j - start[B]);
        udst[idx]            = make_int2((int)slot, count[B]);
      } else {  // B near A but not A near B: the lists are not symmetric
        udst[idx] = make_int2(0, 0);
        *bad      = 1;
      }
    }
    pos += __popcll(m);
  }
  if (usp < 0) usp = pos;
  if (lane == 0) usplit[A] = usp;
}

SymLists& build_sym_lists(System& S, int ts) {
  Octree& T    = S.tree;
  SymLists& Y  = T.sym[sym_slot(ts)];
  if (Y.ts == ts) return Y;
  build_near_expanded(S);
  hipStream_t st = S.stream;
  const int nc   = T.ncells;
  std::vector<int> ids(T.nleaves), cnt(nc), nq(nc, 0);
  CPPE_HIP_CHECK(hipMemcpyAsync(ids.data(), T.leaf_ids.p, sizeof(int) * T.nleaves, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(cnt.data(), T.count.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<int4> items;
  items.reserve(T.nleaves + T.nleaves / 4);
  for (int c : ids) {
    nq[c] = (cnt[c] + ts - 1) / ts;
    for (int q = 0; q < nq[c]; ++q) items.push_back(make_int4(c, q * ts, std::min(ts, cnt[c] - q * ts), q));
  }
  dbuf<int> d_nq, pre;
  d_nq.alloc(nc);
  pre.alloc(std::max<long long>(T.n_p2p, 1));
  Y.nin.alloc(nc);
  CPPE_HIP_CHECK(hipMemcpyAsync(d_nq.p, nq.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sym_pre, dim3(ceil_div(nc, 256)), dim3(256), 0, st, nc, T.p2p_off.p, T.p2p_src.p, d_nq.p,
                     pre.p, Y.nin.p);
  dbuf<long long> ucnt, rsz;
  ucnt.alloc(nc + 1);
  rsz.alloc(nc + 1);
  hipLaunchKernelGGL(k_sym_count, dim3(ceil_div(nc + 1, 256)), dim3(256), 0, st, nc, T.p2p_off.p, T.p2p_src.p,
                     T.count.p, Y.nin.p, ucnt.p, rsz.p);
  Y.uoff.alloc(nc + 1);
  Y.rev_off.alloc(nc + 1);
  dbuf<char> tmp;
  exclusive_scan(ucnt.p, Y.uoff.p, nc + 1, st, tmp);
  exclusive_scan(rsz.p, Y.rev_off.p, nc + 1, st, tmp);
  long long tot[2] = {0, 0};
  CPPE_HIP_CHECK(hipMemcpyAsync(&tot[0], Y.uoff.p + nc, sizeof(long long), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(&tot[1], Y.rev_off.p + nc, sizeof(long long), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  if (tot[1] >= (1ll << 31) - 1)
    throw hip_error("symmetric near field: more reverse slots than a 32-bit index addresses");
  Y.usrc.alloc(std::max<long long>(tot[0], 1));
  Y.udst.alloc(std::max<long long>(tot[0], 1));
  Y.usplit.alloc(nc);
  const long long* xs = S.opts.damping >= T.xdamp ? T.xsplit[1].p : nullptr;
  dbuf<int> bad;
  bad.alloc(1);
  int h_bad = 0;
  CPPE_HIP_CHECK(hipMemcpyAsync(bad.p, &h_bad, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sym_fill, dim3(nc), dim3(64), 0, st, T.xoff.p, T.xsrc.p, xs, T.leaf_of.p, T.start.p,
                     T.count.p, T.p2p_off.p, T.p2p_src.p, pre.p, Y.rev_off.p, Y.uoff.p, Y.usrc.p, Y.udst.p,
                     Y.usplit.p, bad.p);
  Y.items.alloc(std::max<size_t>(items.size(), 1));
  CPPE_HIP_CHECK(hipMemcpyAsync(Y.items.p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  if (h_bad) throw hip_error("symmetric near field: the near lists of this tree are not symmetric");
  Y.nitems = (int)items.size();
  Y.nrev   = tot[1];
  Y.ts     = ts;
  return Y;
}

void build_lists(System& S) {
  Octree& T      = S.tree;
  hipStream_t st = S.stream;
  const double theta = S.opts.theta;
  long long cap_m2l  = 512ll * T.ncells + 4096;
  long long cap_p2p  = 256ll * T.nleaves + 4096;
  long long cap_next = std::max<long long>(cap_m2l, T.level_offset[1]);
  dbuf<int> counters;
  counters.alloc(4);
  for (int attempt = 0; attempt < 6; ++attempt) {
    dbuf<int2> fa, fb;
    dbuf<unsigned long long> m2l, p2p;
    fa.alloc(cap_next);
    fb.alloc(cap_next);
    m2l.alloc(cap_m2l);
    p2p.alloc(cap_p2p);
    // one traversal per snapshot, all started together from the (root, root) pairs
    const int nroots = T.level_offset[1];
    std::vector<int2> seed(nroots);
    for (int c = 0; c < nroots; ++c) seed[c] = make_int2(c, c);
    CPPE_HIP_CHECK(hipMemcpyAsync(fa.p, seed.data(), sizeof(int2) * nroots, hipMemcpyHostToDevice, st));
    int h_cnt[4] = {0, 0, 0, 0};
    CPPE_HIP_CHECK(hipMemcpyAsync(counters.p, h_cnt, sizeof(h_cnt), hipMemcpyHostToDevice, st));
    int nfront = nroots;
    bool overflow = false;
    while (nfront > 0) {
      TravOut o{fb.p, m2l.p, p2p.p, counters.p, (int)cap_next, (int)cap_m2l, (int)cap_p2p};
      hipLaunchKernelGGL(k_traverse, dim3(ceil_div(nfront, 256)), dim3(256), 0, st, nfront, fa.p, o, T.cx.p,
                         T.cy.p, T.cz.p, T.crmax.p, T.nchild.p, T.child_first.p, theta);
      CPPE_HIP_CHECK(hipMemcpyAsync(h_cnt, counters.p, sizeof(h_cnt), hipMemcpyDeviceToHost, st));
      CPPE_HIP_CHECK(hipStreamSynchronize(st));
      if (h_cnt[3]) {
        overflow = true;
        break;
      }
      nfront = h_cnt[0];
      int zero = 0;
      CPPE_HIP_CHECK(hipMemcpyAsync(counters.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
      std::swap(fa.p, fb.p);
    }
    if (overflow) {
      cap_m2l *= 2;
      cap_p2p *= 2;
      cap_next *= 2;
      continue;
    }
    T.n_m2l = h_cnt[1];
    T.n_p2p = h_cnt[2];
    dbuf<char> tmp;
    sort_keys(m2l, T.n_m2l, T.ncells, st, tmp);
    sort_keys(p2p, T.n_p2p, T.ncells, st, tmp);
    T.m2l_off.alloc(T.ncells + 1);
    T.p2p_off.alloc(T.ncells + 1);
    T.m2l_src.alloc(std::max<long long>(T.n_m2l, 1));
    T.p2p_src.alloc(std::max<long long>(T.n_p2p, 1));
    hipLaunchKernelGGL(k_csr_offsets, dim3(ceil_div(T.ncells + 1, 256)), dim3(256), 0, st, T.ncells, T.n_m2l,
                       m2l.p, T.m2l_off.p);
    hipLaunchKernelGGL(k_csr_offsets, dim3(ceil_div(T.ncells + 1, 256)), dim3(256), 0, st, T.ncells, T.n_p2p,
                       p2p.p, T.p2p_off.p);
    if (T.n_m2l)
      hipLaunchKernelGGL(k_low_words, dim3(ceil_div(T.n_m2l, 256)), dim3(256), 0, st, T.n_m2l, m2l.p,
                         T.m2l_src.p);
    if (T.n_p2p)
      hipLaunchKernelGGL(k_low_words, dim3(ceil_div(T.n_p2p, 256)), dim3(256), 0, st, T.n_p2p, p2p.p,
                         T.p2p_src.p);
    dbuf<unsigned long long> acc;
    acc.alloc(1);
    acc.zero(st);
    if (T.n_p2p)
      hipLaunchKernelGGL(k_pair_sites, dim3(ceil_div(T.n_p2p, 256)), dim3(256), 0, st, T.n_p2p, p2p.p,
                         T.count.p, acc.p);
    unsigned long long h_acc = 0;
    CPPE_HIP_CHECK(hipMemcpyAsync(&h_acc, acc.p, sizeof(h_acc), hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
    T.p2p_site_pairs = (long long)h_acc;
    return;
  }
  throw std::runtime_error("octree: interaction lists exceed the traversal capacity");
}

}  // namespace cppe_dev
