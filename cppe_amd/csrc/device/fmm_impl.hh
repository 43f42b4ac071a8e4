// Fast multipole method on gfx950: kernels and pass drivers, templated on the expansion order.
//
// Passes (the reference's Tree::compute_field_fmm, src/core/fmm/tree.cc:359-385, with its
// shift conventions, fmm/calculate.cc:129-260), all owner-computes (no atomics, deterministic):
//   P2M   one thread per leaf: point multipoles of its sites shifted to the leaf centre
//   M2M   one thread per internal cell, level by level from the deepest level to the root
//   M2L   one wavefront per target cell; lanes take sources of its target-grouped M2L list,
//         each lane evaluates the Coulomb derivatives and contracts them with the source M;
//         the 64 partial expansions are summed by wavefront shuffles
//   L2L   one thread per cell, level by level from the root
//   P2P+L2P one wavefront per target leaf: lanes split as (target site t, source slot s) with
//         s < 64/T; the leaf's near list is walked 64 source leaves at a time (wavefront prefix
//         scan of leaf sizes) and its sites are staged through LDS 64 at a time; the slot
//         partials are summed through LDS, then each target adds its L2P term (and, for the
//         response operator, alpha^-1 mu)
// Damping (Thole) applies to near-field pairs only; the far field is undamped, as in the
// reference (fmm/calculate.cc:48-72). Exclusions apply only to near-field pairs (ibid.).
// FP32 form: M, L and pair arithmetic in FP32 with coordinates taken relative to cell / leaf
// centres in FP64 before rounding, so FP32 sees only local offsets.
#pragma once
#include <type_traits>

#include "fmm_ops.gen.hh"
#include "pairs.hh"
#include "system.hh"

namespace cppe_dev {

// Multipole buffers hold, per right-hand side, the full rows (ncells x msize) followed, at a 16-byte
// aligned offset, by the folded rows M2L reads (ncells x fstride: fsize padded to 4 values).
constexpr int fmm_fstride(int kind_mc, int p) { return (fmm_fsize(kind_mc, p) + 3) & ~3; }
inline size_t fmm_fold_offset(size_t ncells, int kind_mc, int p) {
  return (ncells * fmm_msize(kind_mc, p) + 3) & ~(size_t)3;
}
inline size_t fmm_mbuf_size(size_t ncells, int kind_mc, int p) {
  return fmm_fold_offset(ncells, kind_mc, p) + ncells * fmm_fstride(kind_mc, p);
}

enum FmmKind { KIND_M1 = 0, KIND_MC = 1 };
enum P2PMode { MODE_BAPPLY = 0, MODE_FIELD1 = 1, MODE_STATIC = 2 };

// Launch shapes of the B-apply kernels, chosen by the autotuner (cppe_hip_autotune); index 0 is
// the FP64 apply, 1 the FP32 apply.
struct FmmTune {
  int near_variant[2] = {0, 0};  // 0: one thread per target site, 1: one wavefront per leaf chunk
  int near_block[2]   = {256, 256};
  int m2l_group[2]    = {64, 64};  // lanes per target cell in M2L (16, 32 or 64)
  int overlap[2]      = {0, 0};    // far field on a second stream beside the near field (staged variants):
                                   // 0 none, 1 the whole far field, 2 M2L / L2L after the upward pass,
                                   // 3 the whole far field enqueued ahead of the near field
  int near_splits[2]  = {1, 1};    // overlapped forms: the near field in 1, 2 or 4 launches, so the
                                   // far field's kernels get the device between them
};

struct FmmState {
  int order = 0;
  FmmTune tune;
  // second stream for the far field (default priority) and the events that fork / join it
  hipStream_t far_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~FmmState() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (far_stream) (void)hipStreamDestroy(far_stream);
  }
  // work items (leaf, first target) of the leaf-chunk near kernels, for chunks of 64, 32, 16, 8 targets
  dbuf<int2> near_items[4];
  int n_items[4] = {0, 0, 0, 0};
  // target clusters: runs of whole leaves (device order) of <= 64 sites, and per cluster the
  // union of its leaves' near lists with, per source leaf, the mask of lanes it is near for
  int n_clusters = 0;
  dbuf<int2> cl_range;               // (first site, number of sites)
  dbuf<int> cl_off, cl_src;          // CSR: n_clusters+1 offsets, source leaves
  dbuf<unsigned long long> cl_mask;  // lane masks
  dbuf<float4> posf4;     // FP32 source records: offset from the site's leaf centre, alpha_iso^(-1/6)
  dbuf<float4> m2l_rec;   // FP32 M2L pairs in list order: target - source centre (formed in FP64), source
  dbuf<double> M64, L64;  // dipole kind
  dbuf<float> M32, L32;
  dbuf<double> Mc, Lc;    // combined kind (static fields, m = 0 / 2)
  dbuf<double> src10;     // staging of combined sources
  dbuf<double4> pos4;     // near-field source records: x, y, z, alpha_iso^(-1/6)
  // per site, the exclusion partners whose pairs the near field adds (CSR, built once per tree)
  dbuf<int> xn_off, xn_idx;
  bool xn_built = false;
  int reserved_for[2] = {-1, -1};  // near variants (FP64, FP32) fmm_reserve last prepared for
  dbuf<double4> mu4;      // near-field dipole records (FP64 form)
  dbuf<float4> mu4f;      // near-field dipole records (FP32 form)
  // several right-hand sides per apply (kMultiRhs at a time): their expansions and dipole records
  dbuf<double> M64k, L64k;
  dbuf<float> M32k, L32k;
  dbuf<double4> mu4k;
  dbuf<float4> mu4fk;
  // symmetric near field (near_sym.hh): per-site target records -- FP32: offset from the leaf
  // centre and damping rate (times log2 e for the branch-free pair), FP64: position and damping
  // rate -- for the damping factor sym_tdamp, and the reverse slots of each precision
  dbuf<float4> sym_t32;
  dbuf<double4> sym_t64;
  double sym_tdamp[2] = {-1.0, -1.0};
  bool sym_tb2        = false;
  dbuf<float> rev32;
  dbuf<double> rev64;
  // symmetric direct summation (one leaf): tiles of W TS targets, items (tile, source segment,
  // index within the tile), forward-block offsets per tile, and the blocks of each precision
  struct DSym {
    int ts = 0, tb = 0, nitems = 0, ntile = 0;
    long long nfwd = 0, nrev = 0;
    double c[3] = {0.0, 0.0, 0.0};  // the leaf's centre
    dbuf<int4> items;
    dbuf<long long> fwd_off;
  } dsym[2];
  dbuf<float> dfwd32, drev32;
  dbuf<double> dfwd64, drev64;
};

// right-hand sides whose near fields one staged launch evaluates together
constexpr int kMultiRhs = 3;

// near-field dipole record type of each precision
template <typename R>
using MU4Of = typename std::conditional<sizeof(R) == 8, double4, float4>::type;

template <int KIND, int P, typename R>
struct Ops;

#define CPPE_DEFINE_OPS(P)                                                                                    \
  template <typename R>                                                                                       \
  struct Ops<KIND_M1, P, R> {                                                                                 \
    static constexpr int MS = fmm_msize(0, P), LS = fmm_lsize(0, P), FS = fmm_fsize(0, P), NS = 3;            \
    static constexpr int FP = fmm_fstride(0, P);                                                              \
    static __device__ __forceinline__ void p2m(R x, R y, R z, const R* S, R* M) {                             \
      fmmops::p2m_m1_p##P<R>(x, y, z, S, M);                                                                  \
    }                                                                                                         \
    static __device__ __forceinline__ void m2m(R x, R y, R z, const R* Mc, R* Mp) {                           \
      fmmops::m2m_m1_p##P<R>(x, y, z, Mc, Mp);                                                                \
    }                                                                                                         \
    static __device__ __forceinline__ void fold(R* M) { fmmops::fold_m1_p##P<R>(M); }                         \
    static __device__ __forceinline__ void unfold(R* L) { fmmops::unfold_m1_p##P<R>(L); }                     \
    static constexpr bool lkept(int k) { return fmmops::lkept_m1_p##P(k); }                                  \
    static __device__ __forceinline__ void m2l(double x, double y, double z, const double* M, double* L) {    \
      fmmops::m2l_m1_p##P##_f64(x, y, z, M, L);                                                               \
    }                                                                                                         \
    static __device__ __forceinline__ void m2l(float x, float y, float z, const float* M, float* L) {         \
      fmmops::m2l_m1_p##P##_f32(x, y, z, M, L);                                                               \
    }                                                                                                         \
    static __device__ __forceinline__ void l2l(R x, R y, R z, const R* Lp, R* Lc) {                           \
      fmmops::l2l_m1_p##P<R>(x, y, z, Lp, Lc);                                                                \
    }                                                                                                         \
    static __device__ __forceinline__ void l2p(R x, R y, R z, const R* L, R& fx, R& fy, R& fz) {              \
      fmmops::l2p_m1_p##P<R>(x, y, z, L, fx, fy, fz);                                                         \
    }                                                                                                         \
  };                                                                                                          \
  template <typename R>                                                                                       \
  struct Ops<KIND_MC, P, R> {                                                                                 \
    static constexpr int MS = fmm_msize(1, P), LS = fmm_lsize(1, P), FS = fmm_fsize(1, P), NS = 10;           \
    static constexpr int FP = fmm_fstride(1, P);                                                              \
    static __device__ __forceinline__ void p2m(R x, R y, R z, const R* S, R* M) {                             \
      fmmops::p2m_mc_p##P<R>(x, y, z, S, M);                                                                  \
    }                                                                                                         \
    static __device__ __forceinline__ void m2m(R x, R y, R z, const R* Mc, R* Mp) {                           \
      fmmops::m2m_mc_p##P<R>(x, y, z, Mc, Mp);                                                                \
    }                                                                                                         \
    static __device__ __forceinline__ void fold(R* M) { fmmops::fold_mc_p##P<R>(M); }                         \
    static __device__ __forceinline__ void unfold(R* L) { fmmops::unfold_mc_p##P<R>(L); }                     \
    static constexpr bool lkept(int k) { return fmmops::lkept_mc_p##P(k); }                                  \
    static __device__ __forceinline__ void m2l(double x, double y, double z, const double* M, double* L) {    \
      fmmops::m2l_mc_p##P##_f64(x, y, z, M, L);                                                               \
    }                                                                                                         \
    static __device__ __forceinline__ void m2l(float x, float y, float z, const float* M, float* L) {         \
      fmmops::m2l_mc_p##P##_f32(x, y, z, M, L);                                                               \
    }                                                                                                         \
    static __device__ __forceinline__ void l2l(R x, R y, R z, const R* Lp, R* Lc) {                           \
      fmmops::l2l_mc_p##P<R>(x, y, z, Lp, Lc);                                                                \
    }                                                                                                         \
    static __device__ __forceinline__ void l2p(R x, R y, R z, const R* L, R& fx, R& fy, R& fz) {              \
      fmmops::l2p_mc_p##P<R>(x, y, z, L, fx, fy, fz);                                                         \
    }                                                                                                         \
  };

#define CPPE_DEFINE_OPS_EXPAND(P) CPPE_DEFINE_OPS(P)
#ifdef CPPE_FMM_ORDER
CPPE_DEFINE_OPS_EXPAND(CPPE_FMM_ORDER)
#endif

// Sources in device order, SoA: dipole kind = 3 arrays (mu); combined kind = 10 arrays
// (q, mu, packed quadrupole); the far field of the combined kind uses -mu (static field
// convention F(q) - F(mu) + F(theta')).
template <int KIND, typename R>
__device__ __forceinline__ void load_source(const double* __restrict__ src, int n, int i, R* S) {
  if (KIND == KIND_M1) {
    S[0] = (R)src[i];
    S[1] = (R)src[n + i];
    S[2] = (R)src[2 * (size_t)n + i];
  } else {
    S[0] = (R)src[i];
    for (int c = 1; c < 4; ++c) S[c] = -(R)src[(size_t)c * n + i];
    for (int c = 4; c < 10; ++c) S[c] = (R)src[(size_t)c * n + i];
  }
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sum over the G consecutive lanes of a group (G a power of two <= 64), by xor shuffles.
template <int G, typename R>
__device__ __forceinline__ R group_sum(R v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// P2M: G lanes per leaf (64/G leaves per wavefront); lane g expands sites g, g + G, ... of the
// leaf about its centre, the G partial expansions are summed by xor shuffles inside the group and
// the owner lane of each coefficient stores it. Roots are skipped (their expansions never enter
// an M2L pair).
template <int KIND, int P, typename R, int G>
__global__ __launch_bounds__(64) void k_fmm_p2m(int nleaves, const int* __restrict__ leaf_ids,
                                                const int* __restrict__ level,
                                                const int* __restrict__ start, const int* __restrict__ count,
                                                const double* __restrict__ cx, const double* __restrict__ cy,
                                                const double* __restrict__ cz, const double* __restrict__ x,
                                                const double* __restrict__ y, const double* __restrict__ z,
                                                const double* __restrict__ src, int n, R* __restrict__ M) {
  using O          = Ops<KIND, P, R>;
  const int lane   = threadIdx.x;
  const int gl     = lane % G;
  const int t      = blockIdx.x * (64 / G) + lane / G;
  const int c      = t < nleaves ? leaf_ids[t] : 0;
  const bool valid = t < nleaves && level[c] != 0;
  R m[O::MS];
#pragma unroll
  for (int k = 0; k < O::MS; ++k) m[k] = R(0);
  if (valid) {
    const int i0 = start[c], i1 = i0 + count[c];
    const double ccx = cx[c], ccy = cy[c], ccz = cz[c];
    for (int i = i0 + gl; i < i1; i += G) {
      R S[O::NS];
      load_source<KIND, R>(src, n, i, S);
      O::p2m((R)(ccx - x[i]), (R)(ccy - y[i]), (R)(ccz - z[i]), S, m);
    }
  }
  constexpr int NCH = (O::MS + G - 1) / G;
  R mine[NCH];
#pragma unroll
  for (int k = 0; k < NCH; ++k) mine[k] = R(0);
#pragma unroll
  for (int k = 0; k < O::MS; ++k) {
    const R v = group_sum<G>(m[k]);
    if (gl == (k % G)) mine[k / G] = v;
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int k = q * G + gl;
      if (k < O::MS) M[(size_t)c * O::MS + k] = mine[q];
    }
  }
}

// M2M: 8 lanes per parent cell, lane g translates child g (if any); the partial expansions are
// summed by xor shuffles inside the group. Cells without children (leaves) keep their P2M.
template <int KIND, int P, typename R>
__global__ __launch_bounds__(64) void k_fmm_m2m(int c0, int c1, const int* __restrict__ nchild,
                                                const int* __restrict__ child_first, const double* __restrict__ cx,
                                                const double* __restrict__ cy, const double* __restrict__ cz,
                                                R* __restrict__ M) {
  using O          = Ops<KIND, P, R>;
  constexpr int G  = 8;
  const int lane   = threadIdx.x;
  const int gl     = lane % G;
  const int c      = c0 + blockIdx.x * (64 / G) + lane / G;
  const int nc     = c < c1 ? nchild[c] : 0;
  const bool valid = nc > 0;
  R m[O::MS];
#pragma unroll
  for (int k = 0; k < O::MS; ++k) m[k] = R(0);
  if (gl < nc) {
    const int ch = child_first[c] + gl;
    O::m2m((R)(cx[c] - cx[ch]), (R)(cy[c] - cy[ch]), (R)(cz[c] - cz[ch]), &M[(size_t)ch * O::MS], m);
  }
  constexpr int NCH = (O::MS + G - 1) / G;
  R mine[NCH];
#pragma unroll
  for (int k = 0; k < NCH; ++k) mine[k] = R(0);
#pragma unroll
  for (int k = 0; k < O::MS; ++k) {
    const R v = group_sum<G>(m[k]);
    if (gl == (k % G)) mine[k / G] = v;
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int k = q * G + gl;
      if (k < O::MS) M[(size_t)c * O::MS + k] = mine[q];
    }
  }
}

// Folds each cell's multipole expansion onto its harmonic components (see fmm_ops.gen.hh) into
// the compact rows Mf (stride FP) that M2L reads: FS instead of MS coefficients per source.
template <int KIND, int P, typename R>
__global__ __launch_bounds__(64) void k_fmm_fold(int c0, int c1, const R* __restrict__ M, R* __restrict__ Mf) {
  using O     = Ops<KIND, P, R>;
  const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  R m[O::MS];
#pragma unroll
  for (int k = 0; k < O::MS; ++k) m[k] = M[(size_t)c * O::MS + k];
  O::fold(m);
#pragma unroll
  for (int k = 0; k < O::FS; ++k) Mf[(size_t)c * O::FP + k] = m[k];
}

// M2L produces the local coefficients L_n with n_z < 2 only; the others follow from them (the
// local expansion is harmonic) and are restored here, per cell, before L2L.
template <int KIND, int P, typename R>
__global__ __launch_bounds__(64) void k_fmm_unfold(int ncells, R* __restrict__ L) {
  using O     = Ops<KIND, P, R>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  O::unfold(&L[(size_t)c * O::LS]);
}

// M2L: G lanes per target cell (G = 16, 32, 64; 64/G cells per wavefront). Lanes of a group take
// the sources of the target-grouped list in turn; the G partial expansions are summed by xor
// shuffles inside the group and the owner lane of each coefficient stores it. At most four
// wavefronts per SIMD: more (the register count allows five in FP32) was measured slower, the
// concurrent source rows then miss the caches more.
template <int KIND, int P, typename R, int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_fmm_m2l(
    int ncells, const int* __restrict__ off, const int* __restrict__ srcs, const double* __restrict__ cx,
    const double* __restrict__ cy, const double* __restrict__ cz, const R* __restrict__ Mf, R* __restrict__ L,
    const float4* __restrict__ rec) {
  using O         = Ops<KIND, P, R>;
  const int lane  = threadIdx.x;
  const int gl    = lane % G;
  const int A     = blockIdx.x * (64 / G) + lane / G;
  const bool valid = A < ncells;
  R l[O::LS];
#pragma unroll
  for (int k = 0; k < O::LS; ++k) l[k] = R(0);
  if (valid) {
    const int o0 = off[A], o1 = off[A + 1];
    if constexpr (sizeof(R) == 4) {  // one coalesced record per pair: no dependent centre gathers
      for (int k = o0 + gl; k < o1; k += G) {
        const float4 q = rec[k];
        const int B    = __float_as_int(q.w);
        const R* mrow  = static_cast<const R*>(__builtin_assume_aligned(&Mf[(size_t)B * O::FP], 16));
        O::m2l(q.x, q.y, q.z, mrow, l);
      }
    } else {
      (void)rec;
      const double ax = cx[A], ay = cy[A], az = cz[A];
      for (int k = o0 + gl; k < o1; k += G) {
        const int B = srcs[k];
        const R* mrow = static_cast<const R*>(__builtin_assume_aligned(&Mf[(size_t)B * O::FP], 16));
        O::m2l((R)(ax - cx[B]), (R)(ay - cy[B]), (R)(az - cz[B]), mrow, l);
      }
    }
  }
  constexpr int NCH = (O::LS + G - 1) / G;
  R mine[NCH];
#pragma unroll
  for (int k = 0; k < NCH; ++k) mine[k] = R(0);
#pragma unroll
  for (int k = 0; k < O::LS; ++k) {
    if (!O::lkept(k)) continue;  // stored as 0 here, restored by the downward pass (down_cell)
    const R v = group_sum<G>(l[k]);
    if (gl == (k % G)) mine[k / G] = v;
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
      const int k = q * G + gl;
      if (k < O::LS) L[(size_t)A * O::LS + k] = mine[q];
    }
  }
}

// Downward pass of one cell: its M2L result unfolded (the coefficients M2L left out restored, as
// k_fmm_unfold does) and, below the root, its parent's complete expansion translated into it.
template <int KIND, int P, typename R>
__device__ __forceinline__ void down_cell(int c, bool child, const int* __restrict__ parent,
                                          const double* __restrict__ cx, const double* __restrict__ cy,
                                          const double* __restrict__ cz, R* __restrict__ L) {
  using O = Ops<KIND, P, R>;
  R l[O::LS];
#pragma unroll
  for (int k = 0; k < O::LS; ++k) l[k] = L[(size_t)c * O::LS + k];
  O::unfold(l);
  if (child) {
    const int pa = parent[c];
    O::l2l((R)(cx[c] - cx[pa]), (R)(cy[c] - cy[pa]), (R)(cz[c] - cz[pa]), &L[(size_t)pa * O::LS], l);
  }
#pragma unroll
  for (int k = 0; k < O::LS; ++k) L[(size_t)c * O::LS + k] = l[k];
}

// Cells [c0, c1) of one level below the top levels (their parents complete).
template <int KIND, int P, typename R>
__global__ __launch_bounds__(64) void k_fmm_l2l(int c0, int c1, const int* __restrict__ parent,
                                                const double* __restrict__ cx, const double* __restrict__ cy,
                                                const double* __restrict__ cz, R* __restrict__ L) {
  const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  down_cell<KIND, P, R>(c, true, parent, cx, cy, cz, L);
}

// The top levels of the downward pass (root first, few cells each) in one workgroup: level by
// level with a barrier between, so one launch replaces a launch per level.
constexpr int kTopLevels = 8, kTopCells = 1024;
struct TopLevels {
  int nlev;                    // levels 0 .. nlev - 1
  int off[kTopLevels + 1];     // first cell of each level, then the end
};

// 512 threads for the dipole-field operators (one pass over a level of up to 512 cells); 256 for
// the static-field ones, whose FP64 expansions need twice the registers
template <int KIND>
constexpr int down_top_threads() {
  return KIND == KIND_M1 ? 512 : 256;
}

template <int KIND, int P, typename R>
__global__ __launch_bounds__(down_top_threads<KIND>()) void k_fmm_down_top(TopLevels t, const int* __restrict__ parent,
                                                      const double* __restrict__ cx, const double* __restrict__ cy,
                                                      const double* __restrict__ cz, R* __restrict__ L) {
  for (int lev = 0; lev < t.nlev; ++lev) {
    for (int c = t.off[lev] + threadIdx.x; c < t.off[lev + 1]; c += blockDim.x)
      down_cell<KIND, P, R>(c, lev > 0, parent, cx, cy, cz, L);
    __syncthreads();
  }
}

struct P2PArgs {
  const int* leaf_ids;
  const int* start;
  const int* count;
  const int* p2p_off;
  const int* p2p_src;
  const double *cx, *cy, *cz;
  const double *x, *y, *z;
  const double* s;  // alpha_iso^(-1/6) (0: undamped site)
  double damping;
  const int* eoff;
  const int* eidx;
  const double* src;   // sources, SoA in device order (3 or 10 arrays)
  const double* ainv;  // alpha^-1 packed SoA (response operator only)
  int n;
  double* out;         // 3n SoA
};

template <int KIND, int P, typename R, bool DAMP, int MODE>
__global__ __launch_bounds__(64) void k_fmm_p2p(P2PArgs a, const R* __restrict__ L) {
  using O            = Ops<KIND, P, R>;
  constexpr int NS   = (MODE == MODE_STATIC) ? 10 : 3;
  __shared__ R sx[64], sy[64], sz[64], ss[64];
  __shared__ R sS[NS][64];
  __shared__ int sj[64], spre[64], sst[64];
  __shared__ double px[64], py[64], pz[64];
  const int lane = threadIdx.x;
  const int A    = a.leaf_ids[blockIdx.x];
  const int T    = a.count[A], i0 = a.start[A];
  const int o0 = a.p2p_off[A], o1 = a.p2p_off[A + 1];
  const double cax = a.cx[A], cay = a.cy[A], caz = a.cz[A];
  const int n = a.n;
  for (int tb = 0; tb < T; tb += 64) {
    const int Tb = min(64, T - tb), Sb = 64 / Tb;
    const int t = lane % Tb, sl = lane / Tb;
    const bool act = sl < Sb;
    const int i    = i0 + tb + t;
    // target: FP64 kernels use absolute coordinates, FP32 kernels offsets from the leaf centre
    const R xi = sizeof(R) == 8 ? (R)a.x[i] : (R)(a.x[i] - cax);
    const R yi = sizeof(R) == 8 ? (R)a.y[i] : (R)(a.y[i] - cay);
    const R zi = sizeof(R) == 8 ? (R)a.z[i] : (R)(a.z[i] - caz);
    const R di = DAMP ? (R)(a.damping * a.s[i]) : R(0);
    int lo = i, hi = i;
    if (a.eoff)
      for (int k = a.eoff[i]; k < a.eoff[i + 1]; ++k) {
        const int e = a.eidx[k];
        if (e >= 0) {
          lo = min(lo, e);
          hi = max(hi, e);
        }
      }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int g0 = o0; g0 < o1; g0 += 64) {
      const int gl = min(64, o1 - g0);
      int cnt = 0, st = 0;
      if (lane < gl) {
        const int B = a.p2p_src[g0 + lane];
        cnt         = a.count[B];
        st          = a.start[B];
      }
      int incl = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
      }
      const int total = __shfl(incl, 63);
      __syncthreads();
      spre[lane] = incl - cnt;
      sst[lane]  = st;
      __syncthreads();
      for (int c0 = 0; c0 < total; c0 += 64) {
        const int q = c0 + lane;
        if (q < total) {
          int lo2 = 0, hi2 = gl;  // largest k with spre[k] <= q
          while (hi2 - lo2 > 1) {
            const int mid = (lo2 + hi2) >> 1;
            if (spre[mid] <= q)
              lo2 = mid;
            else
              hi2 = mid;
          }
          const int j = sst[lo2] + (q - spre[lo2]);
          sj[lane]    = j;
          if (sizeof(R) == 8) {
            sx[lane] = (R)a.x[j];
            sy[lane] = (R)a.y[j];
            sz[lane] = (R)a.z[j];
          } else {
            sx[lane] = (R)(a.x[j] - cax);
            sy[lane] = (R)(a.y[j] - cay);
            sz[lane] = (R)(a.z[j] - caz);
          }
          ss[lane] = DAMP ? (R)a.s[j] : R(0);
#pragma unroll
          for (int c = 0; c < NS; ++c) sS[c][lane] = (R)a.src[(size_t)c * n + j];
        }
        __syncthreads();
        const int cn = min(64, total - c0);
        if (act) {
          R gx = R(0), gy = R(0), gz = R(0);
          for (int u = sl; u < cn; u += Sb) {
            const int j = sj[u];
            if (j >= lo && j <= hi && excluded(i, j, a.eoff, a.eidx)) continue;
            const R dx = xi - sx[u], dy = yi - sy[u], dz = zi - sz[u];
            if (MODE == MODE_STATIC) {
              double Sq[6];
#pragma unroll
              for (int c = 0; c < 6; ++c) Sq[c] = (double)sS[4 + c][u];
              double hx = 0.0, hy = 0.0, hz = 0.0;
              pair_static((double)dx, (double)dy, (double)dz, (double)sS[0][u], (double)sS[1][u],
                          (double)sS[2][u], (double)sS[3][u], Sq, (double)(di * ss[u]), DAMP, hx, hy, hz);
              gx += (R)hx;
              gy += (R)hy;
              gz += (R)hz;
            } else if (sizeof(R) == 8) {
              double hx = 0.0, hy = 0.0, hz = 0.0;
              pair_m1_f64<DAMP>((double)dx, (double)dy, (double)dz, (double)sS[0][u], (double)sS[1][u],
                                (double)sS[2][u], (double)(di * ss[u]), hx, hy, hz);
              gx += (R)hx;
              gy += (R)hy;
              gz += (R)hz;
            } else {
              float hx = 0.f, hy = 0.f, hz = 0.f;
              pair_m1_f32<DAMP>((float)dx, (float)dy, (float)dz, (float)sS[0][u], (float)sS[1][u],
                                (float)sS[2][u], (float)(di * ss[u]), hx, hy, hz);
              gx += (R)hx;
              gy += (R)hy;
              gz += (R)hz;
            }
          }
          fx += (double)gx;
          fy += (double)gy;
          fz += (double)gz;
        }
        __syncthreads();
      }
    }
    px[lane] = act ? fx : 0.0;
    py[lane] = act ? fy : 0.0;
    pz[lane] = act ? fz : 0.0;
    __syncthreads();
    if (lane < Tb) {
      double Fx = 0.0, Fy = 0.0, Fz = 0.0;
      for (int s2 = 0; s2 < Sb; ++s2) {
        Fx += px[lane + Tb * s2];
        Fy += py[lane + Tb * s2];
        Fz += pz[lane + Tb * s2];
      }
      R lx = R(0), ly = R(0), lz = R(0);
      O::l2p((R)(a.x[i] - cax), (R)(a.y[i] - cay), (R)(a.z[i] - caz), &L[(size_t)A * O::LS], lx, ly, lz);
      Fx += (double)lx;
      Fy += (double)ly;
      Fz += (double)lz;
      if (MODE == MODE_BAPPLY) {
        const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
        const double* ai = a.ainv;
        const double a0 = ai[i], a1 = ai[n + i], a2 = ai[2 * (size_t)n + i], a3 = ai[3 * (size_t)n + i],
                     a4 = ai[4 * (size_t)n + i], a5 = ai[5 * (size_t)n + i];
        Fx += a0 * m0 + a1 * m1 + a2 * m2;
        Fy += a1 * m0 + a3 * m1 + a4 * m2;
        Fz += a2 * m0 + a4 * m1 + a5 * m2;
      }
      a.out[i]                 = Fx;
      a.out[n + i]             = Fy;
      a.out[2 * (size_t)n + i] = Fz;
    }
    __syncthreads();
  }
}

// Near field for dipole sources, one thread per target site (device order). Each lane walks the
// near list of its own leaf with a private cursor over the concatenated source sites; lanes of
// the same leaf (consecutive in Morton order) load identical records, which the memory
// pipeline coalesces, so no LDS staging or barriers are needed. Source records are 32-byte
// (x, y, z, s) and (mu, 0) vectors. FP32 form: the displacement is formed in FP64 and rounded,
// the rest of the pair arithmetic is FP32.
struct NearArgs {
  const int* leaf_of;
  const int* start;
  const int* count;
  const int* p2p_off;
  const int* p2p_src;
  const double *cx, *cy, *cz;
  const double4* pos4;
  const int* eoff;
  const int* eidx;
  const double* src;   // mu, 3 SoA arrays (alpha^-1 epilogue)
  const double* ainv;  // packed alpha^-1, 6 SoA arrays
  double damping;
  int n;
  double* out;
};

template <int P, typename R, bool DAMP, int MODE, typename MU4>
__global__ __launch_bounds__(256) void k_fmm_near(NearArgs a, const MU4* __restrict__ mu4, const R* __restrict__ L) {
  using O     = Ops<KIND_M1, P, R>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int A     = a.leaf_of[i];
  const double4 pi = a.pos4[i];
  const R di       = DAMP ? (R)(a.damping * pi.w) : R(0);
  int lo = i, hi = i;
  if (a.eoff)
    for (int k = a.eoff[i]; k < a.eoff[i + 1]; ++k) {
      const int e = a.eidx[k];
      if (e >= 0) {
        lo = min(lo, e);
        hi = max(hi, e);
      }
    }
  int k = a.p2p_off[A];
  const int kend = a.p2p_off[A + 1];
  R fx = R(0), fy = R(0), fz = R(0);
  if (k < kend) {
    int B = a.p2p_src[k];
    int j = a.start[B], jend = j + a.count[B];
    while (true) {
      const double4 pj = a.pos4[j];
      const MU4 mj     = mu4[j];
      if (!(j >= lo && j <= hi && excluded(i, j, a.eoff, a.eidx))) {
        if constexpr (sizeof(R) == 8) {
          double hx = 0.0, hy = 0.0, hz = 0.0;
          pair_m1_f64<DAMP>(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, (double)mj.x, (double)mj.y, (double)mj.z,
                            (double)di * pj.w, hx, hy, hz);
          fx += (R)hx;
          fy += (R)hy;
          fz += (R)hz;
        } else {
          float hx = 0.f, hy = 0.f, hz = 0.f;
          pair_m1_f32<DAMP>((float)(pi.x - pj.x), (float)(pi.y - pj.y), (float)(pi.z - pj.z), (float)mj.x,
                            (float)mj.y, (float)mj.z, (float)di * (float)pj.w, hx, hy, hz);
          fx += (R)hx;
          fy += (R)hy;
          fz += (R)hz;
        }
      }
      if (++j == jend) {
        if (++k == kend) break;
        B    = a.p2p_src[k];
        j    = a.start[B];
        jend = j + a.count[B];
      }
    }
  }
  R lx = R(0), ly = R(0), lz = R(0);
  O::l2p((R)(pi.x - a.cx[A]), (R)(pi.y - a.cy[A]), (R)(pi.z - a.cz[A]), &L[(size_t)A * O::LS], lx, ly, lz);
  double Fx = (double)fx + (double)lx, Fy = (double)fy + (double)ly, Fz = (double)fz + (double)lz;
  const int n = a.n;
  if (MODE == MODE_BAPPLY) {
    const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
    const double* ai = a.ainv;
    const double a0 = ai[i], a1 = ai[n + i], a2 = ai[2 * (size_t)n + i], a3 = ai[3 * (size_t)n + i],
                 a4 = ai[4 * (size_t)n + i], a5 = ai[5 * (size_t)n + i];
    Fx += a0 * m0 + a1 * m1 + a2 * m2;
    Fy += a1 * m0 + a3 * m1 + a4 * m2;
    Fz += a2 * m0 + a4 * m1 + a5 * m2;
  }
  a.out[i]                 = Fx;
  a.out[n + i]             = Fy;
  a.out[2 * (size_t)n + i] = Fz;
}

// Near field, one wavefront per 64-target chunk of a leaf (work item = (leaf, first target)).
// The source walk is uniform across the wavefront -- every lane visits the same source site at
// the same time -- so the source records are read once per wavefront (scalar loads) and the
// lanes differ only in their target. FP32 form: per source leaf, each lane forms its target's
// offset from that leaf's centre in FP64 and rounds it; sources carry FP32 offsets from their
// own leaf centre (posf4), so FP32 only ever sees offsets of a few leaf widths.
template <int P, typename R, bool DAMP, int MODE, typename MU4, typename POSR>
__global__ __launch_bounds__(256) void k_fmm_near_lw(NearArgs a, const int2* __restrict__ items, int nitems,
                                                     const POSR* __restrict__ posr, const MU4* __restrict__ mu4,
                                                     const R* __restrict__ L) {
  using O        = Ops<KIND_M1, P, R>;
  const int w    = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (w >= nitems) return;  // uniform per wavefront
  const int2 item  = items[w];
  const int A      = __builtin_amdgcn_readfirstlane(item.x);
  const int t      = __builtin_amdgcn_readfirstlane(item.y) + lane;
  const bool act   = t < a.count[A];
  const int i      = a.start[A] + (act ? t : 0);  // idle lanes shadow the first target, store nothing
  const double4 pi = a.pos4[i];
  const R di       = DAMP ? (R)(a.damping * pi.w) : R(0);
  int lo = i, hi = i;
  if (a.eoff)
    for (int k = a.eoff[i]; k < a.eoff[i + 1]; ++k) {
      const int e = a.eidx[k];
      if (e >= 0) {
        lo = min(lo, e);
        hi = max(hi, e);
      }
    }
  R fx = R(0), fy = R(0), fz = R(0);
  const int k1 = a.p2p_off[A + 1];
  for (int k = a.p2p_off[A]; k < k1; ++k) {
    const int B  = a.p2p_src[k];
    const int j0 = a.start[B], j1 = j0 + a.count[B];
    R tx, ty, tz;
    if constexpr (sizeof(R) == 8) {
      tx = pi.x;
      ty = pi.y;
      tz = pi.z;
    } else {
      tx = (R)(pi.x - a.cx[B]);
      ty = (R)(pi.y - a.cy[B]);
      tz = (R)(pi.z - a.cz[B]);
    }
    for (int j = j0; j < j1; ++j) {
      const POSR pj = posr[j];
      const MU4 mj  = mu4[j];
      if (j >= lo && j <= hi && excluded(i, j, a.eoff, a.eidx)) continue;
      if constexpr (sizeof(R) == 8) {
        double hx = 0.0, hy = 0.0, hz = 0.0;
        pair_m1_f64<DAMP>(tx - pj.x, ty - pj.y, tz - pj.z, (double)mj.x, (double)mj.y, (double)mj.z,
                          (double)di * pj.w, hx, hy, hz);
        fx += hx;
        fy += hy;
        fz += hz;
      } else {
        pair_m1_f32<DAMP>(tx - pj.x, ty - pj.y, tz - pj.z, mj.x, mj.y, mj.z, di * pj.w, fx, fy, fz);
      }
    }
  }
  R lx = R(0), ly = R(0), lz = R(0);
  O::l2p((R)(pi.x - a.cx[A]), (R)(pi.y - a.cy[A]), (R)(pi.z - a.cz[A]), &L[(size_t)A * O::LS], lx, ly, lz);
  double Fx = (double)fx + (double)lx, Fy = (double)fy + (double)ly, Fz = (double)fz + (double)lz;
  if (!act) return;
  const int n = a.n;
  if (MODE == MODE_BAPPLY) {
    const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
    const double* ai = a.ainv;
    const double a0 = ai[i], a1 = ai[n + i], a2 = ai[2 * (size_t)n + i], a3 = ai[3 * (size_t)n + i],
                 a4 = ai[4 * (size_t)n + i], a5 = ai[5 * (size_t)n + i];
    Fx += a0 * m0 + a1 * m1 + a2 * m2;
    Fy += a1 * m0 + a3 * m1 + a4 * m2;
    Fz += a2 * m0 + a4 * m1 + a5 * m2;
  }
  a.out[i]                 = Fx;
  a.out[n + i]             = Fy;
  a.out[2 * (size_t)n + i] = Fz;
}

// Near field, one wavefront per 64-target chunk of a leaf, lanes split as (target t, source slot
// s): T = targets of the chunk, S = 64 / T slots, so up to T*S lanes evaluate a pair each step.
// The sources are the leaf's expanded near list (the concatenated site ranges of its near leaves,
// built once per tree); slot s takes positions s, s + S, ... of it, so consecutive slots read
// consecutive records. Slot partials are summed through LDS at the end. The pair arithmetic is
// that of the thread-per-target kernel (FP32 form: displacement formed in FP64, then rounded).
template <int P, typename R, bool DAMP, int MODE, typename MU4>
__global__ __launch_bounds__(256) void k_fmm_near_sl(NearArgs a, const int2* __restrict__ items, int nitems,
                                                     const long long* __restrict__ xoff,
                                                     const int* __restrict__ xsrc, const MU4* __restrict__ mu4,
                                                     const R* __restrict__ L) {
  using O = Ops<KIND_M1, P, R>;
  __shared__ R red[3][256];
  const int wl   = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int w    = blockIdx.x * (blockDim.x >> 6) + wl;
  const bool wv  = w < nitems;
  int A = 0, t0 = 0, T = 1;
  if (wv) {
    const int2 item = items[w];
    A               = item.x;
    t0              = item.y;
    T               = min(64, a.count[A] - t0);
  }
  const int S    = 64 / T;
  const int t    = lane % T, sl = lane / T;
  const bool act = wv && sl < S;
  const int i    = a.start[A] + t0 + t;
  R fx = R(0), fy = R(0), fz = R(0);
  double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
  if (wv) pi = a.pos4[i];
  if (act) {
    const R di = DAMP ? (R)(a.damping * pi.w) : R(0);
    int lo = i, hi = i;
    if (a.eoff)
      for (int k = a.eoff[i]; k < a.eoff[i + 1]; ++k) {
        const int e = a.eidx[k];
        if (e >= 0) {
          lo = min(lo, e);
          hi = max(hi, e);
        }
      }
    const long long q1 = xoff[A + 1];
    for (long long q = xoff[A] + sl; q < q1; q += S) {
      const int j       = xsrc[q];
      const double4 pj  = a.pos4[j];
      const MU4 mj      = mu4[j];
      if (j >= lo && j <= hi && excluded(i, j, a.eoff, a.eidx)) continue;
      if constexpr (sizeof(R) == 8) {
        double hx = 0.0, hy = 0.0, hz = 0.0;
        pair_m1_f64<DAMP>(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, (double)mj.x, (double)mj.y, (double)mj.z,
                          (double)di * pj.w, hx, hy, hz);
        fx += hx;
        fy += hy;
        fz += hz;
      } else {
        pair_m1_f32<DAMP>((float)(pi.x - pj.x), (float)(pi.y - pj.y), (float)(pi.z - pj.z), (float)mj.x,
                          (float)mj.y, (float)mj.z, di * (float)pj.w, fx, fy, fz);
      }
    }
  }
  red[0][threadIdx.x] = fx;
  red[1][threadIdx.x] = fy;
  red[2][threadIdx.x] = fz;
  __syncthreads();
  if (!wv || lane >= T) return;
  const int base = wl * 64 + lane;
  double Fx = 0.0, Fy = 0.0, Fz = 0.0;
  for (int s2 = 0; s2 < S; ++s2) {
    Fx += (double)red[0][base + T * s2];
    Fy += (double)red[1][base + T * s2];
    Fz += (double)red[2][base + T * s2];
  }
  R lx = R(0), ly = R(0), lz = R(0);
  O::l2p((R)(pi.x - a.cx[A]), (R)(pi.y - a.cy[A]), (R)(pi.z - a.cz[A]), &L[(size_t)A * O::LS], lx, ly, lz);
  Fx += (double)lx;
  Fy += (double)ly;
  Fz += (double)lz;
  const int n = a.n;
  if (MODE == MODE_BAPPLY) {
    const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
    const double* ai = a.ainv;
    const double a0 = ai[i], a1 = ai[n + i], a2 = ai[2 * (size_t)n + i], a3 = ai[3 * (size_t)n + i],
                 a4 = ai[4 * (size_t)n + i], a5 = ai[5 * (size_t)n + i];
    Fx += a0 * m0 + a1 * m1 + a2 * m2;
    Fy += a1 * m0 + a3 * m1 + a4 * m2;
    Fz += a2 * m0 + a4 * m1 + a5 * m2;
  }
  a.out[i]                 = Fx;
  a.out[n + i]             = Fy;
  a.out[2 * (size_t)n + i] = Fz;
}

// Near field over target clusters: one wavefront per cluster of whole leaves (<= 64 consecutive
// target sites, one per lane). The cluster walks the union of its leaves' near lists; each entry
// carries the mask of lanes for which that source leaf is near (the others take it through M2L
// and skip it). The walk is uniform across the wavefront, so source records arrive through
// scalar loads and lanes differ only in their target. Lanes whose exclusion range touches the
// source leaf take a checked loop; all others a loop without tests. FP32 form: target offsets
// from the source leaf centre are formed in FP64 and rounded once per source leaf; sources carry
// FP32 offsets from their own leaf centre.
template <typename R, bool DAMP, typename MU4, typename POSR, bool CHECK>
__device__ __forceinline__ void cluster_leaf(int i, int lo, int hi, int j0, int j1, R tx, R ty, R tz, R di,
                                             const POSR* __restrict__ posr, const MU4* __restrict__ mu4,
                                             const int* __restrict__ eoff, const int* __restrict__ eidx,
                                             R& fx, R& fy, R& fz) {
  for (int j = j0; j < j1; ++j) {
    const POSR pj = posr[j];
    const MU4 mj  = mu4[j];
    if (CHECK && j >= lo && j <= hi && excluded(i, j, eoff, eidx)) continue;
    if constexpr (sizeof(R) == 8) {
      pair_m1_f64<DAMP>(tx - pj.x, ty - pj.y, tz - pj.z, mj.x, mj.y, mj.z, di * pj.w, fx, fy, fz);
    } else {
      pair_m1_f32<DAMP>(tx - pj.x, ty - pj.y, tz - pj.z, mj.x, mj.y, mj.z, di * pj.w, fx, fy, fz);
    }
  }
}

template <int P, typename R, bool DAMP, int MODE, typename MU4, typename POSR>
__global__ __launch_bounds__(256) void k_fmm_near_cl(NearArgs a, int ncl, const int2* __restrict__ range,
                                                     const int* __restrict__ uoff, const int* __restrict__ usrc,
                                                     const unsigned long long* __restrict__ umask,
                                                     const POSR* __restrict__ posr, const MU4* __restrict__ mu4,
                                                     const R* __restrict__ L) {
  using O        = Ops<KIND_M1, P, R>;
  const int g    = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (g >= ncl) return;  // uniform per wavefront
  const int2 rg   = range[g];
  const int g0    = __builtin_amdgcn_readfirstlane(rg.x);
  const int nt    = __builtin_amdgcn_readfirstlane(rg.y);
  const bool act  = lane < nt;
  const int i     = g0 + (act ? lane : 0);
  const double4 pi = a.pos4[i];
  const R di       = DAMP ? (R)(a.damping * pi.w) : R(0);
  int lo = i, hi = i;
  if (a.eoff)
    for (int k = a.eoff[i]; k < a.eoff[i + 1]; ++k) {
      const int e = a.eidx[k];
      if (e >= 0) {
        lo = min(lo, e);
        hi = max(hi, e);
      }
    }
  const unsigned long long bit = 1ull << lane;
  R fx = R(0), fy = R(0), fz = R(0);
  const int k1 = uoff[g + 1];
  for (int k = uoff[g]; k < k1; ++k) {
    const int B = usrc[k];
    if (!act || !(umask[k] & bit)) continue;
    const int j0 = a.start[B], j1 = j0 + a.count[B];
    R tx, ty, tz;
    if constexpr (sizeof(R) == 8) {
      tx = pi.x;
      ty = pi.y;
      tz = pi.z;
    } else {
      tx = (R)(pi.x - a.cx[B]);
      ty = (R)(pi.y - a.cy[B]);
      tz = (R)(pi.z - a.cz[B]);
    }
    if (lo < j1 && hi >= j0)
      cluster_leaf<R, DAMP, MU4, POSR, true>(i, lo, hi, j0, j1, tx, ty, tz, di, posr, mu4, a.eoff, a.eidx, fx, fy,
                                             fz);
    else
      cluster_leaf<R, DAMP, MU4, POSR, false>(i, lo, hi, j0, j1, tx, ty, tz, di, posr, mu4, a.eoff, a.eidx, fx,
                                              fy, fz);
  }
  if (!act) return;
  const int A = a.leaf_of[i];
  R lx = R(0), ly = R(0), lz = R(0);
  O::l2p((R)(pi.x - a.cx[A]), (R)(pi.y - a.cy[A]), (R)(pi.z - a.cz[A]), &L[(size_t)A * O::LS], lx, ly, lz);
  double Fx = (double)fx + (double)lx, Fy = (double)fy + (double)ly, Fz = (double)fz + (double)lz;
  const int n = a.n;
  if (MODE == MODE_BAPPLY) {
    const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
    const double* ai = a.ainv;
    const double a0 = ai[i], a1 = ai[n + i], a2 = ai[2 * (size_t)n + i], a3 = ai[3 * (size_t)n + i],
                 a4 = ai[4 * (size_t)n + i], a5 = ai[5 * (size_t)n + i];
    Fx += a0 * m0 + a1 * m1 + a2 * m2;
    Fy += a1 * m0 + a3 * m1 + a4 * m2;
    Fz += a2 * m0 + a4 * m1 + a5 * m2;
  }
  a.out[i]                 = Fx;
  a.out[n + i]             = Fy;
  a.out[2 * (size_t)n + i] = Fz;
}

// Orders LDS accesses of one wavefront (no cross-wavefront synchronisation).
__device__ __forceinline__ void wave_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS source record of the staged near-field kernel: position (FP32: offset from the target
// leaf centre; FP64: absolute), alpha_iso^(-1/6), and the dipole of each of K right-hand sides.
template <typename R, int K>
struct alignas(K == 1 ? 4 * sizeof(R) : 16) StagedRec {  // K = 1: two 16-byte (FP32) / 32-byte (FP64) reads
  R x, y, z, s;
  R m[3 * K];
};

// Geometry of one near-field pair of the staged kernel and of the exclusion pass (identical
// arithmetic, so an excluded pair is taken out exactly as it was added). B2: the branch-free FP32
// form for systems whose sites are all damped; di then carries the factor log2(e).
// FAR: a source past the precision's far class for every target of the leaf (never the target
// itself): the undamped geometry (FP32: from v >= 18 on, where the screening it leaves out moves
// the field by less than the FP32 rounding; FP64 takes no far stages).
template <typename R, bool DAMP, bool B2, bool FAR = false>
__device__ __forceinline__ void staged_geo(R dx, R dy, R dz, R a, R& c3, R& cf) {
  if constexpr (sizeof(R) == 8)
    geo_m1_f64<DAMP && !FAR, !FAR>(dx, dy, dz, a, c3, cf);
  else if constexpr (FAR)
    geo_m1_f32<false, false>(dx, dy, dz, a, c3, cf);
  else if constexpr (B2)
    geo_m1_f32_b2(dx, dy, dz, a, c3, cf);
  else
    geo_m1_f32<DAMP, true>(dx, dy, dz, a, c3, cf);
}

// one pair against the K dipoles of a staged record; f holds 3 K accumulators
template <typename R, bool DAMP, bool B2, int K, bool FAR = false>
__device__ __forceinline__ void staged_pair(R xi, R yi, R zi, R di, const StagedRec<R, K>& r, R* f) {
  const R dx = xi - r.x, dy = yi - r.y, dz = zi - r.z;
  if constexpr (FAR && sizeof(R) == 4) {
    float ri3, m3;
    geo_far_f32(dx, dy, dz, ri3, m3);
#pragma unroll
    for (int k = 0; k < K; ++k)
      acc_far_f32(ri3, m3, dx, dy, dz, r.m[3 * k], r.m[3 * k + 1], r.m[3 * k + 2], f[3 * k], f[3 * k + 1],
                  f[3 * k + 2]);
  } else {
    R c3, cf;
    staged_geo<R, DAMP, B2, FAR>(dx, dy, dz, di * r.s, c3, cf);
#pragma unroll
    for (int k = 0; k < K; ++k)
      acc_m1(c3, cf, dx, dy, dz, r.m[3 * k], r.m[3 * k + 1], r.m[3 * k + 2], f[3 * k], f[3 * k + 1], f[3 * k + 2]);
  }
}

template <typename R, bool B2>
__device__ __forceinline__ R staged_di(double damping, double s) {
  constexpr double kLog2e = 1.44269504088896340736;
  return (R)(B2 ? damping * s * kLog2e : damping * s);
}

// Near field, one wavefront per leaf chunk of at most `chunk` (64, 32, 16 or 8) targets, several
// wavefronts per block, each with its own LDS region and no block-wide barrier. Lanes are (target
// t, source slot s): T = targets of the chunk, S = 64 / T; smaller chunks spread a large leaf over
// several wavefronts with more slots each. The leaf's expanded near list is consumed 64 sources at a time: each
// lane gathers one source record into LDS (coalesced index load), then slot s evaluates sources
// s, s + S, ... of the chunk against its target. FP32 form: positions staged as offsets from the
// target leaf centre, formed in FP64 and rounded. Slot partials are summed through LDS.
// CHT = 8: built for chunks of 8 -- a full chunk against a full stage of 64 sources runs as a
// straight 8-pair sequence per lane with constant LDS offsets (no per-pair loop arithmetic).
// K right-hand sides (B applies only): each pair's geometry is evaluated once and applied to K
// dipoles; dipole records of right-hand side k at mu4 + k n, its L at L + k lstride, and
// src / out at + 3 k n.
template <int P, typename R, bool DAMP, int MODE, typename MU4, bool WITH_L2P = true, int CHT = 0, bool B2 = false,
          int K = 1>
__global__ __launch_bounds__(256) void k_fmm_near_st(NearArgs a, const int2* __restrict__ items, int nitems,
                                                     int chunk, const long long* __restrict__ xoff,
                                                     const int* __restrict__ xsrc, const MU4* __restrict__ mu4,
                                                     const R* __restrict__ L, size_t lstride = 0,
                                                     const long long* __restrict__ xsplit = nullptr) {
  using O = Ops<KIND_M1, P, R>;
  static_assert(3 * K * 64 * sizeof(R) <= 64 * sizeof(StagedRec<R, K>), "slot sums reuse the record LDS");
  __shared__ StagedRec<R, K> rec[4][64];
  const int wl   = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int w    = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wl);
  if (w >= nitems) return;  // uniform per wavefront; no block-wide barrier follows
  if (CHT) chunk = CHT;
  const int2 item = items[w];
  const int A     = __builtin_amdgcn_readfirstlane(item.x);
  const int t0    = __builtin_amdgcn_readfirstlane(item.y);
  const int T     = min(chunk, a.count[A] - t0);
  const int S     = 64 / T;
  const int t = lane % T, sl = lane / T;
  const bool act = sl < S;
  const int i    = a.start[A] + t0 + t;
  const int n    = a.n;
  const double4 pi = a.pos4[i];
  const double cax = a.cx[A], cay = a.cy[A], caz = a.cz[A];
  const R xi = sizeof(R) == 8 ? (R)pi.x : (R)(pi.x - cax);
  const R yi = sizeof(R) == 8 ? (R)pi.y : (R)(pi.y - cay);
  const R zi = sizeof(R) == 8 ? (R)pi.z : (R)(pi.z - caz);
  const R di = DAMP ? staged_di<R, B2>(a.damping, pi.w) : R(0);
  StagedRec<R, K>* sr = rec[wl];
  R f[3 * K];
#pragma unroll
  for (int c = 0; c < 3 * K; ++c) f[c] = R(0);
  const long long q1 = xoff[A + 1];
  // stages from qf on hold far sources only (FP32: about half of the near pairs at 300k sites; FP64,
  // with its higher cutoff, only a few percent, which would not pay for the extra code path)
  const long long qf = (DAMP && sizeof(R) == 4 && xsplit) ? xsplit[A] : q1;
  // FP32 records: software pipeline over the stages -- the source index of stage s + 2 and the
  // position / dipole values of stage s + 1 are loaded while stage s is evaluated (lanes beyond a
  // stage's count load nothing and keep values they never store). FP64 records are gathered just
  // before their stage is stored: the pipeline's registers cost that kernel more than it hides.
  constexpr bool kPipe = sizeof(R) == 4;
  const long long qa = xoff[A];
  int jnn = 0;  // source index of the stage after the next (pipelined form)
  double4 pn = make_double4(0.0, 0.0, 0.0, 0.0);
  MU4 mn[K];
  auto gather = [&](int j) {
    pn = a.pos4[j];
#pragma unroll
    for (int k = 0; k < K; ++k) mn[k] = mu4[(size_t)k * n + j];
  };
  if (kPipe) {
    if (lane < q1 - qa) gather(xsrc[qa + lane]);
    if (lane < q1 - qa - 64) jnn = xsrc[qa + 64 + lane];
  }
  for (long long q0 = qa; q0 < q1; q0 += 64) {
    const int cn = (int)min<long long>(64, q1 - q0);
    if (!kPipe && lane < cn) gather(xsrc[q0 + lane]);
    if (lane < cn) {
      StagedRec<R, K> r;
      r.x = sizeof(R) == 8 ? (R)pn.x : (R)(pn.x - cax);
      r.y = sizeof(R) == 8 ? (R)pn.y : (R)(pn.y - cay);
      r.z = sizeof(R) == 8 ? (R)pn.z : (R)(pn.z - caz);
      r.s = (R)pn.w;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        r.m[3 * k]     = (R)mn[k].x;
        r.m[3 * k + 1] = (R)mn[k].y;
        r.m[3 * k + 2] = (R)mn[k].z;
      }
      sr[lane] = r;
    }
    wave_lds_barrier();
    if (kPipe) {
      if (lane < q1 - q0 - 64) gather(jnn);                    // stage s + 1
      if (lane < q1 - q0 - 128) jnn = xsrc[q0 + 128 + lane];  // stage s + 2: its source index
    }
    // every pair of the near list, the self pair contributing zero; excluded near pairs are
    // taken out afterwards by k_fmm_excl_sub with the same arithmetic
    if (DAMP && sizeof(R) == 4 && q0 >= qf) {  // far stage (uniform): undamped pairs, branch-free
      if (CHT && T == CHT && cn == 64) {
#pragma unroll
        for (int k = 0; k < CHT; ++k)
          staged_pair<R, DAMP, B2, K, true>(xi, yi, zi, di, sr[sl + k * (64 / (CHT ? CHT : 1))], f);
      } else if (act) {
        for (int u = sl; u < cn; u += S) staged_pair<R, DAMP, B2, K, true>(xi, yi, zi, di, sr[u], f);
      }
    } else if (CHT && T == CHT && cn == 64) {  // uniform: all 64 lanes active, 64 / CHT slots
      if constexpr (B2 || !DAMP) {  // branch-free pair: straight-line code, registers stay low
#pragma unroll
        for (int k = 0; k < CHT; ++k)
          staged_pair<R, DAMP, B2, K>(xi, yi, zi, di, sr[sl + k * (64 / (CHT ? CHT : 1))], f);
      } else {  // a branch per pair: unrolling would hoist the records into registers
#pragma unroll 1
        for (int k = 0; k < CHT; ++k)
          staged_pair<R, DAMP, B2, K>(xi, yi, zi, di, sr[sl + k * (64 / (CHT ? CHT : 1))], f);
      }
    } else if (act) {
      for (int u = sl; u < cn; u += S) staged_pair<R, DAMP, B2, K>(xi, yi, zi, di, sr[u], f);
    }
    wave_lds_barrier();
  }
  // slot sums through the (now unused) record LDS of this wavefront: red[c][lane]
  R* red = reinterpret_cast<R*>(sr);
#pragma unroll
  for (int c = 0; c < 3 * K; ++c) red[c * 64 + lane] = act ? f[c] : R(0);
  wave_lds_barrier();
  if (lane >= T) return;
  const double* ai = a.ainv;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double F[3] = {0.0, 0.0, 0.0};
    for (int s2 = 0; s2 < S; ++s2)
#pragma unroll
      for (int c = 0; c < 3; ++c) F[c] += (double)red[(3 * k + c) * 64 + lane + T * s2];
    if constexpr (WITH_L2P) {
      R lx = R(0), ly = R(0), lz = R(0);
      O::l2p((R)(pi.x - cax), (R)(pi.y - cay), (R)(pi.z - caz), &L[k * lstride + (size_t)A * O::LS], lx, ly, lz);
      F[0] += (double)lx;
      F[1] += (double)ly;
      F[2] += (double)lz;
    } else {
      (void)L;
      (void)lstride;
    }
    const size_t o = 3 * (size_t)n * k;
    if (MODE == MODE_BAPPLY) {  // alpha^-1 loaded after the L2P: fewer registers live across it
      double a6[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) a6[c] = ai[(size_t)c * n + i];
      const double m0 = a.src[o + i], m1 = a.src[o + n + i], m2 = a.src[o + 2 * (size_t)n + i];
      F[0] += a6[0] * m0 + a6[1] * m1 + a6[2] * m2;
      F[1] += a6[1] * m0 + a6[3] * m1 + a6[4] * m2;
      F[2] += a6[2] * m0 + a6[4] * m1 + a6[5] * m2;
    }
    a.out[o + i]                 = F[0];
    a.out[o + n + i]             = F[1];
    a.out[o + 2 * (size_t)n + i] = F[2];
  }
}

// FP32 source records: offset of each site from its leaf centre, and alpha_iso^(-1/6).
static __global__ void k_pack_posf4(int n, const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const double* __restrict__ s,
                                    const int* __restrict__ leaf_of, const double* __restrict__ cx,
                                    const double* __restrict__ cy, const double* __restrict__ cz,
                                    float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = leaf_of[i];
  out[i] = make_float4((float)(x[i] - cx[c]), (float)(y[i] - cy[c]), (float)(z[i] - cz[c]), s ? (float)s[i] : 0.f);
}

static __global__ void k_pack_pos4(int n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const double* __restrict__ s, double4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double4(x[i], y[i], z[i], s ? s[i] : 0.0);
}

template <typename MU4>
__global__ void k_pack_mu4(int n, const double* __restrict__ mu, MU4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  MU4 v;
  v.x    = mu[i];
  v.y    = mu[n + i];
  v.z    = mu[2 * (size_t)n + i];
  v.w    = 0;
  out[i] = v;
}

// Whether exclusion partner k of site i is a pair the near field adds: a valid index other than
// i, its first occurrence in i's list, and its leaf in the near list of i's leaf A (o0, o1: A's
// list, sorted by source leaf; otherwise a far-field pair, which stays in).
__device__ __forceinline__ bool excl_in_near(const NearArgs& a, int i, int k0, int k, int o0, int o1) {
  const int e = a.eidx[k];
  if (e < 0 || e >= a.n || e == i) return false;
  for (int q = k0; q < k; ++q)
    if (a.eidx[q] == e) return false;
  const int B = a.leaf_of[e];
  int lo = o0, hi = o1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a.p2p_src[mid] < B)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < o1 && a.p2p_src[lo] == B;
}

// The partners excl_in_near accepts, per site in list order: their number into out[i] (off
// null), or the partners themselves into out[off[i] ...].
static __global__ __launch_bounds__(256) void k_excl_near(NearArgs a, const int* __restrict__ off,
                                                         int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int k0 = a.eoff[i], k1 = a.eoff[i + 1];
  const int A = a.leaf_of[i], o0 = a.p2p_off[A], o1 = a.p2p_off[A + 1];
  int m = 0;
  for (int k = k0; k < k1; ++k) {
    if (!excl_in_near(a, i, k0, k, o0, o1)) continue;
    if (off) out[off[i] + m] = a.eidx[k];
    ++m;
  }
  if (!off) out[i] = m;
}

// Takes the excluded near-field pairs out of a near-field sum that included them: for each site
// i and each distinct exclusion partner e != i whose leaf is in the near list of i's leaf, out[i]
// -= pair(i, e), evaluated as the staged kernel does (same record, same arithmetic; FP32: offsets
// from i's leaf centre). K right-hand sides as in the staged kernel. LISTED: a.eoff / a.eidx hold
// those partners already (FmmState::xn_off / xn_idx), so no partner is tested here.
template <typename R, bool DAMP, typename MU4, bool B2 = false, int K = 1, bool LISTED = false>
__global__ __launch_bounds__(256) void k_fmm_excl_sub(NearArgs a, const MU4* __restrict__ mu4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int k0 = a.eoff[i], k1 = a.eoff[i + 1];
  if (k0 == k1) return;
  const int n      = a.n;
  const int A      = a.leaf_of[i];
  const int o0 = LISTED ? 0 : a.p2p_off[A], o1 = LISTED ? 0 : a.p2p_off[A + 1];
  const double4 pi = a.pos4[i];
  const double cax = a.cx[A], cay = a.cy[A], caz = a.cz[A];
  const R xi = sizeof(R) == 8 ? (R)pi.x : (R)(pi.x - cax);
  const R yi = sizeof(R) == 8 ? (R)pi.y : (R)(pi.y - cay);
  const R zi = sizeof(R) == 8 ? (R)pi.z : (R)(pi.z - caz);
  const R di = DAMP ? staged_di<R, B2>(a.damping, pi.w) : R(0);
  R f[3 * K];
#pragma unroll
  for (int c = 0; c < 3 * K; ++c) f[c] = R(0);
  for (int k = k0; k < k1; ++k) {
    if (!LISTED && !excl_in_near(a, i, k0, k, o0, o1)) continue;
    const int e      = a.eidx[k];
    const double4 pj = a.pos4[e];
    StagedRec<R, K> r;
    r.x = sizeof(R) == 8 ? (R)pj.x : (R)(pj.x - cax);
    r.y = sizeof(R) == 8 ? (R)pj.y : (R)(pj.y - cay);
    r.z = sizeof(R) == 8 ? (R)pj.z : (R)(pj.z - caz);
    r.s = (R)pj.w;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const MU4 mj   = mu4[(size_t)q * n + e];
      r.m[3 * q]     = (R)mj.x;
      r.m[3 * q + 1] = (R)mj.y;
      r.m[3 * q + 2] = (R)mj.z;
    }
    staged_pair<R, DAMP, (sizeof(R) == 4 && B2), K>(xi, yi, zi, di, r, f);
  }
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const size_t o = 3 * (size_t)n * q;
    a.out[o + i] -= (double)f[3 * q];
    a.out[o + n + i] -= (double)f[3 * q + 1];
    a.out[o + 2 * (size_t)n + i] -= (double)f[3 * q + 2];
  }
}

// out += L2P of each site's leaf expansion (the far-field term when it runs beside the near field)
// (grid.y = right-hand side: L + y lstride, out + 3 y n)
// With d.p set (one right-hand side): each block also writes the partials p.out and out.out of
// its 256 sites -- the CG step's p.Ap and Ap.Ap, summed as the CG's own chunk kernel sums them.
struct L2pDots {
  const double* p = nullptr;
  double* part    = nullptr;
  int nchunk      = 0;
};

template <int P, typename R>
__global__ __launch_bounds__(256) void k_fmm_l2p_add(int n, const int* __restrict__ leaf_of,
                                                     const double4* __restrict__ pos4, const double* __restrict__ cx,
                                                     const double* __restrict__ cy, const double* __restrict__ cz,
                                                     const R* __restrict__ L, double* __restrict__ out,
                                                     size_t lstride = 0, L2pDots d = {}) {
  using O     = Ops<KIND_M1, P, R>;
  const int b = xcd_chunk(blockIdx.x, gridDim.x);  // the CG chunk of the same 256 sites
  const int i = b * blockDim.x + threadIdx.x;
  L += blockIdx.y * lstride;
  out += 3 * (size_t)n * blockIdx.y;
  double s = 0.0, q = 0.0;
  if (i < n) {
    const int A      = leaf_of[i];
    const double4 pi = pos4[i];
    R lx = R(0), ly = R(0), lz = R(0);
    O::l2p((R)(pi.x - cx[A]), (R)(pi.y - cy[A]), (R)(pi.z - cz[A]), &L[(size_t)A * O::LS], lx, ly, lz);
    const size_t k1 = n + (size_t)i, k2 = k1 + n;
    const double a0 = out[i] + (double)lx, a1 = out[k1] + (double)ly, a2 = out[k2] + (double)lz;
    out[i]  = a0;
    out[k1] = a1;
    out[k2] = a2;
    if (d.p) {
      s = d.p[i] * a0 + d.p[k1] * a1 + d.p[k2] * a2;
      q = a0 * a0 + a1 * a1 + a2 * a2;
    }
  }
  if (d.p) {
    __shared__ double red[4];
    s = block_sum(s, red);
    q = block_sum(q, red);
    if (threadIdx.x == 0) {
      d.part[2 * (size_t)d.nchunk + b] = s;
      d.part[3 * (size_t)d.nchunk + b] = q;
    }
  }
}

// The L2P pass of a single right-hand side's apply into out, with the CG step's partials when
// the step armed them for this output (System::dots).
template <int P, typename R>
void launch_l2p_add(System& S, const R* L, double* out, hipStream_t st) {
  const Octree& T = S.tree;
  L2pDots d;
  if (S.dots.p && S.dots.ap == out && S.dots.nchunk == ceil_div(S.n, 256)) {
    d = L2pDots{S.dots.p, S.dots.part, S.dots.nchunk};
    S.dots.done = true;
  }
  hipLaunchKernelGGL((k_fmm_l2p_add<P, R>), dim3(ceil_div(S.n, 256)), dim3(256), 0, st, S.n, T.leaf_of.p,
                     S.fmm->pos4.p, T.cx.p, T.cy.p, T.cz.p, L, out, (size_t)0, d);
}

#include "near_sym.hh"

// Target records of the symmetric near field (see FmmState::sym_t32 / sym_t64).
static __global__ void k_sym_trec32(int n, const double4* __restrict__ pos4, const int* __restrict__ leaf_of,
                                    const double* __restrict__ cx, const double* __restrict__ cy,
                                    const double* __restrict__ cz, double damping, int b2, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c      = leaf_of[i];
  const double4 p  = pos4[i];
  const float di   = b2 ? staged_di<float, true>(damping, p.w) : staged_di<float, false>(damping, p.w);
  out[i]           = make_float4((float)(p.x - cx[c]), (float)(p.y - cy[c]), (float)(p.z - cz[c]), di);
}

static __global__ void k_sym_trec64(int n, const double4* __restrict__ pos4, double damping,
                                    double4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 p = pos4[i];
  out[i]          = make_double4(p.x, p.y, p.z, staged_di<double, false>(damping, p.w));
}

// ---- host drivers ----------------------------------------------------------------------------
// FP32 M2L pair records (target-grouped list order): the displacement is formed in FP64 and
// rounded, as the FP64-centre form of the kernel does, so both give the same FP32 operands.
static __global__ void k_m2l_rec(int ncells, const int* __restrict__ off, const int* __restrict__ srcs,
                                 const double* __restrict__ cx, const double* __restrict__ cy,
                                 const double* __restrict__ cz, float4* __restrict__ rec) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= ncells) return;
  const double ax = cx[A], ay = cy[A], az = cz[A];
  for (int k = off[A]; k < off[A + 1]; ++k) {
    const int B = srcs[k];
    rec[k]      = make_float4((float)(ax - cx[B]), (float)(ay - cy[B]), (float)(az - cz[B]), __int_as_float(B));
  }
}

// The records are filled on stream st: the stream of the M2L that reads them (the far field may
// run on a second stream beside the near field).
inline void ensure_m2l_rec(System& S, hipStream_t st) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  if (F.m2l_rec.p && F.m2l_rec.n >= (size_t)std::max<long long>(T.n_m2l, 1)) return;
  F.m2l_rec.alloc((size_t)std::max<long long>(T.n_m2l, 1));
  hipLaunchKernelGGL(k_m2l_rec, dim3(ceil_div(T.ncells, 256)), dim3(256), 0, st, T.ncells, T.m2l_off.p,
                     T.m2l_src.p, T.cx.p, T.cy.p, T.cz.p, F.m2l_rec.p);
}

// Upward half of the far field: P2M, M2M level by level, fold into the M2L rows.
template <int KIND, int P, typename R>
void fmm_upward_pass(System& S, const double* src, R* M, hipStream_t st) {
  const Octree& T = S.tree;
  const auto& lo = T.level_offset;
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[0], st));
  constexpr int GP = 16;  // lanes per leaf in P2M
  hipLaunchKernelGGL((k_fmm_p2m<KIND, P, R, GP>), dim3(ceil_div(T.nleaves, 64 / GP)), dim3(64), 0, st, T.nleaves,
                     T.leaf_ids.p, T.level.p, T.start.p, T.count.p, T.cx.p, T.cy.p, T.cz.p, S.x.p, S.y.p, S.z.p, src, S.n, M);
  // roots (level 0) are skipped: their expansions never enter an M2L pair
  for (int lev = T.max_level - 1; lev >= 1; --lev) {
    const int c0 = lo[lev], c1 = lo[lev + 1];
    hipLaunchKernelGGL((k_fmm_m2m<KIND, P, R>), dim3(ceil_div(c1 - c0, 8)), dim3(64), 0, st, c0, c1,
                       T.nchild.p, T.child_first.p, T.cx.p, T.cy.p, T.cz.p, M);
  }
  R* Mf = M + fmm_fold_offset(T.ncells, KIND, P);
  if (T.ncells > lo[1])
    hipLaunchKernelGGL((k_fmm_fold<KIND, P, R>), dim3(ceil_div(T.ncells - lo[1], 64)), dim3(64), 0, st, lo[1],
                       T.ncells, M, Mf);
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[1], st));
}

// Downward half: M2L from the folded rows, unfold, L2L level by level.
template <int KIND, int P, typename R>
void fmm_down_pass(System& S, R* M, R* L, hipStream_t st) {
  using O         = Ops<KIND, P, R>;
  const Octree& T = S.tree;
  const auto& lo = T.level_offset;
  R* Mf = M + fmm_fold_offset(T.ncells, KIND, P);
  const int grp = S.fmm ? S.fmm->tune.m2l_group[sizeof(R) == 4] : 64;
  const float4* rec = nullptr;
  if constexpr (sizeof(R) == 4) {
    if (!S.fmm) throw std::runtime_error("FP32 far field without FMM state");
    ensure_m2l_rec(S, st);  // allocated beforehand by fmm_reserve when the apply is captured
    rec = S.fmm->m2l_rec.p;
  }
  if constexpr (KIND == KIND_M1) {
    if (grp == 16)
      hipLaunchKernelGGL((k_fmm_m2l<KIND, P, R, 16>), dim3(ceil_div(T.ncells, 4)), dim3(64), 0, st, T.ncells,
                         T.m2l_off.p, T.m2l_src.p, T.cx.p, T.cy.p, T.cz.p, Mf, L, rec);
    else if (grp == 32)
      hipLaunchKernelGGL((k_fmm_m2l<KIND, P, R, 32>), dim3(ceil_div(T.ncells, 2)), dim3(64), 0, st, T.ncells,
                         T.m2l_off.p, T.m2l_src.p, T.cx.p, T.cy.p, T.cz.p, Mf, L, rec);
    else
      hipLaunchKernelGGL((k_fmm_m2l<KIND, P, R, 64>), dim3(T.ncells), dim3(64), 0, st, T.ncells, T.m2l_off.p,
                         T.m2l_src.p, T.cx.p, T.cy.p, T.cz.p, Mf, L, rec);
  } else {
    (void)grp;
    hipLaunchKernelGGL((k_fmm_m2l<KIND, P, R, 64>), dim3(T.ncells), dim3(64), 0, st, T.ncells, T.m2l_off.p,
                       T.m2l_src.p, T.cx.p, T.cy.p, T.cz.p, Mf, L, rec);
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[2], st));
  // unfold + L2L: the top levels (at most kTopCells cells each) in one workgroup, then one launch
  // per level
  TopLevels top{0, {}};
  top.off[0] = lo[0];
  while (top.nlev < kTopLevels && top.nlev <= T.max_level && lo[top.nlev + 1] - lo[top.nlev] <= kTopCells) {
    top.off[top.nlev + 1] = lo[top.nlev + 1];
    ++top.nlev;
  }
  if (top.nlev > 0)
    hipLaunchKernelGGL((k_fmm_down_top<KIND, P, R>), dim3(1), dim3(down_top_threads<KIND>()), 0, st, top, T.parent.p,
                       T.cx.p, T.cy.p, T.cz.p, L);
  for (int lev = top.nlev; lev <= T.max_level; ++lev) {
    const int c0 = lo[lev], c1 = lo[lev + 1];
    if (lev == 0) {  // more roots than kTopCells (a batch of that many snapshots): unfold only
      hipLaunchKernelGGL((k_fmm_unfold<KIND, P, R>), dim3(ceil_div(c1 - c0, 64)), dim3(64), 0, st, c1, L);
      continue;
    }
    hipLaunchKernelGGL((k_fmm_l2l<KIND, P, R>), dim3(ceil_div(c1 - c0, 64)), dim3(64), 0, st, c0, c1,
                       T.parent.p, T.cx.p, T.cy.p, T.cz.p, L);
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[3], st));
  (void)sizeof(O);
}

template <int KIND, int P, typename R>
void fmm_far_pass(System& S, const double* src, R* M, R* L, hipStream_t st) {
  fmm_upward_pass<KIND, P, R>(S, src, M, st);
  fmm_down_pass<KIND, P, R>(S, M, L, st);
}

inline P2PArgs make_p2p_args(System& S, const double* src, double damping, double* out) {
  const Octree& T = S.tree;
  P2PArgs a;
  a.leaf_ids = T.leaf_ids.p;
  a.start    = T.start.p;
  a.count    = T.count.p;
  a.p2p_off  = T.p2p_off.p;
  a.p2p_src  = T.p2p_src.p;
  a.cx = T.cx.p;
  a.cy = T.cy.p;
  a.cz = T.cz.p;
  a.x = S.x.p;
  a.y = S.y.p;
  a.z = S.z.p;
  a.s       = S.s.p;
  a.damping = damping;
  a.eoff    = S.excl_off.p;
  a.eidx    = S.excl_idx.p;
  a.src     = src;
  a.ainv    = S.ainv6.p;
  a.n       = S.n;
  a.out     = out;
  return a;
}

template <int KIND, int P, typename R, int MODE>
void fmm_p2p_pass(System& S, const double* src, double damping, const R* L, double* out) {
  P2PArgs a      = make_p2p_args(S, src, damping, out);
  const int nl   = S.tree.nleaves;
  hipStream_t st = S.stream;
  if (damping > 0.0)
    hipLaunchKernelGGL((k_fmm_p2p<KIND, P, R, true, MODE>), dim3(nl), dim3(64), 0, st, a, L);
  else
    hipLaunchKernelGGL((k_fmm_p2p<KIND, P, R, false, MODE>), dim3(nl), dim3(64), 0, st, a, L);
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[4], st));
}

template <typename T>
void ensure_buf(dbuf<T>& b, size_t count) {
  if (b.n < count || !b.p) b.alloc(count);
}

inline void ensure_pos4(System& S) {
  FmmState& F = *S.fmm;
  if (F.pos4.p && F.pos4.n >= (size_t)S.n) return;
  F.pos4.alloc(S.n);
  hipLaunchKernelGGL(k_pack_pos4, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, S.x.p, S.y.p, S.z.p, S.s.p,
                     F.pos4.p);
}

// Work items of the leaf-chunk near kernels: (leaf, first target) per chunk of at most C targets of
// every leaf (C = 64 >> ci, ci = 0..3). Returns the item count.
inline int ensure_near_items(System& S, int ci = 0) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  if (F.near_items[ci].p && F.n_items[ci] > 0) return F.n_items[ci];
  const int C = 64 >> ci;
  std::vector<int> ids(T.nleaves), cnt(T.ncells);
  CPPE_HIP_CHECK(hipMemcpyAsync(ids.data(), T.leaf_ids.p, sizeof(int) * T.nleaves, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(cnt.data(), T.count.p, sizeof(int) * T.ncells, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  std::vector<int2> items;
  items.reserve(T.nleaves);
  for (int c : ids)
    for (int t0 = 0; t0 < cnt[c]; t0 += C) items.push_back(make_int2(c, t0));
  F.near_items[ci].alloc(items.size());
  CPPE_HIP_CHECK(hipMemcpyAsync(F.near_items[ci].p, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice,
                                S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  F.n_items[ci] = (int)items.size();
  return F.n_items[ci];
}

inline void ensure_posf4(System& S) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  if (F.posf4.p && F.posf4.n >= (size_t)S.n) return;
  F.posf4.alloc(S.n);
  hipLaunchKernelGGL(k_pack_posf4, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, S.x.p, S.y.p, S.z.p, S.s.p,
                     T.leaf_of.p, T.cx.p, T.cy.p, T.cz.p, F.posf4.p);
}

// The exclusion partners whose pairs the near field adds (k_excl_near), per site, once per tree;
// nothing when the system has no exclusions. Allocates: the CG reserves it before any capture.
inline void ensure_excl_near(System& S) {
  FmmState& F = *S.fmm;
  if (F.xn_built || !S.excl_off.p) return;
  const Octree& T = S.tree;
  NearArgs a{};
  a.leaf_of = T.leaf_of.p;
  a.p2p_off = T.p2p_off.p;
  a.p2p_src = T.p2p_src.p;
  a.eoff    = S.excl_off.p;
  a.eidx    = S.excl_idx.p;
  a.n       = S.n;
  const dim3 g(ceil_div(S.n, 256)), b(256);
  F.xn_off.alloc((size_t)S.n + 1);
  hipLaunchKernelGGL(k_excl_near, g, b, 0, S.stream, a, nullptr, F.xn_off.p);
  std::vector<int> off((size_t)S.n + 1, 0);
  CPPE_HIP_CHECK(hipMemcpyAsync(off.data(), F.xn_off.p, sizeof(int) * S.n, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  int tot = 0;
  for (int i = 0; i < S.n; ++i) {
    const int c = off[i];
    off[i]      = tot;
    tot += c;
  }
  off[S.n] = tot;
  CPPE_HIP_CHECK(hipMemcpyAsync(F.xn_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, S.stream));
  F.xn_idx.alloc((size_t)std::max(tot, 1));
  hipLaunchKernelGGL(k_excl_near, g, b, 0, S.stream, a, F.xn_off.p, F.xn_idx.p);
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  F.xn_built = true;
}

// k_fmm_excl_sub over the listed partners (ensure_excl_near)
template <typename R, bool DAMP, typename MU4, int K>
void launch_excl_sub(System& S, const NearArgs& a, const MU4* mu4, bool b2, hipStream_t st) {
  if (!a.eoff) return;
  ensure_excl_near(S);
  NearArgs ax             = a;
  ax.eoff                 = S.fmm->xn_off.p;
  ax.eidx                 = S.fmm->xn_idx.p;
  constexpr bool kB2      = sizeof(R) == 4 && DAMP;
  const dim3 g(ceil_div(S.n, 256)), b(256);
  if (b2)
    hipLaunchKernelGGL((k_fmm_excl_sub<R, DAMP, MU4, kB2, K, true>), g, b, 0, st, ax, mu4);
  else
    hipLaunchKernelGGL((k_fmm_excl_sub<R, DAMP, MU4, false, K, true>), g, b, 0, st, ax, mu4);
}

void ensure_clusters(System& S);  // fmm.hip

// targets per work item of the symmetric near-field variants 9, 10, 11: FP64 8, 6, 4 and FP32 16,
// 12, 8 (larger chunks receive fewer reverse blocks; smaller ones leave registers for more
// wavefronts per SIMD)
constexpr int kSymTs[2][3] = {{8, 6, 4}, {16, 12, 8}};
inline bool near_sym_variant(int v) { return v >= 9 && v <= 11; }

// Target records of the symmetric forms for this damping factor (precision prec: 0 FP64, 1 FP32).
inline void ensure_sym_trec(System& S, int prec, double damping) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  ensure_pos4(S);
  const dim3 g(ceil_div(S.n, 256)), b(256);
  if (prec == 1) {
    const bool b2 = S.all_damped && damping > 0.0;
    if (!F.sym_t32.p || F.sym_tdamp[1] != damping || F.sym_tb2 != b2) {
      F.sym_t32.alloc(S.n);
      hipLaunchKernelGGL(k_sym_trec32, g, b, 0, S.stream, S.n, F.pos4.p, T.leaf_of.p, T.cx.p, T.cy.p, T.cz.p,
                         damping, b2 ? 1 : 0, F.sym_t32.p);
      F.sym_tdamp[1] = damping;
      F.sym_tb2      = b2;
    }
  } else if (!F.sym_t64.p || F.sym_tdamp[0] != damping) {
    F.sym_t64.alloc(S.n);
    hipLaunchKernelGGL(k_sym_trec64, g, b, 0, S.stream, S.n, F.pos4.p, damping, F.sym_t64.p);
    F.sym_tdamp[0] = damping;
  }
}

// Lists (chunks of ts targets), target records (for this damping factor) and reverse slots of the
// symmetric near field in precision prec (0 FP64, 1 FP32).
inline SymLists& ensure_sym(System& S, int prec, int ts, double damping) {
  FmmState& F = *S.fmm;
  SymLists& Y = build_sym_lists(S, ts);
  ensure_sym_trec(S, prec, damping);
  if (prec)
    ensure_buf(F.rev32, (size_t)std::max<long long>(Y.nrev, 1));
  else
    ensure_buf(F.rev64, (size_t)std::max<long long>(Y.nrev, 1));
  return Y;
}

// out += the reverse blocks of each site's leaf (and its L2P term when L is given).
template <int P, typename R>
void launch_sym_gather(System& S, const SymLists& Y, const R* L, double* out) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  const R* rev;
  if constexpr (sizeof(R) == 4)
    rev = F.rev32.p;
  else
    rev = F.rev64.p;
  hipLaunchKernelGGL((k_fmm_sym_gather<P, R>), dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, T.leaf_of.p,
                     T.start.p, T.count.p, Y.rev_off.p, Y.nin.p, rev, F.pos4.p, T.cx.p, T.cy.p, T.cz.p, L, out);
}

// wavefronts per block of the symmetric direct-summation kernel
constexpr int kDSymW = 8;
// sources per item of the symmetric direct-summation kernel (rounded to whole tiles)
constexpr int kDSymSeg = 8192;

// Items and blocks of the symmetric direct summation with TS targets per wavefront (systems whose
// tree is one leaf); built on the host once per chunk size.
inline FmmState::DSym& ensure_dsym(System& S, int prec, int ts, double damping) {
  FmmState& F       = *S.fmm;
  FmmState::DSym& D = F.dsym[prec];
  ensure_sym_trec(S, prec, damping);
  if (D.ts != ts) {
    const int n = S.n, TB = kDSymW * ts;
    const int ntile  = (n + TB - 1) / TB;
    const long long seg = (long long)TB * std::max(1, kDSymSeg / TB);
    std::vector<int4> items;
    std::vector<long long> off(ntile + 1, 0);
    for (int I = 0; I < ntile; ++I) {
      const long long e = std::min<long long>(n, (long long)(I + 1) * TB);
      items.push_back(make_int4(I, I * TB, (int)e, 0));
      int k = 1;
      for (long long s0 = (long long)(I + 1) * TB; s0 < n; s0 += seg, ++k)
        items.push_back(make_int4(I, (int)s0, (int)std::min<long long>(n, s0 + seg), k));
      off[I + 1] = off[I] + 3ll * TB * k;
    }
    // longest items first (the tile's own sites are short)
    std::stable_sort(items.begin(), items.end(), [](const int4& x, const int4& y) { return x.z - x.y > y.z - y.y; });
    D.items.alloc(items.size());
    D.fwd_off.alloc(off.size());
    CPPE_HIP_CHECK(hipMemcpyAsync(D.items.p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, S.stream));
    CPPE_HIP_CHECK(hipMemcpyAsync(D.fwd_off.p, off.data(), sizeof(long long) * off.size(), hipMemcpyHostToDevice,
                                  S.stream));
    const Octree& T = S.tree;
    CPPE_HIP_CHECK(hipMemcpyAsync(&D.c[0], T.cx.p, sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CPPE_HIP_CHECK(hipMemcpyAsync(&D.c[1], T.cy.p, sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CPPE_HIP_CHECK(hipMemcpyAsync(&D.c[2], T.cz.p, sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
    D.ts     = ts;
    D.tb     = TB;
    D.ntile  = ntile;
    D.nitems = (int)items.size();
    D.nfwd   = off[ntile];
    D.nrev   = 3ll * TB * ((long long)ntile * (ntile - 1) / 2);
  }
  const size_t nf = (size_t)std::max<long long>(D.nfwd, 1), nr = (size_t)std::max<long long>(D.nrev, 1);
  if (prec) {
    ensure_buf(F.dfwd32, nf);
    ensure_buf(F.drev32, nr);
  } else {
    ensure_buf(F.dfwd64, nf);
    ensure_buf(F.drev64, nr);
  }
  return D;
}

// a system whose tree is one leaf holding every site (direct summation)
inline bool one_leaf_system(const System& S) { return S.nseg == 1 && S.tree.nleaves == 1; }

// Symmetric direct summation of a B apply: the tile kernel, the gather (which writes out), the
// exclusion pass, then -- when the far field has run before it (with_l2p) -- the (empty) L2P pass.
template <int P, typename R, bool DAMP, typename MU4, int TS>
void launch_direct_sym(System& S, const NearArgs& a, const MU4* mu4, const R* L, bool with_l2p) {
  FmmState& F        = *S.fmm;
  const Octree& T    = S.tree;
  constexpr int prec = sizeof(R) == 4;
  FmmState::DSym& D  = ensure_dsym(S, prec, TS, a.damping);
  constexpr bool kB2 = prec && DAMP;
  const bool b2      = kB2 && S.all_damped;
  hipStream_t st     = S.stream;
  const DSymArgs da{D.items.p, D.nitems, S.n, D.fwd_off.p, D.c[0], D.c[1], D.c[2]};
  R *fwd, *rev;
  if constexpr (prec) {
    fwd = F.dfwd32.p;
    rev = F.drev32.p;
  } else {
    fwd = F.dfwd64.p;
    rev = F.drev64.p;
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[6], st));
  const dim3 g(D.nitems), b(64 * kDSymW);
  if constexpr (prec) {
    if (b2)
      hipLaunchKernelGGL((k_direct_sym<R, DAMP, kB2, TS, kDSymW, MU4, float4>), g, b, 0, st, a, da, F.sym_t32.p, mu4,
                         fwd, rev);
    else
      hipLaunchKernelGGL((k_direct_sym<R, DAMP, false, TS, kDSymW, MU4, float4>), g, b, 0, st, a, da, F.sym_t32.p,
                         mu4, fwd, rev);
  } else {
    hipLaunchKernelGGL((k_direct_sym<R, DAMP, false, TS, kDSymW, MU4, double4>), g, b, 0, st, a, da, F.sym_t64.p,
                       mu4, fwd, rev);
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[7], st));
  hipLaunchKernelGGL(k_direct_sym_gather<R>, dim3(ceil_div(S.n, 256)), dim3(256), 0, st, a, D.tb, D.fwd_off.p, fwd,
                     rev);
  launch_excl_sub<R, DAMP, MU4, 1>(S, a, mu4, b2, st);
  if (with_l2p) launch_l2p_add<P, R>(S, L, a.out, st);
}

// Symmetric near field of a B apply (variants 9-11, TS targets per chunk): the pair kernel, the
// exclusion pass and the gather of the reverse blocks (beside the far field when it runs on the
// second stream), then -- when the far field has run before it (with_l2p) -- the L2P pass;
// otherwise the caller adds L2P after joining the far field.
template <int P, typename R, bool DAMP, typename MU4, int TS>
void launch_near_sym(System& S, const NearArgs& a, const MU4* mu4, const R* L, bool with_l2p) {
  FmmState& F        = *S.fmm;
  const Octree& T    = S.tree;
  constexpr int prec = sizeof(R) == 4;
  SymLists& Y        = ensure_sym(S, prec, TS, a.damping);
  constexpr bool kB2 = prec && DAMP;
  const bool b2      = kB2 && S.all_damped;
  const long long* usp = (DAMP && prec && a.damping >= T.xdamp) ? Y.usplit.p : nullptr;
  const int block    = F.tune.near_block[prec];
  const int splits   = with_l2p ? 1 : std::max(1, F.tune.near_splits[prec]);
  hipStream_t st     = S.stream;
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[6], st));
  for (int k = 0; k < splits; ++k) {
    const int i0 = (int)((long long)Y.nitems * k / splits), i1 = (int)((long long)Y.nitems * (k + 1) / splits);
    if (i1 <= i0) continue;
    const SymArgs sa{Y.items.p + i0, i1 - i0, Y.uoff.p, usp, Y.usrc.p, Y.udst.p};
    const dim3 g(ceil_div(i1 - i0, block / 64)), b(block);
    if constexpr (prec) {
      if (b2)
        hipLaunchKernelGGL((k_fmm_near_sym<R, DAMP, kB2, TS, MU4, float4>), g, b, 0, st, a, sa, F.sym_t32.p, mu4,
                           F.rev32.p);
      else
        hipLaunchKernelGGL((k_fmm_near_sym<R, DAMP, false, TS, MU4, float4>), g, b, 0, st, a, sa, F.sym_t32.p, mu4,
                           F.rev32.p);
    } else {
      hipLaunchKernelGGL((k_fmm_near_sym<R, DAMP, false, TS, MU4, double4>), g, b, 0, st, a, sa, F.sym_t64.p, mu4,
                         F.rev64.p);
    }
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[7], st));
  launch_excl_sub<R, DAMP, MU4, 1>(S, a, mu4, b2, st);
  launch_sym_gather<P, R>(S, Y, nullptr, a.out);
  if (with_l2p) launch_l2p_add<P, R>(S, L, a.out, st);
}

template <int P, typename R, bool DAMP, typename MU4>
void launch_near_sym_v(System& S, int var, const NearArgs& a, const MU4* mu4, const R* L, bool with_l2p) {
  constexpr int prec = sizeof(R) == 4;
  const bool direct  = one_leaf_system(S);
  auto go = [&](auto ts) {
    constexpr int TS = decltype(ts)::value;
    if (direct)
      launch_direct_sym<P, R, DAMP, MU4, TS>(S, a, mu4, L, with_l2p);
    else
      launch_near_sym<P, R, DAMP, MU4, TS>(S, a, mu4, L, with_l2p);
  };
  if (var == 9)
    go(std::integral_constant<int, kSymTs[prec][0]>{});
  else if (var == 10)
    go(std::integral_constant<int, kSymTs[prec][1]>{});
  else
    go(std::integral_constant<int, kSymTs[prec][2]>{});
}


template <int P, typename R, int MODE, bool DAMP, typename MU4, int K = 1>
void launch_near(System& S, const NearArgs& a, const MU4* mu4, const R* L, bool with_l2p, size_t lstride = 0) {
  FmmState& F     = *S.fmm;
  const int prec  = sizeof(R) == 4;
  const int block = F.tune.near_block[prec];
  int var         = F.tune.near_variant[prec];
  hipStream_t st  = S.stream;
  if (near_sym_variant(var)) {
    if constexpr (K == 1 && MODE == MODE_BAPPLY) {
      launch_near_sym_v<P, R, DAMP, MU4>(S, var, a, mu4, L, with_l2p);
      return;
    }
    var = 8;
  }
  if (var >= 5) {  // 5, 6, 7, 8: staged, chunks of 64, 32, 16, 8 targets
    const int ci = var - 5;
    const int ni = ensure_near_items(S, ci);
    build_near_expanded(S);
    const Octree& T = S.tree;
    const dim3 b(block);
    const int chunk = 64 >> ci;
    // the branch-free FP32 pair needs every site damped
    constexpr bool kB2 = sizeof(R) == 4 && DAMP;
    const bool b2      = kB2 && S.all_damped;
    // far-source stages: valid for damping factors at least the classification's
    const long long* xs = (DAMP && a.damping >= T.xdamp) ? T.xsplit[prec].p : nullptr;
    // Beside a far field on the second stream (no fused L2P), the work items may be issued in
    // several launches: the device dispatches the far field's small dependent kernels between
    // them instead of behind the whole near field.
    const int splits = with_l2p ? 1 : std::max(1, F.tune.near_splits[prec]);
    auto go = [&](auto l2p, auto cht, auto bb) {
      for (int k = 0; k < splits; ++k) {
        const int i0 = (int)((long long)ni * k / splits), i1 = (int)((long long)ni * (k + 1) / splits);
        if (i1 <= i0) continue;
        hipLaunchKernelGGL((k_fmm_near_st<P, R, DAMP, MODE, MU4, decltype(l2p)::value, decltype(cht)::value,
                                          decltype(bb)::value, K>),
                           dim3(ceil_div(i1 - i0, block / 64)), b, 0, st, a, F.near_items[ci].p + i0, i1 - i0, chunk,
                           T.xoff.p, T.xsrc.p, mu4, L, lstride, xs);
      }
    };
    auto go2 = [&](auto l2p, auto bb) {
      // the unrolled full-chunk path pays for branch-free pairs only (a damped FP64 pair branches
      // around its exponential, and unrolling it costs occupancy)
      if constexpr (decltype(bb)::value || !DAMP) {
        if (chunk == 8) {
          go(l2p, std::integral_constant<int, 8>{}, bb);
          return;
        }
      }
      go(l2p, std::integral_constant<int, 0>{}, bb);
    };
    auto go1 = [&](auto bb) {
      if (with_l2p)
        go2(std::true_type{}, bb);
      else
        go2(std::false_type{}, bb);
    };
    if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[6], st));
    if (b2)
      go1(std::integral_constant<bool, kB2>{});
    else
      go1(std::false_type{});
    if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[7], st));
    launch_excl_sub<R, DAMP, MU4, K>(S, a, mu4, b2, st);
  } else if constexpr (K > 1) {
    throw hip_error("several right-hand sides per launch need a staged near-field variant");
  } else if (var == 4) {
    ensure_clusters(S);
    const dim3 g(ceil_div(F.n_clusters, block / 64)), b(block);
    if constexpr (sizeof(R) == 8) {
      hipLaunchKernelGGL((k_fmm_near_cl<P, R, DAMP, MODE, MU4, double4>), g, b, 0, st, a, F.n_clusters, F.cl_range.p,
                         F.cl_off.p, F.cl_src.p, F.cl_mask.p, F.pos4.p, mu4, L);
    } else {
      ensure_posf4(S);
      hipLaunchKernelGGL((k_fmm_near_cl<P, R, DAMP, MODE, MU4, float4>), g, b, 0, st, a, F.n_clusters, F.cl_range.p,
                         F.cl_off.p, F.cl_src.p, F.cl_mask.p, F.posf4.p, mu4, L);
    }
  } else if (var == 1 || var == 3) {
    const int ni    = ensure_near_items(S, 0);
    const int waves = block / 64;
    const dim3 g(ceil_div(ni, waves)), b(block);
    if (var == 3) {
      build_near_expanded(S);
      const Octree& T = S.tree;
      hipLaunchKernelGGL((k_fmm_near_sl<P, R, DAMP, MODE, MU4>), g, b, 0, st, a, F.near_items[0].p, ni,
                         T.xoff.p, T.xsrc.p, mu4, L);
    } else if constexpr (sizeof(R) == 8) {
      hipLaunchKernelGGL((k_fmm_near_lw<P, R, DAMP, MODE, MU4, double4>), g, b, 0, st, a, F.near_items[0].p, ni,
                         F.pos4.p, mu4, L);
    } else {
      ensure_posf4(S);
      hipLaunchKernelGGL((k_fmm_near_lw<P, R, DAMP, MODE, MU4, float4>), g, b, 0, st, a, F.near_items[0].p, ni,
                         F.posf4.p, mu4, L);
    }
  } else if (var == 2) {
    // LDS-staged kernel (one wavefront per leaf); its sources are the SoA dipoles
    P2PArgs pa = make_p2p_args(S, a.src, a.damping, a.out);
    hipLaunchKernelGGL((k_fmm_p2p<KIND_M1, P, R, DAMP, MODE>), dim3(S.tree.nleaves), dim3(64), 0, st, pa, L);
  } else {
    hipLaunchKernelGGL((k_fmm_near<P, R, DAMP, MODE, MU4>), dim3(ceil_div(S.n, block)), dim3(block), 0, st, a, mu4, L);
  }
}

template <int P, typename R, int MODE>
void fmm_near_pass(System& S, const double* mu, double damping, const R* L, double* out, bool with_l2p = true) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  hipStream_t st  = S.stream;
  ensure_pos4(S);
  NearArgs a{T.leaf_of.p, T.start.p, T.count.p, T.p2p_off.p, T.p2p_src.p, T.cx.p, T.cy.p, T.cz.p, F.pos4.p,
             S.excl_off.p, S.excl_idx.p, mu, S.ainv6.p, damping, S.n, out};
  const dim3 g(ceil_div(S.n, 256)), b(256);
  if constexpr (sizeof(R) == 8) {
    ensure_buf(F.mu4, S.n);
    if (mu != S.prepacked) hipLaunchKernelGGL(k_pack_mu4<double4>, g, b, 0, st, S.n, mu, F.mu4.p);
    if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[3], st));
    if (damping > 0.0)
      launch_near<P, R, MODE, true>(S, a, F.mu4.p, L, with_l2p);
    else
      launch_near<P, R, MODE, false>(S, a, F.mu4.p, L, with_l2p);
  } else {
    ensure_buf(F.mu4f, S.n);
    if (mu != S.prepacked) hipLaunchKernelGGL(k_pack_mu4<float4>, g, b, 0, st, S.n, mu, F.mu4f.p);
    if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[3], st));
    if (damping > 0.0)
      launch_near<P, R, MODE, true>(S, a, F.mu4f.p, L, with_l2p);
    else
      launch_near<P, R, MODE, false>(S, a, F.mu4f.p, L, with_l2p);
  }
  if (S.prof) CPPE_HIP_CHECK(hipEventRecord(S.prof[4], st));
}

// Far field then near field (fused L2P), or -- when tuned so and the near variant is a staged
// one -- the far field on a second stream beside the near field, joined before an L2P pass.
// Overlap 1: the whole far field beside the near field. The near-field kernels are enqueued
// before the far pass: the launch order carries into a captured graph, and the near field then
// starts together with the (latency-bound) upward pass instead of behind it. Overlap 2: the
// upward pass first, alone on the system's stream, then only M2L / L2L beside the near field --
// the near field's workgroups otherwise occupy the device while the upward pass's small launches
// wait for slots, and M2L then ends well after the near field. Overlap 3: as 1, but the far pass
// enqueued ahead of the near field, so that the upward pass's workgroups are dispatched first.
template <int P, typename R>
void fmm_apply_b_prec(System& S, const double* x, double* out, R* M, R* L) {
  FmmState& F      = *S.fmm;
  const int prec   = sizeof(R) == 4;
  const double dmp = S.opts.damping;
  const int ovl    = F.tune.near_variant[prec] >= 5 && !S.prof && F.far_stream ? F.tune.overlap[prec] : 0;
  if (!ovl) {
    fmm_far_pass<KIND_M1, P, R>(S, x, M, L, S.stream);
    fmm_near_pass<P, R, MODE_BAPPLY>(S, x, dmp, L, out);
    return;
  }
  if (ovl == 2) fmm_upward_pass<KIND_M1, P, R>(S, x, M, S.stream);
  CPPE_HIP_CHECK(hipEventRecord(F.ev_fork, S.stream));
  CPPE_HIP_CHECK(hipStreamWaitEvent(F.far_stream, F.ev_fork, 0));
  if (ovl == 3) fmm_far_pass<KIND_M1, P, R>(S, x, M, L, F.far_stream);
  fmm_near_pass<P, R, MODE_BAPPLY>(S, x, dmp, L, out, false);
  if (ovl == 2)
    fmm_down_pass<KIND_M1, P, R>(S, M, L, F.far_stream);
  else if (ovl == 1)
    fmm_far_pass<KIND_M1, P, R>(S, x, M, L, F.far_stream);
  CPPE_HIP_CHECK(hipEventRecord(F.ev_join, F.far_stream));
  CPPE_HIP_CHECK(hipStreamWaitEvent(S.stream, F.ev_join, 0));
  launch_l2p_add<P, R>(S, L, out, S.stream);
}

template <int P>
void fmm_apply_b_t(System& S, const double* x, double* out, bool fp32) {
  FmmState& F     = *S.fmm;
  const size_t nc = S.tree.ncells;
  using O64       = Ops<KIND_M1, P, double>;
  if (fp32) {
    ensure_buf(F.M32, fmm_mbuf_size(nc, KIND_M1, P));
    ensure_buf(F.L32, nc * O64::LS);
    fmm_apply_b_prec<P, float>(S, x, out, F.M32.p, F.L32.p);
  } else {
    ensure_buf(F.M64, fmm_mbuf_size(nc, KIND_M1, P));
    ensure_buf(F.L64, nc * O64::LS);
    fmm_apply_b_prec<P, double>(S, x, out, F.M64.p, F.L64.p);
  }
}

// kMultiRhs right-hand sides ([rhs][component][site]) through one near-field launch: the far
// passes run one right-hand side after another into their own expansions, the staged near-field
// kernel evaluates each pair's geometry once for all of them. Staged near variants only.
template <int P, typename R>
void fmm_apply_b_multi_prec(System& S, const double* x, double* out, R* M, R* L, MU4Of<R>* mu4) {
  constexpr int K  = kMultiRhs;
  FmmState& F      = *S.fmm;
  const int prec   = sizeof(R) == 4;
  const double dmp = S.opts.damping;
  const Octree& T  = S.tree;
  using O          = Ops<KIND_M1, P, R>;
  const size_t ms = fmm_mbuf_size(T.ncells, KIND_M1, P), ls = T.ncells * (size_t)O::LS, n3 = 3 * (size_t)S.n;
  const int ovl   = !S.prof && F.far_stream ? F.tune.overlap[prec] : 0;
  hipStream_t fst = ovl ? F.far_stream : S.stream;
  auto far = [&] {  // (overlap 2: the upward passes already ran)
    for (int k = 0; k < K; ++k) {
      if (ovl == 2)
        fmm_down_pass<KIND_M1, P, R>(S, M + k * ms, L + k * ls, fst);
      else
        fmm_far_pass<KIND_M1, P, R>(S, x + k * n3, M + k * ms, L + k * ls, fst);
    }
  };
  if (ovl) {
    if (ovl == 2)
      for (int k = 0; k < K; ++k) fmm_upward_pass<KIND_M1, P, R>(S, x + k * n3, M + k * ms, S.stream);
    CPPE_HIP_CHECK(hipEventRecord(F.ev_fork, S.stream));
    CPPE_HIP_CHECK(hipStreamWaitEvent(F.far_stream, F.ev_fork, 0));
    if (ovl == 3) far();  // ahead of the near field
  } else {
    far();  // the fused L2P of the near-field kernel reads the local expansions
  }
  ensure_pos4(S);
  NearArgs a{T.leaf_of.p, T.start.p, T.count.p, T.p2p_off.p, T.p2p_src.p, T.cx.p, T.cy.p, T.cz.p, F.pos4.p,
             S.excl_off.p, S.excl_idx.p, x, S.ainv6.p, dmp, S.n, out};
  const dim3 g(ceil_div(S.n, 256)), b(256);
  for (int k = 0; k < K; ++k)
    hipLaunchKernelGGL(k_pack_mu4<MU4Of<R>>, g, b, 0, S.stream, S.n, x + k * n3, mu4 + (size_t)k * S.n);
  if (dmp > 0.0)
    launch_near<P, R, MODE_BAPPLY, true, MU4Of<R>, K>(S, a, mu4, L, !ovl, ls);
  else
    launch_near<P, R, MODE_BAPPLY, false, MU4Of<R>, K>(S, a, mu4, L, !ovl, ls);
  if (ovl) {
    if (ovl != 3) far();  // enqueued after the near field so that both start together (see fmm_apply_b_prec)
    CPPE_HIP_CHECK(hipEventRecord(F.ev_join, F.far_stream));
    CPPE_HIP_CHECK(hipStreamWaitEvent(S.stream, F.ev_join, 0));
    hipLaunchKernelGGL((k_fmm_l2p_add<P, R>), dim3(ceil_div(S.n, 256), K), dim3(256), 0, S.stream, S.n, T.leaf_of.p,
                       F.pos4.p, T.cx.p, T.cy.p, T.cz.p, L, out, ls);
  }
}

template <int P>
void fmm_apply_b_multi_t(System& S, const double* x, double* out, bool fp32) {
  FmmState& F = *S.fmm;
  if (fp32)
    fmm_apply_b_multi_prec<P, float>(S, x, out, F.M32k.p, F.L32k.p, F.mu4fk.p);
  else
    fmm_apply_b_multi_prec<P, double>(S, x, out, F.M64k.p, F.L64k.p, F.mu4k.p);
}

template <int P>
void fmm_field1_t(System& S, const double* src, double* out, double damping) {
  FmmState& F     = *S.fmm;
  const size_t nc = S.tree.ncells;
  using O64       = Ops<KIND_M1, P, double>;
  ensure_buf(F.M64, fmm_mbuf_size(nc, KIND_M1, P));
  ensure_buf(F.L64, nc * O64::LS);
  fmm_far_pass<KIND_M1, P, double>(S, src, F.M64.p, F.L64.p, S.stream);
  fmm_near_pass<P, double, MODE_FIELD1>(S, src, damping, F.L64.p, out);
}

// src10: 10 SoA arrays (q, mu, packed quadrupole), device order
template <int P>
void fmm_static_t(System& S, const double* src10, double* out, double damping) {
  FmmState& F     = *S.fmm;
  const size_t nc = S.tree.ncells;
  using OC        = Ops<KIND_MC, P, double>;
  ensure_buf(F.Mc, fmm_mbuf_size(nc, KIND_MC, P));
  ensure_buf(F.Lc, nc * OC::LS);
  fmm_far_pass<KIND_MC, P, double>(S, src10, F.Mc.p, F.Lc.p, S.stream);
  fmm_p2p_pass<KIND_MC, P, double, MODE_STATIC>(S, src10, damping, F.Lc.p, out);
}

#define CPPE_INSTANTIATE_FMM(P)                                                          \
  template void fmm_apply_b_t<P>(System&, const double*, double*, bool);                 \
  template void fmm_apply_b_multi_t<P>(System&, const double*, double*, bool);           \
  template void fmm_field1_t<P>(System&, const double*, double*, double);                \
  template void fmm_static_t<P>(System&, const double*, double*, double);

}  // namespace cppe_dev
