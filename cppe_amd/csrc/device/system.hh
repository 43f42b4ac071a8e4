// Device-resident state behind one cppe_hip_system handle.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../../include/cppe_hip.h"
#include "common.hh"

namespace cppe_dev {

// Device buffer with RAII release.
template <typename T>
struct dbuf {
  T* p = nullptr;
  size_t n = 0;
  dbuf() = default;
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  ~dbuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    CPPE_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void zero(hipStream_t s) { CPPE_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
  operator T*() const { return p; }
};

// Lists of the symmetric near-field kernel (near_sym.hh) for chunks of ts targets: the work items;
// per leaf A its upper list (A's own sites, then the sites of its near leaves B > A in the order of
// the expanded near list), with the reverse slot of each entry (for chunk 0 and component x, and
// |B|); per leaf B the layout of its incoming reverse blocks (first slot, block count).
struct SymLists {
  int ts = 0;            // targets per chunk (0: not built)
  int nitems = 0;
  long long nrev = 0;    // reverse slots
  dbuf<int4> items;      // (leaf, first target, targets, chunk index)
  dbuf<long long> uoff;  // ncells + 1
  dbuf<long long> usplit;  // per cell: first entry of the far class (FP32 cutoff)
  dbuf<long long> rev_off; // ncells + 1
  dbuf<int> usrc, nin;
  dbuf<int2> udst;
};

// Octree in level order (level 0 = root). Sites are sorted by Morton key ("device order");
// every cell covers the contiguous device-order range [start, start + count).
struct Octree {
  int ncells = 0, nleaves = 0, max_level = 0;
  int max_leaf = 0;  // sites in the largest leaf
  std::vector<int> level_offset;  // host copy: first cell index of each level (+ end)
  dbuf<double> cx, cy, cz, cr, crmax;
  dbuf<int> start, count, parent, level, child_first, nchild, octmask;
  dbuf<int> leaf_ids;  // indices of leaf cells
  dbuf<int> leaf_of;   // leaf cell of each site (device order)
  // interaction lists grouped by target (owner computes): CSR over target cells
  long long n_m2l = 0, n_p2p = 0, p2p_site_pairs = 0;
  dbuf<int> m2l_off, m2l_src;  // ncells+1 offsets, sources
  dbuf<int> p2p_off, p2p_src;  // ncells+1 offsets (non-empty only for leaves), source leaves
  // expanded near lists: per target leaf, the concatenated site ranges of its near leaves
  long long nx = -1;           // total length (-1: not built)
  dbuf<long long> xoff;        // ncells+1 offsets
  dbuf<int> xsrc;              // source sites (device order)
  // Each leaf's expanded list is ordered by the smallest Thole argument v of each source over the
  // leaf's targets: v_min < 18, then 18 <= v_min < 40, then v_min >= 40 (stable within each class).
  // xsplit[1] (FP32) / xsplit[0] (FP64): offset of the first source that takes the undamped pair
  // in that precision for every target of the leaf (kFarClass32, kTholeCutoff).
  dbuf<long long> xsplit[2];
  double xdamp = 0.0;  // damping factor of the classification (valid for kernels with damping >= xdamp)
  SymLists sym[5];     // symmetric near field with chunks of 4, 6, 8, 12, 16 targets (sym_slot)
  bool built = false;
};

struct FmmState;  // FMM expansions and operator scratch (fmm.hip)

// Pair list of the local Neumann preconditioner (precond.hip): CSR over sites (device order),
// neighbour index and the FP32 damped dipole tensor block (6 packed values) of each pair.
struct Precond {
  double rc = 0.0, damping = 0.0;
  long long npairs = 0;
  dbuf<int> off, nj;  // pairs of site i: partners nj[off[i] .. off[i+1])
  dbuf<double4> site;  // x, y, z, Thole scale s per site
};
struct MlGrid;    // neighbour grid of the learned warm start (ml_features.hip)

// Outcome of a device CG solve, per segment.
struct CgResult {
  std::vector<int> iters, converged;
  std::vector<double> rnorm;
  std::vector<double> norms0;  // residual norm per iteration of segment 0
  std::vector<std::vector<double>> norms;  // residual norms per iteration, every segment
  int loop_iters = 0;          // iterations of the shared loop (max over segments)
  int reliable_updates = 0;    // FP64 true-residual applies of the mixed form
  int confirmations    = 0;    // FP32-corrected confirmations of the mixed form (one FP32 apply each)
  int fp64_fallback = -1;      // mixed form: iteration at which it continued in FP64 (-1: never)
  int step_launches = 0;       // launches of step sequences (each of 1 to 8 steps)
  int steps_launched = 0;      // steps in them (those that start after an event take none)
  int host_reads = 0;          // status reads of the host loop (one per launch it waits for)
  bool all_converged = false;
};

// Work split, scalars, status block and captured graphs of one device CG (solver.hip). Vectors
// are [rhs][component][site]: nrhs blocks of 3 * na values, component-major; each segment lies
// inside one right-hand side (segments per right-hand side: spr).
struct CgState {
  int nseg = 0, nchunk = 0;
  int na = 0;    // sites per right-hand side
  int nrhs = 1;  // right-hand sides
  int spr = 1;   // segments per right-hand side
  dbuf<int> chunk_lo, chunk_seg, seg_chunk;
  dbuf<double> part, sd, params, hist;
  dbuf<int> si, status;
  int hist_cap = 0;
  int* h_status = nullptr;  // host-mapped mirror of the status block (written by the final kernels)
  int* d_hstat  = nullptr;  // its device address
  int expect    = 0;        // final kernels launched so far: the mirror's count once they have run
  double h_params[8] = {};  // the solve parameters last uploaded to params (valid: params_ok)
  bool params_ok     = false;
  char* h_out        = nullptr;  // host-mapped: a solve's segment tables and residual history
  char* d_out        = nullptr;  // its device address
  size_t out_bytes   = 0;
  // g_step[form][c]: c consecutive steps, 1 <= c <= kMaxGraphSteps (form 0 FP64, 1 mixed)
  static constexpr int kMaxGraphSteps = 8;
  hipGraphExec_t g_step[2][kMaxGraphSteps + 1] = {}, g_rel = nullptr, g_conf = nullptr, g_dir[2] = {nullptr, nullptr};
  hipGraphExec_t g_init[2] = {nullptr, nullptr};  // initial residual and the first step (form 0 FP64, 1 mixed)
  double *x = nullptr, *rhs = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *t = nullptr;
  float4* rec32   = nullptr;  // near-field dipole records of p written by the direction kernels
  double4* rec64 = nullptr;   // (one of them, in the precision of the step's apply; none: nrhs > 1)
  dbuf<double> store;  // backing of the vectors when this state owns them
  dbuf<double> wbuf;   // alpha r of the local preconditioner (same layout as r)
  // mixed form: moments and residual at the last FP64 true residual and the correction since
  // (FP32-corrected confirmation), in rstore
  double *xrel = nullptr, *rrel = nullptr, *dv = nullptr;
  dbuf<double> rstore;
  void release_graphs();
  ~CgState();
};

struct System {
  int n = 0;
  cppe_hip_options opts{};
  hipStream_t stream = nullptr;
  bool has_alpha = false;
  bool all_damped = false;  // every site has alpha_iso > 0 (Thole damping applies to every pair)

  // site data in device order (Morton-sorted)
  dbuf<double> x, y, z;          // positions
  dbuf<double> s;                // alpha_iso^(-1/6), 0 = undamped site
  dbuf<double> alpha6, ainv6;    // packed 3x3 (xx,xy,xz,yy,yz,zz), SoA [k*n + i]
  dbuf<int> excl_off, excl_idx;  // CSR in device order
  dbuf<int> perm;                // device order -> caller index
  dbuf<int> iperm;               // caller index -> device order
  std::vector<int> h_perm;
  std::vector<int> h_iperm;  // caller index -> device order (filled on first use)
  int iperm_host(int i) {
    if (h_iperm.empty()) {
      h_iperm.resize(n);
      for (int k = 0; k < n; ++k) h_iperm[h_perm[k]] = k;
    }
    return h_iperm[i];
  }
  std::vector<int> h_excl_off, h_excl_idx;  // caller order, values as given

  Octree tree;
  FmmState* fmm = nullptr;
  MlGrid* ml_grid = nullptr;

  // snapshots ("segments"): contiguous site ranges that form independent systems (own root
  // cell, own ncrit, own CG recurrence). A plain system is one segment.
  int nseg = 1;
  std::vector<int> h_seg_off;            // nseg + 1 (caller order; the same ranges in device order)
  dbuf<int> seg_off, seg_ncrit, seg_of;  // seg_of: segment of each site

  // solver work vectors: 3n SoA (component-major) in device order
  dbuf<double> v_mu, v_rhs, v_r, v_z, v_p, v_Ap, v_t, v_d;
  // CG state: cg for the system's own solve (one right-hand side, one segment per snapshot),
  // cgm for solves of several right-hand sides at once (built on first use)
  CgState cg, cgm;
  bool use_graphs = true;
  int graph_steps = 8;  // most CG steps launched as one sequence (1 to 8)
  // CG preconditioner: 0 = alpha r (the reference's Jacobi form), 1 = local first-order Neumann
  // alpha r + weight alpha T_loc alpha r over pairs closer than precond_rc (precond.hip)
  int precond = 0;
  double precond_rc = 5.6691783738;  // 3 A in bohr
  double precond_weight = 0.8;
  Precond* pc = nullptr;
  double reliable_drop = 1e-5;  // mixed form: residual drop that triggers an FP64 true residual
  int refine_stall = 10;        // mixed form: steps without halving that trigger one as well
  double refine_min_progress = 0.5;  // mixed form: a true residual above this fraction of the
                                     // previous one continues the solve in FP64
  double confirm_eta = 1e-5;  // mixed form: relative-error bound of an FP32 apply (0: FP64 confirmations only)
  int confirm_forms  = 3;     // bit 0: the gap bound of the recursive residual, bit 1: corrected confirmation
  CgResult last;            // outcome of the last solve
  dbuf<double> io_a, io_b;  // AoS3 staging in caller order
  dbuf<float> f32_src;      // FP32 staging for mixed-precision applies
  bool mu_valid = false;
  // When set, the near-field dipole records of the next B apply's precision already hold this
  // vector (the CG direction kernels write them beside p), and the apply does not pack them.
  const double* prepacked = nullptr;
  // When p is set (by a CG step around its apply into ap, one right-hand side, chunks of 256 sites
  // from site 0): the L2P pass that ends the apply also writes the chunk partials p.ap and ap.ap
  // at part[2 nchunk + c] and part[3 nchunk + c], and sets done.
  struct StepDots {
    const double* p = nullptr;
    const double* ap = nullptr;
    double* part = nullptr;
    int nchunk = 0;
    bool done = false;
  } dots;
  hipEvent_t* prof = nullptr;  // when set (8 events): [0..4] FMM phase boundaries, [6..7] around the
                               // staged near-field kernel alone

  ~System();
};

// ---- kernels/launchers implemented across the .hip translation units -----------------------
void build_octree(System& S);
void build_lists(System& S);
void build_near_expanded(System& S);
// lists of the symmetric near-field kernel with chunks of ts targets (4, 6, 8, 12 or 16) into
// S.tree.sym[sym_slot(ts)]
inline int sym_slot(int ts) { return ts <= 4 ? 0 : ts <= 6 ? 1 : ts <= 8 ? 2 : ts <= 12 ? 3 : 4; }
SymLists& build_sym_lists(System& S, int ts);

// field of order-m sources given in device-order SoA (ncomp(m) arrays of n) into F (3n SoA)
void direct_field(System& S, int m, const double* src, double* F, double damping);
void direct_static(System& S, const double* q, const double* mu, const double* th, double* F,
                   double damping);
// out = field_m1(x) + alpha^-1 x, FP64 (or FP32 arithmetic when fp32 is set)
void apply_b(System& S, const double* x, double* out, bool fp32);
void precondition(System& S, const double* in, double* out);
void fill_f64(System& S, double* a, size_t n, double v);
// segment tables and CG state; seg_off has nseg + 1 entries
void setup_segments(System& S, const std::vector<int>& seg_off, const std::vector<int>& seg_ncrit);
// CG on C.rhs from the guess in C.x (S.cg: the system's vectors S.v_rhs / S.v_mu); returns the
// shared loop's iteration count
int device_cg(System& S, CgState& C, double thresh, int maxiter, bool mixed, CgResult& res);
// S.cgm for nrhs right-hand sides of a one-snapshot system, its vectors allocated
void setup_multi_rhs(System& S, int nrhs);
void cg_graphs_release(System& S);
// out = B x for nrhs right-hand sides ([rhs][component][site] layout), one launch sequence
void apply_b_multi(System& S, const double* x, double* out, int nrhs, bool fp32);
void fmm_reserve_multi(System& S);
// near-field dipole records of a single-right-hand-side B apply (allocated by fmm_reserve)
void fmm_dipole_records(System& S, float4** rec32, double4** rec64);

void fmm_field(System& S, int m, const double* src, double* F, double damping, bool fp32);
void fmm_static(System& S, const double* q, const double* mu, const double* th, double* F, double damping);
void fmm_destroy(System& S);
int fmm_near_variant(const System& S, bool fp32);  // tuned near-field variant (0 before the first apply)
void ml_grid_destroy(System& S);
void precond_build(System& S);
void precond_destroy(System& S);

// layout conversions between caller AoS (caller order) and device SoA (device order)
void aos_to_soa(System& S, const double* aos, double* soa, int ncomp);
void soa_to_aos(System& S, const double* soa, double* aos, int ncomp);

}  // namespace cppe_dev
