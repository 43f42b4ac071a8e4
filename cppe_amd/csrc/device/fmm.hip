// FMM entry points: order dispatch and per-system FMM state (kernels: fmm_impl.hh).
#define CPPE_FMMOPS_ONLY_P 0  // no operator code in the dispatch unit
#include "fmm_impl.hh"

#include <algorithm>
#include <vector>

namespace cppe_dev {

// instantiated per order in fmm_order.hip
#define CPPE_EXTERN_FMM(P)                                                                \
  extern template void fmm_apply_b_t<P>(System&, const double*, double*, bool);           \
  extern template void fmm_apply_b_multi_t<P>(System&, const double*, double*, bool);     \
  extern template void fmm_field1_t<P>(System&, const double*, double*, double);          \
  extern template void fmm_static_t<P>(System&, const double*, double*, double);
CPPE_EXTERN_FMM(2)
CPPE_EXTERN_FMM(3)
CPPE_EXTERN_FMM(4)
CPPE_EXTERN_FMM(5)
CPPE_EXTERN_FMM(6)
CPPE_EXTERN_FMM(7)
CPPE_EXTERN_FMM(8)

#define CPPE_FMM_DISPATCH(call)                                                          \
  switch (S.opts.order) {                                                                \
    case 2: call(2); break;                                                              \
    case 3: call(3); break;                                                              \
    case 4: call(4); break;                                                              \
    case 5: call(5); break;                                                              \
    case 6: call(6); break;                                                              \
    case 7: call(7); break;                                                              \
    case 8: call(8); break;                                                              \
    default: throw std::runtime_error("tree_expansion_order must be in 2..8 for fmm summation"); \
  }

static FmmState& state(System& S) {
  if (!S.tree.built || S.tree.m2l_off.p == nullptr)
    throw std::runtime_error("fmm summation needs the interaction lists of this system");
  if (!S.fmm) {
    auto* F = new FmmState();
    try {
      CPPE_HIP_CHECK(hipStreamCreateWithFlags(&F->far_stream, hipStreamNonBlocking));
      CPPE_HIP_CHECK(hipEventCreateWithFlags(&F->ev_fork, hipEventDisableTiming));
      CPPE_HIP_CHECK(hipEventCreateWithFlags(&F->ev_join, hipEventDisableTiming));
    } catch (...) {
      delete F;
      throw;
    }
    S.fmm = F;
  }
  S.fmm->order = S.opts.order;
  return *S.fmm;
}

void apply_b_fmm(System& S, const double* x, double* out, bool fp32) {
  state(S);
#define CALL(P) fmm_apply_b_t<P>(S, x, out, fp32)
  CPPE_FMM_DISPATCH(CALL)
#undef CALL
}

// Groups of kMultiRhs right-hand sides share one near-field launch (staged near variants); the
// rest, and every right-hand side under another variant, are applied one at a time.
void apply_b_multi(System& S, const double* x, double* out, int nrhs, bool fp32) {
  FmmState& F     = state(S);
  const size_t n3 = 3 * (size_t)S.n;
  int k           = 0;
  if (F.tune.near_variant[fp32 ? 1 : 0] >= 5)
    for (; k + kMultiRhs <= nrhs; k += kMultiRhs) {
#define CALL(P) fmm_apply_b_multi_t<P>(S, x + k * n3, out + k * n3, fp32)
      CPPE_FMM_DISPATCH(CALL)
#undef CALL
    }
  for (; k < nrhs; ++k) apply_b_fmm(S, x + k * n3, out + k * n3, fp32);
}

void fmm_dipole_records(System& S, float4** rec32, double4** rec64) {
  FmmState& F = state(S);
  ensure_buf(F.mu4, S.n);
  ensure_buf(F.mu4f, S.n);
  *rec32 = F.mu4f.p;
  *rec64 = F.mu4.p;
}

void fmm_reserve_multi(System& S) {
  FmmState& F     = state(S);
  const size_t nc = S.tree.ncells;
  const int p     = S.opts.order;
  ensure_buf(F.M64k, kMultiRhs * fmm_mbuf_size(nc, 0, p));
  ensure_buf(F.L64k, kMultiRhs * nc * fmm_lsize(0, p));
  ensure_buf(F.M32k, kMultiRhs * fmm_mbuf_size(nc, 0, p));
  ensure_buf(F.L32k, kMultiRhs * nc * fmm_lsize(0, p));
  ensure_buf(F.mu4k, kMultiRhs * (size_t)S.n);
  ensure_buf(F.mu4fk, kMultiRhs * (size_t)S.n);
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
}

__global__ void k_pack10(int n, int m, const double* __restrict__ src, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 10; ++c) out[(size_t)c * n + i] = 0.0;
  if (m == 0) out[i] = src[i];
  if (m == 2)
    for (int c = 0; c < 6; ++c) out[(size_t)(4 + c) * n + i] = src[(size_t)c * n + i];
}

void fmm_field(System& S, int m, const double* src, double* F, double damping, bool fp32) {
  (void)fp32;
  FmmState& st = state(S);
  if (m == 1) {
#define CALL(P) fmm_field1_t<P>(S, src, F, damping)
    CPPE_FMM_DISPATCH(CALL)
#undef CALL
    return;
  }
  ensure_buf(st.src10, 10 * (size_t)S.n);
  hipLaunchKernelGGL(k_pack10, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, m, src, st.src10.p);
#define CALL(P) fmm_static_t<P>(S, st.src10.p, F, damping)
  CPPE_FMM_DISPATCH(CALL)
#undef CALL
}

void fmm_static(System& S, const double* q, const double* mu, const double* th, double* F, double damping) {
  (void)mu;
  (void)th;
  state(S);
  // q, mu, th are consecutive SoA blocks of one 10-component source array
#define CALL(P) fmm_static_t<P>(S, q, F, damping)
  CPPE_FMM_DISPATCH(CALL)
#undef CALL
}

// Target clusters for the cluster near-field kernel: leaves in device order packed greedily into
// runs of <= 64 sites (a leaf with more sites is split into 64-site chunks); per cluster, the
// union of its leaves' near lists in memory order of the source leaves, each with the mask of
// lanes (targets) for which it is near.
void ensure_clusters(System& S) {
  FmmState& F     = *S.fmm;
  const Octree& T = S.tree;
  if (F.n_clusters > 0) return;
  hipStream_t st = S.stream;
  const int nc   = T.ncells;
  std::vector<int> ids(T.nleaves), start(nc), count(nc), off(nc + 1), src(std::max<long long>(T.n_p2p, 1));
  CPPE_HIP_CHECK(hipMemcpyAsync(ids.data(), T.leaf_ids.p, sizeof(int) * T.nleaves, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(start.data(), T.start.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(count.data(), T.count.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipMemcpyAsync(off.data(), T.p2p_off.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost, st));
  if (T.n_p2p)
    CPPE_HIP_CHECK(hipMemcpyAsync(src.data(), T.p2p_src.p, sizeof(int) * T.n_p2p, hipMemcpyDeviceToHost, st));
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  std::sort(ids.begin(), ids.end(), [&](int a, int b) { return start[a] < start[b]; });
  std::vector<int2> ranges;
  std::vector<int> coff{0}, csrc;
  std::vector<unsigned long long> cmask;
  std::vector<int> slot(nc, -1);
  struct Member { int leaf, lane0, n; };
  std::vector<Member> cur;
  int cur_first = 0, cur_n = 0;
  std::vector<std::pair<int, unsigned long long>> uni;  // (source leaf, mask) of the open cluster
  auto flush = [&]() {
    if (cur.empty()) return;
    uni.clear();
    for (const Member& m : cur) {
      const unsigned long long bits = (m.n >= 64 ? ~0ull : ((1ull << m.n) - 1)) << m.lane0;
      for (int k = off[m.leaf]; k < off[m.leaf + 1]; ++k) {
        const int B = src[k];
        if (slot[B] < 0) {
          slot[B] = (int)uni.size();
          uni.push_back({B, 0ull});
        }
        uni[slot[B]].second |= bits;
      }
    }
    for (auto& e : uni) slot[e.first] = -1;
    std::sort(uni.begin(), uni.end(), [&](const auto& x, const auto& y) { return start[x.first] < start[y.first]; });
    for (auto& e : uni) {
      csrc.push_back(e.first);
      cmask.push_back(e.second);
    }
    coff.push_back((int)csrc.size());
    ranges.push_back(make_int2(cur_first, cur_n));
    cur.clear();
    cur_n = 0;
  };
  for (int a : ids) {
    const int ca = count[a];
    if (ca > 64) {  // chunks of a large leaf, each a cluster of its own
      flush();
      for (int t0 = 0; t0 < ca; t0 += 64) {
        cur_first = start[a] + t0;
        cur_n     = std::min(64, ca - t0);
        cur.push_back({a, 0, cur_n});
        flush();
      }
      continue;
    }
    if (cur_n + ca > 64) flush();
    if (cur.empty()) cur_first = start[a];
    cur.push_back({a, cur_n, ca});
    cur_n += ca;
  }
  flush();
  auto up = [&](auto& b, const auto& h) {
    b.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty())
      CPPE_HIP_CHECK(hipMemcpyAsync(b.p, h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, st));
  };
  up(F.cl_range, ranges);
  up(F.cl_off, coff);
  up(F.cl_src, csrc);
  up(F.cl_mask, cmask);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  F.n_clusters = (int)ranges.size();
}

// Allocates every buffer an FMM B apply touches (both precisions), so that applies can be
// captured into a graph.
void fmm_reserve(System& S) {
  FmmState& F     = state(S);
  // everything below depends on the tree, the order and the near variants only: once reserved for
  // these variants, a later solve neither allocates nor waits for the stream here
  if (F.reserved_for[0] == F.tune.near_variant[0] && F.reserved_for[1] == F.tune.near_variant[1]) return;
  const size_t nc = S.tree.ncells;
  const int p     = S.opts.order;
  ensure_buf(F.M64, fmm_mbuf_size(nc, 0, p));
  ensure_buf(F.L64, nc * fmm_lsize(0, p));
  ensure_buf(F.M32, fmm_mbuf_size(nc, 0, p));
  ensure_buf(F.L32, nc * fmm_lsize(0, p));
  ensure_buf(F.mu4, S.n);
  ensure_buf(F.mu4f, S.n);
  ensure_pos4(S);
  ensure_posf4(S);
  ensure_excl_near(S);
  ensure_m2l_rec(S, S.stream);  // the synchronisation below orders it before any stream's M2L
  ensure_near_items(S, 0);
  for (int prec = 0; prec < 2; ++prec) {
    const int v = F.tune.near_variant[prec];
    if (v == 3 || v >= 5) build_near_expanded(S);
    if (v >= 5) ensure_near_items(S, near_sym_variant(v) ? 3 : v - 5);  // symmetric: other launches use chunks of 8
    if (near_sym_variant(v)) {
      if (one_leaf_system(S))
        ensure_dsym(S, prec, kSymTs[prec][v - 9], S.opts.damping);
      else
        ensure_sym(S, prec, kSymTs[prec][v - 9], S.opts.damping);
    }
  }
  if (F.tune.near_variant[0] == 4 || F.tune.near_variant[1] == 4) ensure_clusters(S);
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  F.reserved_for[0] = F.tune.near_variant[0];
  F.reserved_for[1] = F.tune.near_variant[1];
}

// Autotuner of the B-apply launch shapes: for each precision, every near-field candidate
// (variant x block size) and every M2L group width is timed on this system's own tree with HIP
// events around the kernel (mean of reps applies after one warm-up), and the fastest is kept.
// Near-field variants: 0 thread per target, 1 wavefront per leaf chunk with a uniform source
// walk, 2 wavefront per leaf with LDS-staged sources, 3 wavefront per leaf chunk with (target,
// source slot) lanes over the expanded near list, 4 wavefront per target cluster (<= 64 sites of
// consecutive leaves) over the union of their near lists with lane masks, 5 wavefront per leaf
// chunk (5: 64 targets, 6: 32, 7: 16, 8: 8), several per block, sources of the expanded near list staged
// through wave-private LDS, 9, 10, 11 symmetric (near_sym.hh: each unordered pair once, lanes =
// sources, chunks of 16 / 12 / 8 targets in FP32 and 8 / 6 / 4 in FP64; the system's other
// launches then use variant 8).
// Then, for a staged near variant, the whole apply is timed with the far field in series, beside
// the near field on a second stream (enqueued after or ahead of it), and with only M2L / L2L beside
// it after the upward pass; the fastest is kept (the overlapped timing reported is the fastest
// overlapped form); each
// overlapped form is timed with the near field in 1, 2 and 4 launches.
// ms receives the timings: per precision (FP64 then FP32) kNearCand near values, kM2lCand M2L
// values, then the apply in series and overlapped; choice receives (near variant, near block, m2l
// group, overlap) per precision.
static const int kNearCand[][2] = {{0, 64}, {0, 128}, {0, 256}, {1, 64}, {1, 128}, {1, 256},
                                   {2, 64}, {3, 64},  {3, 128}, {3, 256}, {4, 64}, {4, 128},
                                   {4, 256}, {5, 64}, {5, 128}, {5, 256}, {6, 64}, {6, 128},
                                   {6, 256}, {7, 64}, {7, 128}, {7, 256}, {8, 64}, {8, 128},
                                   {8, 256}, {9, 64}, {9, 128}, {9, 256}, {10, 64}, {10, 128},
                                   {10, 256}, {11, 64}, {11, 128}, {11, 256}};
static const int kM2lCand[]     = {16, 32, 64};
constexpr int kNNear = 34, kNM2l = 3;

// The LDS-staged wavefront-per-leaf kernel (variant 2) gives one wavefront all targets of a leaf:
// its time grows with the square of the leaf size, so it is only a candidate for small leaves.
constexpr int kMaxLeafVariant2 = 1024;

int max_leaf_size(System& S) {
  Octree& T = S.tree;
  if (T.max_leaf > 0) return T.max_leaf;
  std::vector<int> ids(T.nleaves), cnt(T.ncells);
  CPPE_HIP_CHECK(hipMemcpyAsync(ids.data(), T.leaf_ids.p, sizeof(int) * T.nleaves, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipMemcpyAsync(cnt.data(), T.count.p, sizeof(int) * T.ncells, hipMemcpyDeviceToHost, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  int m = 0;
  for (int c : ids) m = std::max(m, cnt[c]);
  T.max_leaf = m;
  return m;
}

bool near_variant_allowed(System& S, int variant) {
  return variant != 2 || max_leaf_size(S) <= kMaxLeafVariant2;
}

// The overlap entry of a launch-shape choice: the far/near form (0 .. 3) plus 4 log2 of the
// near-field launches (1, 2 or 4; overlapped forms only).
int encode_overlap(int form, int splits) { return form + 4 * (splits >= 4 ? 2 : splits >= 2 ? 1 : 0); }
bool decode_overlap(int code, int& form, int& splits) {
  form         = code & 3;
  const int sl = code >> 2;
  splits       = 1 << sl;
  return code >= 0 && form <= 3 && sl <= 2 && (form > 0 || sl == 0);
}

void fmm_autotune(System& S, int reps, int* choice, double* ms) {
  fmm_reserve(S);
  FmmState& F    = *S.fmm;
  hipStream_t st = S.stream;
  hipEvent_t ev[8];
  for (auto& e : ev) CPPE_HIP_CHECK(hipEventCreate(&e));
  auto timed = [&](bool fp32, int phase) {  // phase 0: near (events 3-4), 1: m2l (events 1-2)
    apply_b_fmm(S, S.v_d.p, S.v_t.p, fp32);  // warm-up
    double acc = 0.0;
    for (int r = 0; r < reps; ++r) {
      S.prof = ev;
      apply_b_fmm(S, S.v_d.p, S.v_t.p, fp32);
      S.prof = nullptr;
      CPPE_HIP_CHECK(hipEventSynchronize(ev[4]));
      float t = 0.f;
      if (phase == 0)
        CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[3], ev[4]));
      else
        CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[1], ev[2]));
      acc += t;
    }
    return acc / reps;
  };
  auto timed_total = [&](bool fp32) {
    apply_b_fmm(S, S.v_d.p, S.v_t.p, fp32);  // warm-up
    CPPE_HIP_CHECK(hipEventRecord(ev[0], S.stream));
    for (int r = 0; r < reps; ++r) apply_b_fmm(S, S.v_d.p, S.v_t.p, fp32);
    CPPE_HIP_CHECK(hipEventRecord(ev[1], S.stream));
    CPPE_HIP_CHECK(hipEventSynchronize(ev[1]));
    float t = 0.f;
    CPPE_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
    return (double)t / reps;
  };
  try {  // S.v_d holds the caller's test dipoles
    for (int prec = 0; prec < 2; ++prec) {
      const bool fp32 = prec == 1;
      double* m       = ms + prec * (kNNear + kNM2l + 2);
      F.tune.overlap[prec] = 0;
      int best        = 0;
      for (int c = 0; c < kNNear; ++c) {
        if (!near_variant_allowed(S, kNearCand[c][0])) {
          m[c] = 0.0;  // not timed
          continue;
        }
        F.tune.near_variant[prec] = kNearCand[c][0];
        F.tune.near_block[prec]   = kNearCand[c][1];
        m[c]                      = timed(fp32, 0);
        if (m[best] <= 0.0 || m[c] < m[best]) best = c;
      }
      F.tune.near_variant[prec] = kNearCand[best][0];
      F.tune.near_block[prec]   = kNearCand[best][1];
      int bm                    = 0;
      for (int c = 0; c < kNM2l; ++c) {
        F.tune.m2l_group[prec] = kM2lCand[c];
        m[kNNear + c]          = timed(fp32, 1);
        if (m[kNNear + c] < m[kNNear + bm]) bm = c;
      }
      F.tune.m2l_group[prec] = kM2lCand[bm];
      m[kNNear + kNM2l]      = timed_total(fp32);
      m[kNNear + kNM2l + 1]  = 0.0;
      F.tune.near_splits[prec] = 1;
      if (F.tune.near_variant[prec] >= 5) {  // overlapped: the fastest of the three forms x near splits
        int bo = 0, bs = 1;
        double bt = 0.0;
        for (int form = 1; form <= 3; ++form)
          for (int sp = 1; sp <= 4; sp *= 2) {
            F.tune.overlap[prec]     = form;
            F.tune.near_splits[prec] = sp;
            const double t           = timed_total(fp32);
            if (bo == 0 || t < bt) {
              bo = form;
              bs = sp;
              bt = t;
            }
          }
        m[kNNear + kNM2l + 1]    = bt;
        const bool ovl           = bt < m[kNNear + kNM2l];
        F.tune.overlap[prec]     = ovl ? bo : 0;
        F.tune.near_splits[prec] = ovl ? bs : 1;
      }
      choice[4 * prec]     = F.tune.near_variant[prec];
      choice[4 * prec + 1] = F.tune.near_block[prec];
      choice[4 * prec + 2] = F.tune.m2l_group[prec];
      choice[4 * prec + 3] = encode_overlap(F.tune.overlap[prec], F.tune.near_splits[prec]);
    }
  } catch (...) {
    S.prof = nullptr;
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  cg_graphs_release(S);  // captured graphs hold the previous launch shapes
  F.reserved_for[0] = F.reserved_for[1] = -1;  // the sweeps may have rebuilt per-variant structures
}

void fmm_set_tune(System& S, const int* choice) {
  FmmState& F = state(S);
  for (int prec = 0; prec < 2; ++prec) {
    F.tune.near_variant[prec] = choice[4 * prec];
    F.tune.near_block[prec]   = choice[4 * prec + 1];
    F.tune.m2l_group[prec]    = choice[4 * prec + 2];
    decode_overlap(choice[4 * prec + 3], F.tune.overlap[prec], F.tune.near_splits[prec]);
  }
  cg_graphs_release(S);
  F.reserved_for[0] = F.reserved_for[1] = -1;
}

void fmm_destroy(System& S) {
  delete S.fmm;
  S.fmm = nullptr;
}

int fmm_near_variant(const System& S, bool fp32) { return S.fmm ? S.fmm->tune.near_variant[fp32 ? 1 : 0] : 0; }

}  // namespace cppe_dev
