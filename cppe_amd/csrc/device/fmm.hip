// FMM entry points: order dispatch and per-system FMM state (kernels: fmm_impl.hh).
#define CPPE_FMMOPS_ONLY_P 0  // no operator code in the dispatch unit
#include "fmm_impl.hh"

namespace cppe_dev {

// instantiated per order in fmm_order.hip
#define CPPE_EXTERN_FMM(P)                                                                \
  extern template void fmm_apply_b_t<P>(System&, const double*, double*, bool);           \
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
  if (!S.fmm) S.fmm = new FmmState();
  S.fmm->order = S.opts.order;
  return *S.fmm;
}

void apply_b_fmm(System& S, const double* x, double* out, bool fp32) {
  state(S);
#define CALL(P) fmm_apply_b_t<P>(S, x, out, fp32)
  CPPE_FMM_DISPATCH(CALL)
#undef CALL
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

// Allocates every buffer an FMM B apply touches (both precisions), so that applies can be
// captured into a graph.
void fmm_reserve(System& S) {
  FmmState& F     = state(S);
  const size_t nc = S.tree.ncells;
  const int p     = S.opts.order;
  ensure_buf(F.M64, nc * fmm_msize(0, p));
  ensure_buf(F.L64, nc * fmm_lsize(0, p));
  ensure_buf(F.M32, nc * fmm_msize(0, p));
  ensure_buf(F.L32, nc * fmm_lsize(0, p));
  ensure_buf(F.mu4, S.n);
  ensure_buf(F.mu4f, S.n);
  ensure_pos4(S);
  ensure_posf4(S);
  ensure_near_items(S);
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
}

// Autotuner of the B-apply launch shapes: for each precision, every near-field candidate
// (variant x block size) and every M2L group width is timed on this system's own tree with HIP
// events around the kernel (mean of reps applies after one warm-up), and the fastest is kept.
// ms receives the timings: per precision (FP64 then FP32) kNearCand near values then kM2lCand
// M2L values; choice receives (near variant, near block, m2l group) per precision.
static const int kNearCand[][2] = {{0, 64}, {0, 128}, {0, 256}, {1, 64}, {1, 128}, {1, 256}};
static const int kM2lCand[]     = {16, 32, 64};
constexpr int kNNear = 6, kNM2l = 3;

void fmm_autotune(System& S, int reps, int* choice, double* ms) {
  fmm_reserve(S);
  FmmState& F    = *S.fmm;
  hipStream_t st = S.stream;
  hipEvent_t ev[6];
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
  try {
    fill_f64(S, S.v_d.p, 3 * (size_t)S.n, 0.01);
    for (int prec = 0; prec < 2; ++prec) {
      const bool fp32 = prec == 1;
      double* m       = ms + prec * (kNNear + kNM2l);
      int best        = 0;
      for (int c = 0; c < kNNear; ++c) {
        F.tune.near_variant[prec] = kNearCand[c][0];
        F.tune.near_block[prec]   = kNearCand[c][1];
        m[c]                      = timed(fp32, 0);
        if (m[c] < m[best]) best = c;
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
      choice[3 * prec]       = F.tune.near_variant[prec];
      choice[3 * prec + 1]   = F.tune.near_block[prec];
      choice[3 * prec + 2]   = F.tune.m2l_group[prec];
    }
  } catch (...) {
    S.prof = nullptr;
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  cg_graphs_release(S);  // captured graphs hold the previous launch shapes
}

void fmm_set_tune(System& S, const int* choice) {
  FmmState& F = state(S);
  for (int prec = 0; prec < 2; ++prec) {
    F.tune.near_variant[prec] = choice[3 * prec];
    F.tune.near_block[prec]   = choice[3 * prec + 1];
    F.tune.m2l_group[prec]    = choice[3 * prec + 2];
  }
  cg_graphs_release(S);
}

void fmm_destroy(System& S) {
  delete S.fmm;
  S.fmm = nullptr;
}

}  // namespace cppe_dev
