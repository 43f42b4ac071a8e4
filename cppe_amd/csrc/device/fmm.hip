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

void fmm_destroy(System& S) {
  delete S.fmm;
  S.fmm = nullptr;
}

}  // namespace cppe_dev
