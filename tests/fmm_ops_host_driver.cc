// Host build of the generated FMM operators (cppe_amd/csrc/device/fmm_ops.gen.hh) for CPU tests:
// device qualifiers and the hardware rsqrt are mapped to host equivalents.
#include <cmath>
#define __device__
#define __forceinline__ inline
#define __restrict__
#define CPPE_FMMOPS_HOST
namespace cppe_dev {
inline double rsqrt_f64(double x) { return 1.0 / std::sqrt(x); }
inline float rsqrt_f32(float x) { return 1.0f / std::sqrt(x); }
}  // namespace cppe_dev
#include "fmm_ops.gen.hh"

using namespace cppe_dev::fmmops;
using cppe_dev::fmm_msize;
using cppe_dev::fmm_lsize;

#define P CPPE_FMMOPS_ONLY_P
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

extern "C" {
void gen_p2m(int mc, double x, double y, double z, const double* S, double* M) {
  if (mc) CAT(p2m_mc_p, P)<double>(x, y, z, S, M); else CAT(p2m_m1_p, P)<double>(x, y, z, S, M);
}
void gen_m2m(int mc, double x, double y, double z, const double* Mc, double* Mp) {
  if (mc) CAT(m2m_mc_p, P)<double>(x, y, z, Mc, Mp); else CAT(m2m_m1_p, P)<double>(x, y, z, Mc, Mp);
}
// M2L as the device runs it: the multipole row is folded onto its harmonic components first (on a
// copy here, in place on the device), contracted into the local coefficients with n_z < 2, and the
// others restored (per cell on the device, after all M2L pairs).
void gen_m2l(int mc, double x, double y, double z, const double* M, double* L) {
  double f[256], l[256] = {};
  const int ms = mc ? fmm_msize(1, P) : fmm_msize(0, P), ls = mc ? fmm_lsize(1, P) : fmm_lsize(0, P);
  for (int k = 0; k < ms; ++k) f[k] = M[k];
  if (mc) {
    CAT(fold_mc_p, P)<double>(f);
    CAT(CAT(m2l_mc_p, P), _f64)(x, y, z, f, l);
    CAT(unfold_mc_p, P)<double>(l);
  } else {
    CAT(fold_m1_p, P)<double>(f);
    CAT(CAT(m2l_m1_p, P), _f64)(x, y, z, f, l);
    CAT(unfold_m1_p, P)<double>(l);
  }
  for (int k = 0; k < ls; ++k) L[k] += l[k];
}
void gen_m2l_f32(int mc, float x, float y, float z, const float* M, float* L) {
  float f[256], l[256] = {};
  const int ms = mc ? fmm_msize(1, P) : fmm_msize(0, P), ls = mc ? fmm_lsize(1, P) : fmm_lsize(0, P);
  for (int k = 0; k < ms; ++k) f[k] = M[k];
  if (mc) {
    CAT(fold_mc_p, P)<float>(f);
    CAT(CAT(m2l_mc_p, P), _f32)(x, y, z, f, l);
    CAT(unfold_mc_p, P)<float>(l);
  } else {
    CAT(fold_m1_p, P)<float>(f);
    CAT(CAT(m2l_m1_p, P), _f32)(x, y, z, f, l);
    CAT(unfold_m1_p, P)<float>(l);
  }
  for (int k = 0; k < ls; ++k) L[k] += l[k];
}
void gen_l2l(int mc, double x, double y, double z, const double* Lp, double* Lc) {
  if (mc) CAT(l2l_mc_p, P)<double>(x, y, z, Lp, Lc); else CAT(l2l_m1_p, P)<double>(x, y, z, Lp, Lc);
}
void gen_l2p(int mc, double x, double y, double z, const double* L, double* F) {
  if (mc) CAT(l2p_mc_p, P)<double>(x, y, z, L, F[0], F[1], F[2]);
  else CAT(l2p_m1_p, P)<double>(x, y, z, L, F[0], F[1], F[2]);
}
}
