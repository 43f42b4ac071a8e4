// Pair field kernels shared by the direct and FMM near-field paths.
#pragma once
#include "common.hh"

namespace cppe_dev {

// ---------------------------------------------------------------------------------------------
// Pair kernels. d = r_target - r_source.
// ---------------------------------------------------------------------------------------------
// -T2 mu with Thole damping: fE mu / R^3 - 3 fT (d.mu) d / R^5  (P2P<1,3>, field_m1.cc:127-193),
// split into the pair's geometry (c3 = fE / R^3, cf = 3 fT / R^5) and the accumulation of one
// dipole, so that several dipole vectors can share one geometry evaluation.
// SELF0: a zero displacement (a site with itself) contributes exactly zero.
template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void geo_m1_f64(double dx, double dy, double dz, double a, double& c3, double& cf) {
  const double r2   = dx * dx + dy * dy + dz * dz;
  const double rinv = SELF0 ? (r2 > 0.0 ? rsqrt_f64(r2) : 0.0) : rsqrt_f64(r2);
  const double ri2  = rinv * rinv;
  const double ri3  = rinv * ri2;
  double fE = 1.0, fT = 1.0;
  if (DAMP) {
    const double u = a * (r2 * rinv);
    if (u < kTholeCutoff && a > 0.0) thole_ET(u, exp(-u), fE, fT);
  }
  c3 = fE * ri3;
  cf = (3.0 * fT) * (ri3 * ri2);
}

template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void geo_m1_f32(float dx, float dy, float dz, float a, float& c3, float& cf) {
  const float r2   = dx * dx + dy * dy + dz * dz;
  const float rinv = SELF0 ? (r2 > 0.0f ? rsqrt_f32(r2) : 0.0f) : rsqrt_f32(r2);
  const float ri2  = rinv * rinv;
  const float ri3  = rinv * ri2;
  float fE = 1.0f, fT = 1.0f;
  if (DAMP) {
    const float u = a * (r2 * rinv);
    if (u < kTholeCutoff32 && a > 0.0f) thole_ET(u, expm_f32(u), fE, fT);
  }
  c3 = fE * ri3;
  cf = (3.0f * fT) * (ri3 * ri2);
}

// FP32 damped pair geometry for systems whose sites are all damped (alpha_iso > 0 everywhere),
// without branches. a2 = damping (s_i s_j) log2(e), so u2 = a2 R is the Thole argument v in
// base-2 units (v = u2 ln 2) and e^-v = 2^-u2 comes straight from v_exp_f32. With v = u2 ln2
//   f_E = 1 - (1 + ln2 u2 + (ln2)^2/2 u2^2) 2^-u2,   f_T = f_E - (ln2)^3/6 u2^3 2^-u2.
// A zero displacement (the site itself) contributes exactly zero: 1/R is clamped to 2^40, so
// R = 0, u2 = 0, 2^-u2 = 1 and f_E = f_T = 0 exactly (this needs a2 finite, i.e. damped sites).
// Large u2 (undamped in effect) gives 2^-u2 = 0 and f_E = f_T = 1.
__device__ __forceinline__ void geo_m1_f32_b2(float dx, float dy, float dz, float a2, float& c3, float& cf) {
  constexpr float kC1 = 0.693147180559945309f;   // ln 2
  constexpr float kC2 = 0.240226506959100712f;   // (ln 2)^2 / 2
  constexpr float kC3 = 0.0555041086648215800f;  // (ln 2)^3 / 6
  const float r2   = dx * dx + dy * dy + dz * dz;
  const float rinv = __builtin_fminf(rsqrt_f32(r2), 1099511627776.0f);  // 2^40
  const float ri2  = rinv * rinv;
  const float ri3  = ri2 * rinv;
  const float u    = a2 * (r2 * rinv);
  const float e    = __builtin_amdgcn_exp2f(-u);
  const float u2   = u * u;
  const float p    = __builtin_fmaf(u2, kC2, __builtin_fmaf(u, kC1, 1.0f));
  const float fE   = __builtin_fmaf(-e, p, 1.0f);
  const float q    = (u2 * u) * e;
  const float fT   = __builtin_fmaf(-kC3, q, fE);
  c3 = fE * ri3;
  cf = (fT * ri3) * (3.0f * ri2);  // in this order: at R = 0, ri3 ri2 = 2^200 would overflow
}

// field += c3 mu - cf (d.mu) d
__device__ __forceinline__ void acc_m1(double c3, double cf, double dx, double dy, double dz, double mx, double my,
                                       double mz, double& fx, double& fy, double& fz) {
  const double c5 = cf * (dx * mx + dy * my + dz * mz);
  fx = __builtin_fma(c3, mx, fx);
  fx = __builtin_fma(-c5, dx, fx);
  fy = __builtin_fma(c3, my, fy);
  fy = __builtin_fma(-c5, dy, fy);
  fz = __builtin_fma(c3, mz, fz);
  fz = __builtin_fma(-c5, dz, fz);
}

__device__ __forceinline__ void acc_m1(float c3, float cf, float dx, float dy, float dz, float mx, float my,
                                       float mz, float& fx, float& fy, float& fz) {
  const float c5 = cf * (dx * mx + dy * my + dz * mz);
  fx = __builtin_fmaf(c3, mx, fx);
  fx = __builtin_fmaf(-c5, dx, fx);
  fy = __builtin_fmaf(c3, my, fy);
  fy = __builtin_fmaf(-c5, dy, fy);
  fz = __builtin_fmaf(c3, mz, fz);
  fz = __builtin_fmaf(-c5, dz, fz);
}

// FP32 undamped pair in the form field += R^-3 (mu - 3 R^-2 (d.mu) d): the same tensor as
// c3 mu - cf (d.mu) d with c3 = R^-3, cf = 3 R^-5, one multiply fewer per pair (no R^-5).
// The displacement must be nonzero.
__device__ __forceinline__ void geo_far_f32(float dx, float dy, float dz, float& ri3, float& m3) {
  const float r2   = dx * dx + dy * dy + dz * dz;
  const float rinv = rsqrt_f32(r2);
  const float ri2  = rinv * rinv;
  ri3 = ri2 * rinv;
  m3  = -3.0f * ri2;
}

__device__ __forceinline__ void acc_far_f32(float ri3, float m3, float dx, float dy, float dz, float mx, float my,
                                            float mz, float& fx, float& fy, float& fz) {
  const float q = m3 * (dx * mx + dy * my + dz * mz);
  fx = __builtin_fmaf(ri3, __builtin_fmaf(q, dx, mx), fx);
  fy = __builtin_fmaf(ri3, __builtin_fmaf(q, dy, my), fy);
  fz = __builtin_fmaf(ri3, __builtin_fmaf(q, dz, mz), fz);
}

template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void pair_m1_f64(double dx, double dy, double dz, double mx, double my, double mz, double a,
                                            double& fx, double& fy, double& fz) {
  double c3, cf;
  geo_m1_f64<DAMP, SELF0>(dx, dy, dz, a, c3, cf);
  acc_m1(c3, cf, dx, dy, dz, mx, my, mz, fx, fy, fz);
}

template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void pair_m1_f32(float dx, float dy, float dz, float mx, float my, float mz, float a,
                                            float& fx, float& fy, float& fz) {
  float c3, cf;
  geo_m1_f32<DAMP, SELF0>(dx, dy, dz, a, c3, cf);
  acc_m1(c3, cf, dx, dy, dz, mx, my, mz, fx, fy, fz);
}

// Static field of charge + dipole + packed quadrupole S (half diagonal), combined as
// F(m0) - F(m1) + F(m2) (electric_fields.cc:163,174,186), with Thole factors from f_V/R:
//   m0: +q fE d/R^3
//   -m1: 3 fT (d.mu) d/R^5 - fE mu/R^3
//   m2: 15 fD P d/R^7 - 3 fT (tr d + Q d)/R^5, P = sum_q S_q d^q, tr = S0+S3+S5,
//       (Q d) = (2S0 dx + S1 dy + S2 dz, S1 dx + 2S3 dy + S4 dz, S2 dx + S4 dy + 2S5 dz)
__device__ __forceinline__ void pair_static(double dx, double dy, double dz, double q, double mx,
                                            double my, double mz, const double* Sq, double a,
                                            bool damp, double& fx, double& fy, double& fz) {
  const double r2   = dx * dx + dy * dy + dz * dz;
  const double rinv = rsqrt_f64(r2);
  const double ri2  = rinv * rinv;
  const double ri3  = rinv * ri2;
  const double ri5  = ri3 * ri2;
  const double ri7  = ri5 * ri2;
  double fE = 1.0, fT = 1.0, fD = 1.0;
  if (damp) {
    const double u = a * (r2 * rinv);
    if (u < kTholeCutoff && a > 0.0) {
      const double e = exp(-u);
      thole_ET(u, e, fE, fT);
      const double u2 = u * u;
      fD = fT - (u2 * u2 * (1.0 / 30.0)) * e;
    }
  }
  const double dmu = dx * mx + dy * my + dz * mz;
  const double P   = Sq[0] * dx * dx + Sq[1] * dx * dy + Sq[2] * dx * dz + Sq[3] * dy * dy +
                   Sq[4] * dy * dz + Sq[5] * dz * dz;
  const double tr  = Sq[0] + Sq[3] + Sq[5];
  const double qx  = 2.0 * Sq[0] * dx + Sq[1] * dy + Sq[2] * dz;
  const double qy  = Sq[1] * dx + 2.0 * Sq[3] * dy + Sq[4] * dz;
  const double qz  = Sq[2] * dx + Sq[4] * dy + 2.0 * Sq[5] * dz;
  const double cd  = q * fE * ri3 + 3.0 * fT * dmu * ri5 + 15.0 * fD * P * ri7 - 3.0 * fT * tr * ri5;
  const double cm  = -fE * ri3;
  const double cq  = -3.0 * fT * ri5;
  fx += cd * dx + cm * mx + cq * qx;
  fy += cd * dy + cm * my + cq * qy;
  fz += cd * dz + cm * mz + cq * qz;
}

// Exclusion helpers: i == j or j listed in the CSR exclusions of i (device order).
__device__ __forceinline__ bool tile_hits(int i, int j0, int j1, const int* __restrict__ eoff,
                                          const int* __restrict__ eidx) {
  bool hit = (i >= j0 && i < j1);
  if (eoff) {
    const int e0 = eoff[i], e1 = eoff[i + 1];
    for (int k = e0; k < e1; ++k) {
      const int e = eidx[k];
      hit |= (e >= j0 && e < j1);
    }
  }
  return hit;
}

__device__ __forceinline__ bool excluded(int i, int j, const int* __restrict__ eoff,
                                         const int* __restrict__ eidx) {
  if (i == j) return true;
  if (eoff) {
    const int e0 = eoff[i], e1 = eoff[i + 1];
    for (int k = e0; k < e1; ++k)
      if (eidx[k] == j) return true;
  }
  return false;
}

}  // namespace cppe_dev
