// Pair field kernels shared by the direct and FMM near-field paths.
#pragma once
#include "common.hh"

namespace cppe_dev {

// ---------------------------------------------------------------------------------------------
// Pair kernels. d = r_target - r_source.
// ---------------------------------------------------------------------------------------------
// -T2 mu with Thole damping: fE mu / R^3 - 3 fT (d.mu) d / R^5  (P2P<1,3>, field_m1.cc:127-193)
// SELF0: a zero displacement (a site with itself) contributes exactly zero.
template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void pair_m1_f64(double dx, double dy, double dz, double mx, double my,
                                            double mz, double a, double& fx, double& fy, double& fz) {
  const double r2   = dx * dx + dy * dy + dz * dz;
  const double rinv = SELF0 ? (r2 > 0.0 ? rsqrt_f64(r2) : 0.0) : rsqrt_f64(r2);
  const double ri2  = rinv * rinv;
  const double ri3  = rinv * ri2;
  double fE = 1.0, fT = 1.0;
  if (DAMP) {
    const double u = a * (r2 * rinv);
    if (u < kTholeCutoff && a > 0.0) thole_ET(u, exp(-u), fE, fT);
  }
  const double dmu = dx * mx + dy * my + dz * mz;
  const double c3  = fE * ri3;
  const double c5  = 3.0 * fT * dmu * ri3 * ri2;
  fx += c3 * mx - c5 * dx;
  fy += c3 * my - c5 * dy;
  fz += c3 * mz - c5 * dz;
}

template <bool DAMP, bool SELF0 = false>
__device__ __forceinline__ void pair_m1_f32(float dx, float dy, float dz, float mx, float my, float mz,
                                            float a, float& fx, float& fy, float& fz) {
  const float r2   = dx * dx + dy * dy + dz * dz;
  const float rinv = SELF0 ? (r2 > 0.0f ? rsqrt_f32(r2) : 0.0f) : rsqrt_f32(r2);
  const float ri2  = rinv * rinv;
  const float ri3  = rinv * ri2;
  float fE = 1.0f, fT = 1.0f;
  if (DAMP) {
    const float u = a * (r2 * rinv);
    if (u < kTholeCutoff32 && a > 0.0f) thole_ET(u, expm_f32(u), fE, fT);
  }
  const float dmu = dx * mx + dy * my + dz * mz;
  const float c3  = fE * ri3;
  const float c5  = 3.0f * fT * dmu * ri3 * ri2;
  fx += c3 * mx - c5 * dx;
  fy += c3 * my - c5 * dy;
  fz += c3 * mz - c5 * dz;
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
