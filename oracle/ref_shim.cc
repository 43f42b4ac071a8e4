// Thin C entry points onto the reference's own generated FMM kernels (fmm/field_m{0,1,2}.cc),
// compiled from /root/reference by oracle/Makefile into oracle/_ref/ (never committed, never
// shipped in the product). Test infrastructure only: used to pin the oracle's operators.
#include "field_m0.hh"
#include "field_m1.hh"
#include "field_m2.hh"

extern "C" {

void ref_p2p(int m, double x, double y, double z, double* S, double* F) {
  if (m == 0) P2P<0, 3>(x, y, z, S, F);
  if (m == 1) P2P<1, 3>(x, y, z, S, F);
  if (m == 2) P2P<2, 3>(x, y, z, S, F);
}

void ref_p2p_damped(int m, double x, double y, double z, double a, double* S, double* F) {
  if (m == 0) P2P<0, 3>(x, y, z, a, S, F);
  if (m == 1) P2P<1, 3>(x, y, z, a, S, F);
  if (m == 2) P2P<2, 3>(x, y, z, a, S, F);
}

void ref_m2m(int m, int order, double x, double y, double z, double* M, double* Ms) {
  if (m == 0) M2M<0, 3>(x, y, z, M, Ms, order);
  if (m == 1) M2M<1, 3>(x, y, z, M, Ms, order);
  if (m == 2) M2M<2, 3>(x, y, z, M, Ms, order);
}

void ref_m2l(int m, int order, double x, double y, double z, double* M, double* L) {
  if (m == 0) M2L<0, 3>(x, y, z, M, L, order);
  if (m == 1) M2L<1, 3>(x, y, z, M, L, order);
  if (m == 2) M2L<2, 3>(x, y, z, M, L, order);
}

void ref_l2l(int m, int order, double x, double y, double z, double* L, double* Ls) {
  if (m == 0) L2L<0, 3>(x, y, z, L, Ls, order);
  if (m == 1) L2L<1, 3>(x, y, z, L, Ls, order);
  if (m == 2) L2L<2, 3>(x, y, z, L, Ls, order);
}

void ref_l2p(int m, int order, double x, double y, double z, double* L, double* F) {
  if (m == 0) L2P<0, 3>(x, y, z, L, F, order);
  if (m == 1) L2P<1, 3>(x, y, z, L, F, order);
  if (m == 2) L2P<2, 3>(x, y, z, L, F, order);
}

}  // extern "C"
