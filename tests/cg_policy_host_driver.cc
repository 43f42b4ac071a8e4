// Host build of the device CG's per-segment decisions (cppe_amd/csrc/device/cg_policy.hh) for CPU
// tests: one segment at a time, tables of one segment (ns = 1). prm = [thresh, drop, min_progress,
// stall, eta, eta_gap].
#include "cg_policy.hh"

using namespace cppe_dev;

static Refine refine(const double* prm) { return Refine{prm[0], prm[1], prm[2], (int)prm[3], prm[4], prm[5]}; }

extern "C" {
void pol_sizes(int* nd, int* ni) {
  *nd = D_COUNT;
  *ni = I_COUNT;
}
void pol_init(double* sd, int* si, double rz, double rr, int mixed, const double* prm) {
  seg_init(SegRef{sd, si, 1, 0}, rz, rr, mixed != 0, refine(prm));
}
void pol_restart(double* sd, int* si, double rz, double rr, int it, const double* prm) {
  seg_restart(SegRef{sd, si, 1, 0}, rz, rr, it, refine(prm));
}
void pol_step(double* sd, int* si, double rr, int it, int mixed, const double* prm, double* hist, int cap,
              double pap, double apap) {
  seg_step(SegRef{sd, si, 1, 0}, rr, it, mixed != 0, refine(prm), Hist{hist, cap}, pap, apap);
}
void pol_true(double* sd, int* si, double rr, int it, int stepped, const double* prm, double* hist, int cap) {
  seg_true_residual(SegRef{sd, si, 1, 0}, rr, it, stepped, refine(prm), Hist{hist, cap});
}
int pol_confirm(double* sd, int* si, double rr, double yy, const double* prm, double* hist, int cap) {
  return seg_confirm32(SegRef{sd, si, 1, 0}, rr, yy, refine(prm), Hist{hist, cap});
}
void pol_beta(double* sd, int* si, double rz_new) { seg_beta(SegRef{sd, si, 1, 0}, rz_new); }
void pol_horizon(double* sd, int* si, int mixed, const double* prm, int* to_rel, int* to_done) {
  seg_horizon(SegRef{sd, si, 1, 0}, mixed != 0, refine(prm), *to_rel, *to_done);
}
int pol_far() { return kFarSteps; }
}
