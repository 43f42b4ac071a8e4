// Per-segment decisions of the device CG (InducedMoments::compute, src/core/electric_fields.cc:191-239)
// and of its FP32-refined form: what a segment does with a new residual norm. Compiled for the
// device (k_sg_final) and for the host (CPU tests of the decision sequences).
//
// A segment's scalars and flags live in two SoA tables, sd[k * ns + g] and si[k * ns + g].
//
// Plain form: the reference recurrence (one step before the first test, stop when ||r|| < thresh).
// Mixed form (FP32 applies inside FP64 iterative refinement, reliable updates after Sleijpen & van
// der Vorst, Computing 56, 141 (1996)):
//  * a recursive residual below the threshold freezes the segment until an FP64 true residual
//    confirms it (PENDING);
//  * a recursive residual that has fallen by `drop` since the last true residual asks for an FP64
//    true residual (NEEDREL); so does one that has not halved for `stall` steps -- with an FP32
//    operator the recursive residual can level off before it reaches the trigger;
//  * a true residual that is not below `min_progress` times the previous true residual after at
//    least one step marks the segment as stagnating (STAG): the refinement is not converging and
//    the solve continues from the current moments in the plain FP64 form;
//  * after an FP64 true residual the gap between true and recursive residual is bounded by
//    eta sum_k |alpha_k| ||B32 p_k|| over the steps since (each FP32 apply's error at most eta times
//    its result; residual replacement after van der Vorst & Ye, SIAM J. Sci. Comput. 22, 835
//    (2000)): a recursive residual with ||r|| + that bound below the threshold has converged for
//    the FP64 operator without a confirmation;
//  * otherwise a pending segment with an FP64 true residual r_t at moments x_t can be confirmed
//    from it: r_c = r_t - B32 (x - x_t), one FP32 apply of the correction since then, accepted
//    when ||r_c|| + eta ||B32 (x - x_t)|| < thresh; else an FP64 true residual decides as before.
//    Either form can be switched off (its eta 0); with both off an FP64 true residual confirms.
// Launch horizon: after every decision a segment also predicts how many more steps it will take
// before it needs the host (seg_horizon), from its mean log reduction per step; the host launches
// that many steps as one graph, and steps launched beyond an event do nothing on the device.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define CPPE_HD __host__ __device__
#else
#define CPPE_HD
#endif

namespace cppe_dev {

using std::fabs;
using std::isfinite;
using std::log;
using std::sqrt;

// per-segment doubles; D_RZN: the new r.z of a step, kept between the final kernel's two passes;
// D_RBEST: smallest recursive residual since the last true residual (halving steps only);
// D_LOGQ: mean log reduction of the residual norm per step (0: not known yet); D_GAP: sum of
// |alpha_k| ||B32 p_k|| since the last true residual
enum SegD { D_RZ = 0, D_BETA, D_RNORM, D_RLAST, D_RZN, D_RBEST, D_LOGQ, D_GAP, D_COUNT };
// per-segment ints. I_PENDING: frozen below the threshold until an FP64 true residual confirms it;
// I_ELIGIBLE: within 100x of the replacement trigger, replaced whenever a replacement runs for
// another segment (segments share FP64 applies); I_SINCE: steps since D_RBEST last halved;
// I_STEPS: steps since the last true residual; I_STAG: stagnating refinement; I_HASREL: an FP64
// true residual of this solve is stored (moments and residual at that point)
enum SegI { I_ACTIVE = 0, I_NEEDREL, I_ITERS, I_BAD, I_PENDING, I_ELIGIBLE, I_SINCE, I_STEPS, I_STAG, I_HASREL, I_COUNT };

struct Refine {
  double thresh;        // ||r||_2 below which a segment has converged
  double drop;          // mixed: recursive-residual drop that asks for a true residual
  double min_progress;  // mixed: a true residual above this fraction of the previous one stagnates
  int stall;            // mixed: steps without halving that ask for a true residual
  double eta;           // mixed: bound of an FP32 apply's relative error in the corrected confirmation (0: off)
  double eta_gap;       // mixed: the same bound in the gap bound of the recursive residual (0: off)
};

struct SegRef {
  double* sd;
  int* si;
  int ns, g;
  CPPE_HD double& d(int k) const { return sd[k * ns + g]; }
  CPPE_HD int& i(int k) const { return si[k * ns + g]; }
};

struct Hist {
  double* h;  // residual norms, iteration-major [it * ns + g]
  int cap;
  CPPE_HD void put(const SegRef& s, int it, double v) const {
    if (it >= 0 && it < cap) h[(size_t)it * s.ns + s.g] = v;
  }
};

// First residual of a solve: rz = r.z, rr = r.r.
CPPE_HD inline void seg_init(const SegRef& s, double rz, double rr, bool mixed, const Refine& p) {
  const double rn = sqrt(rr);
  s.d(D_RZ)    = rz;
  s.d(D_RNORM) = rn;
  s.d(D_RLAST) = rn;
  s.d(D_RBEST) = rn;
  s.d(D_BETA)  = 0.0;
  s.d(D_LOGQ)  = 0.0;
  s.d(D_GAP)   = 0.0;
  const bool bad = !isfinite(rn);
  s.i(I_BAD)      = bad;
  s.i(I_ITERS)    = 0;
  s.i(I_PENDING)  = 0;
  s.i(I_ELIGIBLE) = 0;
  s.i(I_SINCE)    = 0;
  s.i(I_STEPS)    = 0;
  s.i(I_STAG)     = 0;
  s.i(I_HASREL)   = 0;
  if (mixed) {
    // an FP32 residual within one drop of the threshold is replaced by the FP64 true residual
    // before any step (a guess that already converges stops at 0 iterations)
    s.i(I_ACTIVE)  = !bad;
    s.i(I_NEEDREL) = !bad && rn < p.thresh / p.drop;
  } else {
    // the reference takes one step before its first test (electric_fields.cc:213-229), also from
    // a guess that meets the threshold; only an exactly zero residual (where that step would
    // divide 0 by 0) stops at 0 iterations
    s.i(I_ACTIVE)  = !bad && rn > 0.0;
    s.i(I_NEEDREL) = 0;
  }
}

// Restart of an unfinished segment in the plain form from its FP64 true residual (beta = 0, p = z),
// after the mixed form stagnated; `it` is the iteration count so far. Finished segments
// (converged or bad) are left as they are.
CPPE_HD inline void seg_restart(const SegRef& s, double rz, double rr, int it, const Refine& p) {
  if (!s.i(I_ACTIVE) && !s.i(I_PENDING) && !s.i(I_NEEDREL)) return;
  const double rn = sqrt(rr);
  s.d(D_RZ)    = rz;
  s.d(D_RNORM) = rn;
  s.d(D_RLAST) = rn;
  s.d(D_RBEST) = rn;
  s.d(D_BETA)  = 0.0;
  s.i(I_PENDING)  = 0;
  s.i(I_ELIGIBLE) = 0;
  s.i(I_NEEDREL)  = 0;
  s.i(I_SINCE)    = 0;
  s.i(I_STEPS)    = 0;
  s.i(I_STAG)     = 0;
  const bool bad = !isfinite(rn);
  s.i(I_BAD) = bad;
  if (bad || rn < p.thresh) {
    s.i(I_ACTIVE) = 0;
    s.i(I_ITERS)  = it;
  } else {
    s.i(I_ACTIVE) = 1;
  }
}

// Residual norm after the step of iteration `it` (rr = r.r of the recursive residual; pap = p.Ap and
// apap = Ap.Ap of the step, mixed form).
CPPE_HD inline void seg_step(const SegRef& s, double rr, int it, bool mixed, const Refine& p, const Hist& h,
                             double pap = 0.0, double apap = 0.0) {
  if (!s.i(I_ACTIVE)) return;
  const double rn = sqrt(rr), prev = s.d(D_RNORM);
  if (isfinite(rn) && rn > 0.0 && isfinite(prev) && prev > 0.0) {  // running mean, weight 1/2
    const double l = log(rn / prev);
    s.d(D_LOGQ)    = s.d(D_LOGQ) == 0.0 ? l : 0.5 * (s.d(D_LOGQ) + l);
  }
  s.d(D_RNORM)    = rn;
  s.i(I_ELIGIBLE) = mixed && isfinite(rn) && rn < 100.0 * p.drop * s.d(D_RLAST);
  if (!isfinite(rn)) {
    s.i(I_BAD)    = 1;
    s.i(I_ACTIVE) = 0;
    s.i(I_ITERS)  = it + 1;
    h.put(s, it, rn);
    return;
  }
  if (mixed) {
    s.i(I_STEPS) += 1;
    if (p.eta_gap > 0.0) {  // alpha_k = (r.z) / (p.Ap); a step without a finite term voids the bound
      const double g = pap != 0.0 ? fabs(s.d(D_RZ) / pap) * sqrt(apap) : INFINITY;
      s.d(D_GAP)     = isfinite(g) ? s.d(D_GAP) + g : INFINITY;
    }
    if (rn < p.thresh && p.eta_gap > 0.0 && s.i(I_HASREL) && rn + p.eta_gap * s.d(D_GAP) < p.thresh) {
      h.put(s, it, rn);  // converged within the gap bound
      s.i(I_ACTIVE) = 0;
      s.i(I_ITERS)  = it + 1;
      return;
    }
    if (rn < 0.5 * s.d(D_RBEST)) {
      s.d(D_RBEST) = rn;
      s.i(I_SINCE) = 0;
    } else {
      s.i(I_SINCE) += 1;
    }
    if (rn < p.thresh) {  // frozen until an FP64 true residual confirms it; confirmations are batched
      s.i(I_PENDING) = 1;
      s.i(I_ACTIVE)  = 0;
      s.i(I_ITERS)   = it + 1;
      return;
    }
    if (rn < p.drop * s.d(D_RLAST) || s.i(I_SINCE) >= p.stall) {
      s.i(I_NEEDREL) = 1;  // its norm is recorded after the replacement
      return;
    }
  }
  h.put(s, it, rn);
  if (rn < p.thresh) {
    s.i(I_ACTIVE) = 0;
    s.i(I_ITERS)  = it + 1;
  }
}

// Progress test of a mixed-form true residual against the previous one.
CPPE_HD inline void seg_progress(const SegRef& s, double rn, const Refine& p) {
  if (s.i(I_STEPS) > 0 && isfinite(rn) && !(rn < p.thresh) && rn > p.min_progress * s.d(D_RLAST)) s.i(I_STAG) = 1;
  s.d(D_RLAST) = rn;
  s.d(D_RBEST) = rn;
  s.d(D_GAP)   = 0.0;
  s.i(I_SINCE) = 0;
  s.i(I_STEPS) = 0;
}

// FP64 true residual (rr = ||rhs - B x||^2) of the mixed form, at iteration count `it`; `stepped`:
// the current iteration's step has been taken (its direction not yet).
CPPE_HD inline void seg_true_residual(const SegRef& s, double rr, int it, int stepped, const Refine& p,
                                      const Hist& h) {
  if (s.i(I_NEEDREL) || (s.i(I_ELIGIBLE) && s.i(I_ACTIVE))) {
    const double rn = sqrt(rr);
    s.i(I_ELIGIBLE) = 0;
    s.i(I_NEEDREL)  = 0;
    s.i(I_HASREL)   = 1;
    s.d(D_RNORM)    = rn;
    seg_progress(s, rn, p);
    h.put(s, it, rn);
    if (!isfinite(rn)) {
      s.i(I_BAD)    = 1;
      s.i(I_ACTIVE) = 0;
      s.i(I_ITERS)  = it + 1;
    } else if (rn < p.thresh) {
      s.i(I_ACTIVE) = 0;
      s.i(I_ITERS)  = it + stepped;  // 0 when confirmed before any step
    }
  } else if (s.i(I_PENDING)) {
    const double rn = sqrt(rr);
    const int k0    = s.i(I_ITERS) - 1;  // iteration at which it froze
    s.d(D_RNORM)    = rn;
    s.i(I_PENDING)  = 0;
    s.i(I_HASREL)   = 1;
    seg_progress(s, rn, p);
    h.put(s, k0, rn);
    if (!isfinite(rn)) {
      s.i(I_BAD) = 1;
    } else if (!(rn < p.thresh)) {
      s.i(I_ACTIVE) = 1;  // resumes from the true residual
    }
  }
}

// Confirmation of a pending segment from its stored FP64 true residual (mixed form): rr =
// ||r_t - B32 (x - x_t)||^2, yy = ||B32 (x - x_t)||^2. Accepted (the segment has converged, its
// norm recorded at the iteration where it froze) when ||r_c|| + eta ||B32 (x - x_t)|| < thresh;
// otherwise it stays pending for an FP64 true residual. Returns whether it was tried.
CPPE_HD inline bool seg_confirm32(const SegRef& s, double rr, double yy, const Refine& p, const Hist& h) {
  if (!s.i(I_PENDING) || !s.i(I_HASREL)) return false;
  const double rn = sqrt(rr), bound = rn + p.eta * sqrt(yy);
  if (isfinite(bound) && bound < p.thresh) {
    s.i(I_PENDING) = 0;
    s.d(D_RNORM)   = rn;
    h.put(s, s.i(I_ITERS) - 1, rn);
  }
  return true;
}

// Steps an active segment is expected to take before it needs the host, from its mean log
// reduction per step, rounded down (a launch of that many steps ends at or before the event):
// to_rel, steps before it asks for a true residual (mixed form: the replacement trigger or the
// stall limit; kFarSteps when neither applies, always in the plain form), to_done, steps before it
// leaves the active set at the threshold. At least 1 each; 1 while the rate is unknown or the
// residual is not falling.
// Inactive segments: to_rel = kFarSteps, to_done = 0; a pending replacement request: both 0.
constexpr int kFarSteps = 1 << 20;
CPPE_HD inline void seg_horizon(const SegRef& s, bool mixed, const Refine& p, int& to_rel, int& to_done) {
  to_rel  = kFarSteps;
  to_done = 0;
  if (!s.i(I_ACTIVE)) return;
  if (s.i(I_NEEDREL)) {
    to_rel = 0;
    return;
  }
  const double rn = s.d(D_RNORM), q = s.d(D_LOGQ);
  if (!(q < -1e-3) || !isfinite(rn)) {
    if (mixed) to_rel = 1;
    to_done = 1;
    return;
  }
  auto steps = [&](double target) -> int {  // floor(log(target / rn) / q), at least 1
    const double k = log(target / rn) / q;
    return k >= (double)kFarSteps ? kFarSteps : (k < 1.0 ? 1 : (int)k);
  };
  to_done = steps(p.thresh);
  if (mixed) {
    const double trig = p.drop * s.d(D_RLAST);
    if (trig > p.thresh) to_rel = steps(trig);
    const int st = p.stall - s.i(I_SINCE);
    if (st < to_rel) to_rel = st < 1 ? 1 : st;
  }
}

// beta of the new direction from the new r.z
CPPE_HD inline void seg_beta(const SegRef& s, double rz_new) {
  if (!s.i(I_ACTIVE)) return;
  s.d(D_BETA) = rz_new / s.d(D_RZ);
  s.d(D_RZ)   = rz_new;
}

}  // namespace cppe_dev
