// Jacobi-preconditioned conjugate gradients on the device (InducedMoments::compute,
// src/core/electric_fields.cc:191-239), for one system or for a batch of independent snapshots
// ("segments") solved together.
//
// Vectors are 3n SoA in device order. Work is split into chunks of at most kChunk sites that
// never cross a segment boundary; every reduction is one partial per chunk (fixed block tree)
// followed by a per-segment sum over its chunks (fixed lane assignment and shuffle tree), so
// results are deterministic and a segment's scalars do not depend on the other segments. Each segment runs its own
// recurrence (its own step lengths, residual norm and convergence test) and is frozen once it
// has converged; the field applies are shared by all segments.
//
// FP64 form: the reference recurrence with every B apply in FP64.
// Mixed form: the same recurrence with FP32 field applies, made exact by reliable updates
// (residual replacement; Sleijpen & van der Vorst, Computing 56, 141 (1996)): whenever a
// segment's recursively updated residual has fallen by the system's reliable_drop factor
// (default 1e-5) since its last replacement, or below the threshold, its true residual rhs - B x is recomputed with an FP64
// apply and replaces it. Convergence is only ever declared on an FP64 true residual, so both
// forms stop on the same FP64 criterion ||rhs - B x||_2 < thresh, per segment.
//
// One iteration is three fixed launch sequences -- step (apply p, step length, x/r update,
// residual norm, and the direction when no segment asks for a replacement), reliable (FP64 apply
// of x, residual replacement; mixed form, when a segment asks for it) and direction
// (preconditioner, beta, p update, after a replacement). Each is captured once into a HIP graph
// and replayed; all scalars and decisions live on the device, and the host reads a small status
// block (published to host-mapped memory by every final kernel) after a graph to decide what to
// launch next. Several steps run as one graph: each final kernel also publishes a horizon, the
// number of steps every segment can take before one of them needs the host (a replacement, or
// no segment left active), predicted from the segments' residual reduction rates (cg_policy.hh);
// the host launches a graph of up to that many steps (1 to 8), and a step that starts after
// such an event leaves every vector and scalar as it is (its apply result is never read).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <vector>

#include "cg_policy.hh"
#include "pairs.hh"
#include "system.hh"

namespace cppe_dev {

void apply_b_fmm(System& S, const double* xin, double* out, bool fp32);
void fmm_reserve(System& S);

constexpr int kChunk = 256;  // one site per thread: enough blocks in flight to stream the vectors

// status block: [iteration, active segments, segments asking for a reliable update, bad,
// step taken in the current iteration, segments pending an FP64 confirmation, stagnating segments
// (mixed form), steps that can be launched before the host is needed (0: none active), number of
// final kernels run (each final kernel publishes the block to a host-mapped mirror, this count last)]
enum Status { ST_IT = 0, ST_ACTIVE, ST_REL, ST_BAD, ST_STEPPED, ST_PEND, ST_STAG, ST_SAFE, ST_SEQ, ST_COUNT };

// A step that starts after the host is needed (a segment asked for a replacement, or no segment
// is active) does nothing: launches of several steps end with such steps when a prediction fell
// short. Read from the status block the previous final kernel wrote.
__device__ __forceinline__ bool step_halted(const int* status) {
  return status[ST_REL] > 0 || status[ST_ACTIVE] == 0;
}
// solve parameters on the device: [thresh, reliable drop, min progress, stall steps, restart,
// eta of the FP32-corrected confirmation, eta of the gap bound]
enum Param { P_THRESH = 0, P_DROP, P_PROGRESS, P_STALL, P_RESTART, P_ETA, P_GAP, P_COUNT };
// F_RNB: the residual norm and its tests, then -- unless a segment asked for a reliable update --
// F_BETA's update, in one launch; F_CONF: FP32-corrected confirmation of pending segments
enum FinalMode { F_INIT = 0, F_REL, F_BETA, F_RNB, F_CONF };

// Both summations apply B through the tree path: a direct-summation system is a tree of one leaf
// whose near list is itself (every pair, exclusions and Thole damping as in the direct sum).
void apply_b(System& S, const double* x, double* out, bool fp32) { apply_b_fmm(S, x, out, fp32); }

// out = alpha in (3x3 packed SoA per site)
__global__ __launch_bounds__(256) void k_alpha_apply(int n, const double* __restrict__ a6,
                                                     const double* __restrict__ in, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m0 = in[i], m1 = in[n + i], m2 = in[2 * n + i];
  const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
               a5 = a6[5 * n + i];
  out[i]         = a0 * m0 + a1 * m1 + a2 * m2;
  out[n + i]     = a1 * m0 + a3 * m1 + a4 * m2;
  out[2 * n + i] = a2 * m0 + a4 * m1 + a5 * m2;
}

void precondition(System& S, const double* in, double* out) {
  hipLaunchKernelGGL(k_alpha_apply, dim3(ceil_div(S.n, 256)), dim3(256), 0, S.stream, S.n, S.alpha6.p, in, out);
}

__global__ void k_fill_f64(size_t n, double* __restrict__ a, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

void fill_f64(System& S, double* a, size_t n, double v) {
  if (n) hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div((long long)n, 256)), dim3(256), 0, S.stream, n, a, v);
}

// Lane l's share of the chunk partials [lo, hi) of NA arrays: its chunks c_m = lo + l + 64 m (m < M)
// summed into four interleaved accumulators, a_(m mod 4) for m < 4 floor(M / 4) and a_0 for the
// rest, each in increasing m, combined as (a0 + a1) + (a2 + a3) -- one fixed order. The loads of
// B chunks of every array are issued before any sum, with addresses clamped into [lo, hi) and the
// values past a lane's last chunk replaced by 0 (no branch between the loads), so a segment of up
// to 64 B chunks waits for memory once. The 64 lane shares are then combined by an xor-shuffle
// tree.
template <int NA, int B>
__device__ __forceinline__ void lane_chunk_sums(const double* const (&pa)[NA], int lo, int hi, int lane,
                                                double (&res)[NA]) {
  static_assert(B % 4 == 0, "batches start at a multiple of 4");
  const int M  = hi - lo > lane ? (hi - lo - lane + 63) / 64 : 0;
  const int M4 = M & ~3;
  const int Mu = (hi - lo + 63) / 64;  // lane 0's count: the same trip count for every lane
  double a[NA][4];
#pragma unroll
  for (int k = 0; k < NA; ++k) a[k][0] = a[k][1] = a[k][2] = a[k][3] = 0.0;
  for (int m0 = 0; m0 < Mu; m0 += B) {
    double x[NA][B];
#pragma unroll
    for (int t = 0; t < B; ++t) {
      const int c = min(lo + lane + 64 * (m0 + t), hi - 1);
#pragma unroll
      for (int k = 0; k < NA; ++k) x[k][t] = pa[k][c];
    }
#pragma unroll
    for (int t = 0; t < B; ++t) {
      const int m = m0 + t;
#pragma unroll
      for (int k = 0; k < NA; ++k) {
        a[k][t & 3] += m < M4 ? x[k][t] : 0.0;
        a[k][0] += m >= M4 && m < M ? x[k][t] : 0.0;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NA; ++k) res[k] = (a[k][0] + a[k][1]) + (a[k][2] + a[k][3]);
}

__device__ __forceinline__ void z_of(const double* __restrict__ a6, int n, int i, double r0, double r1, double r2,
                                     double& z0, double& z1, double& z2) {
  const double a0 = a6[i], a1 = a6[n + i], a2 = a6[2 * n + i], a3 = a6[3 * n + i], a4 = a6[4 * n + i],
               a5 = a6[5 * n + i];
  z0 = a0 * r0 + a1 * r1 + a2 * r2;
  z1 = a1 * r0 + a3 * r1 + a4 * r2;
  z2 = a2 * r0 + a4 * r1 + a5 * r2;
}

struct CgVecs {
  int na, spr, nseg, nchunk;  // sites per right-hand side, segments per right-hand side
  const int* clo;   // nchunk + 1 offsets of the chunks (sites of the stacked right-hand sides)
  const int* cseg;  // segment of each chunk
  const int* seg_chunk;  // nseg + 1 chunk offsets of the segments
  const double* a6;
  const double* rhs;
  double *x, *r, *z, *p, *Ap, *t;
  double *xrel, *rrel, *dv;  // mixed form: moments and residual at the last FP64 true residual, correction
  double* part;     // 4 * nchunk partials
  double* sd;       // D_COUNT * nseg
  int* si;          // I_COUNT * nseg
  const int* status;  // ST_COUNT ints
  float4* rec32;      // when set: near-field dipole records of p for the next FP32 apply
  double4* rec64;     // when set: the same for the next FP64 apply
  // local preconditioner (pc_off set): the kernels that form z write w = alpha r instead (one
  // double4 record per site, stacked numbering), and k_sg_prec forms z = w + weight alpha T_loc w
  // and the partial r.z
  double4* w;
  const int* pc_off;
  const int* pc_nj;
  const double4* pc_site;  // x, y, z, Thole scale per site
  double pc_weight, pc_damping;
};

// p of site i (components p0, p1, p2) into the near-field dipole record the next apply reads
__device__ __forceinline__ void put_record(const CgVecs& v, int i, double p0, double p1, double p2) {
  if (v.rec32) v.rec32[i] = make_float4((float)p0, (float)p1, (float)p2, 0.0f);
  if (v.rec64) v.rec64[i] = make_double4(p0, p1, p2, 0.0);
}

// Vectors are [rhs][component][site]; site i of chunk c (stacked numbering) has component q at
// base(c) + q * na + i and its polarizability at i - rhs(c) * na.
struct ChunkIdx {
  size_t b;
  int ia;
  __device__ ChunkIdx(const CgVecs& v, int c) {
    const int kr = v.cseg[c] / v.spr;
    b            = 2 * (size_t)kr * v.na;
    ia           = -kr * v.na;
  }
};

// r = rhs - t ; z = alpha r ; p = z ; partials (r.z, r.r)
__global__ __launch_bounds__(256) void k_sg_init(CgVecs v) {
  __shared__ double red[4];
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  double rz = 0.0, rr = 0.0;
  for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x) {
    const size_t k0 = ci.b + i, k1 = k0 + n, k2 = k1 + n;
    const double r0 = v.rhs[k0] - v.t[k0], r1 = v.rhs[k1] - v.t[k1], r2 = v.rhs[k2] - v.t[k2];
    double z0, z1, z2;
    z_of(v.a6, n, i + ci.ia, r0, r1, r2, z0, z1, z2);
    v.r[k0] = r0;
    v.r[k1] = r1;
    v.r[k2] = r2;
    rr += r0 * r0 + r1 * r1 + r2 * r2;
    if (v.pc_off) {  // z, p, the records and r.z follow in k_sg_prec
      v.w[i] = make_double4(z0, z1, z2, 0.0);
      continue;
    }
    v.z[k0] = z0;
    v.z[k1] = z1;
    v.z[k2] = z2;
    v.p[k0] = z0;
    v.p[k1] = z1;
    v.p[k2] = z2;
    put_record(v, i + ci.ia, z0, z1, z2);
    rz += r0 * z0 + r1 * z1 + r2 * z2;
  }
  const double s1 = block_sum(rz, red);
  const double s2 = block_sum(rr, red);
  if (threadIdx.x == 0) {
    if (!v.pc_off) v.part[c] = s1;
    v.part[v.nchunk + c] = s2;
  }
}

// partial p.Ap (active segments), into the third block of the partials (read by k_sg_xrz), and
// partial Ap.Ap into the fourth (the gap bound of the mixed form)
__global__ __launch_bounds__(256) void k_sg_pap(CgVecs v) {
  __shared__ double red[4];
  if (step_halted(v.status)) return;
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  double s = 0.0, q = 0.0;
  if (v.si[I_ACTIVE * v.nseg + v.cseg[c]])
    for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x) {
      const size_t k0 = ci.b + i, k1 = k0 + n, k2 = k1 + n;
      const double a0 = v.Ap[k0], a1 = v.Ap[k1], a2 = v.Ap[k2];
      s += v.p[k0] * a0 + v.p[k1] * a1 + v.p[k2] * a2;
      q += a0 * a0 + a1 * a1 + a2 * a2;
    }
  s = block_sum(s, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    v.part[2 * (size_t)v.nchunk + c] = s;
    v.part[3 * (size_t)v.nchunk + c] = q;
  }
}

// Step update and preconditioner in one pass (active segments): x += alpha p ; r -= alpha Ap ; z = alpha r ;
// partials r.r (part[c]) and r.z (part[nchunk + c]), each summed as in the separate kernels. The
// z of a segment that converges or asks for a reliable update here is not used (the direction
// after a reliable update recomputes it).
__global__ __launch_bounds__(kChunk) void k_sg_xrz(CgVecs v) {
  __shared__ double red[4];
  __shared__ double alpha;
  if (step_halted(v.status)) return;
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na, g = v.cseg[c];
  const ChunkIdx ci(v, c);
  const bool active = v.si[I_ACTIVE * v.nseg + g];
  // one site per thread (chunks hold at most kChunk sites): its vectors are loaded before alpha
  // is known, so the loads overlap the sum below
  const int i     = v.clo[c] + threadIdx.x;
  const bool site = active && i < v.clo[c + 1];
  double xq[3], pq[3], rq[3], aq[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const size_t k = ci.b + (size_t)q * n + (site ? i : v.clo[c]);
    xq[q] = site ? v.x[k] : 0.0;
    pq[q] = site ? v.p[k] : 0.0;
    rq[q] = site ? v.r[k] : 0.0;
    aq[q] = site ? v.Ap[k] : 0.0;
  }
  // alpha = (r.z) / (p.Ap) of the segment, from k_sg_pap's partials: every block sums them in one
  // fixed order (lane_chunk_sums, then an xor-shuffle tree), so all blocks of a segment take the
  // same alpha, without a launch of its own
  if (active && threadIdx.x < 64) {
    const double* const pa[1] = {v.part + 2 * (size_t)v.nchunk};
    double sum[1];
    lane_chunk_sums<1, 8>(pa, v.seg_chunk[g], v.seg_chunk[g + 1], threadIdx.x, sum);
    double s1 = sum[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s1 += __shfl_xor(s1, off);
    if (threadIdx.x == 0) alpha = v.sd[D_RZ * v.nseg + g] / s1;
  }
  __syncthreads();
  double srr = 0.0, srz = 0.0;
  if (site) {
    const double al = alpha;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const size_t k = ci.b + (size_t)q * n + i;
      v.x[k] = xq[q] + al * pq[q];
      rq[q]  = rq[q] - al * aq[q];
      v.r[k] = rq[q];
      srr += rq[q] * rq[q];
    }
    double z0, z1, z2;
    z_of(v.a6, n, i + ci.ia, rq[0], rq[1], rq[2], z0, z1, z2);
    const size_t k0 = ci.b + i;
    if (v.pc_off) {  // local preconditioner: w, then k_sg_prec
      v.w[i] = make_double4(z0, z1, z2, 0.0);
    } else {
      v.z[k0]                 = z0;
      v.z[k0 + n]             = z1;
      v.z[k0 + 2 * (size_t)n] = z2;
      srz += rq[0] * z0 + rq[1] * z1 + rq[2] * z2;
    }
  }
  const double s1 = block_sum(srr, red);
  const double s2 = block_sum(srz, red);
  if (threadIdx.x == 0) {
    v.part[c] = s1;
    if (!v.pc_off) v.part[v.nchunk + c] = s2;
  }
}

// r = rhs - t for the segments that asked for a reliable update or await confirmation; partial r.r.
// With xrel set, the moments and the true residual are also kept for later FP32-corrected
// confirmations.
__global__ __launch_bounds__(256) void k_sg_resid(CgVecs v) {
  __shared__ double red[4];
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na, g = v.cseg[c];
  const ChunkIdx ci(v, c);
  double s = 0.0;
  if (v.si[I_NEEDREL * v.nseg + g] || v.si[I_PENDING * v.nseg + g] || v.si[I_ELIGIBLE * v.nseg + g])
    for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const size_t k  = ci.b + (size_t)q * n + i;
        const double rk = v.rhs[k] - v.t[k];
        v.r[k] = rk;
        if (v.xrel) {
          v.xrel[k] = v.x[k];
          v.rrel[k] = rk;
        }
        s += rk * rk;
      }
  s = block_sum(s, red);
  if (threadIdx.x == 0) v.part[c] = s;
}

__device__ __forceinline__ bool conf_seg(const CgVecs& v, int g) {
  return v.si[I_PENDING * v.nseg + g] && v.si[I_HASREL * v.nseg + g];
}

// dv = x - xrel for pending segments with a stored true residual, 0 elsewhere
__global__ __launch_bounds__(256) void k_sg_defect(CgVecs v) {
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  const bool on = conf_seg(v, v.cseg[c]);
  for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const size_t k = ci.b + (size_t)q * n + i;
      v.dv[k]        = on ? v.x[k] - v.xrel[k] : 0.0;
    }
}

// partials of ||rrel - t||^2 (part[c]) and ||t||^2 (part[nchunk + c]), t = B32 dv, for the same segments
__global__ __launch_bounds__(256) void k_sg_confirm(CgVecs v) {
  __shared__ double red[4];
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  double s1 = 0.0, s2 = 0.0;
  if (conf_seg(v, v.cseg[c]))
    for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const size_t k  = ci.b + (size_t)q * n + i;
        const double tk = v.t[k], rk = v.rrel[k] - tk;
        s1 += rk * rk;
        s2 += tk * tk;
      }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) {
    v.part[c]            = s1;
    v.part[v.nchunk + c] = s2;
  }
}

// The direction (beta in F_RNB, k_sg_p) follows the step in the same launch sequence and does
// nothing while a segment waits for a reliable update; the host then runs the direction kernels
// (k_sg_z, F_BETA, k_sg_p, which test the same condition) after the update.

// z = alpha r ; partial z.r (active segments)
__global__ __launch_bounds__(256) void k_sg_z(CgVecs v) {
  __shared__ double red[4];
  if (v.status[ST_REL] > 0) return;
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  double s = 0.0;
  if (v.si[I_ACTIVE * v.nseg + v.cseg[c]])
    for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x) {
      const size_t k0 = ci.b + i, k1 = k0 + n, k2 = k1 + n;
      const double r0 = v.r[k0], r1 = v.r[k1], r2 = v.r[k2];
      double z0, z1, z2;
      z_of(v.a6, n, i + ci.ia, r0, r1, r2, z0, z1, z2);
      if (v.pc_off) {  // local preconditioner: w, then k_sg_prec
        v.w[i] = make_double4(z0, z1, z2, 0.0);
        continue;
      }
      v.z[k0] = z0;
      v.z[k1] = z1;
      v.z[k2] = z2;
      s += r0 * z0 + r1 * z1 + r2 * z2;
    }
  s = block_sum(s, red);
  if (threadIdx.x == 0 && !v.pc_off) v.part[c] = s;
}

// y += T2d(d) w_j for one close pair, d = r_i - r_j: T2d = 3 fT d d^T / R^5 - fE / R^3 (Thole
// damped, field_m1.cc:127-193 for one pair). The scalars fE / R^3 and 3 fT / R^5 come from FP32
// (displacement rounded to FP32, damping product s_i s_j taken in that order-free form), so the
// pair (j, i) gets exactly the same two scalars as (i, j) and T_loc stays symmetric; d and the
// products with w are FP64. The preconditioner only has to be one fixed symmetric operator: its
// blocks agree with the FP64 tensor to about 1e-6 relative.
__device__ __forceinline__ void pc_pair(const double4& pi, const double4& pj, double damping, const double4& wj,
                                        double& y0, double& y1, double& y2) {
  const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  float c3, cf;
  geo_m1_f32<true>((float)dx, (float)dy, (float)dz, (float)(damping * (pi.w * pj.w)), c3, cf);
  const double c5 = (double)cf * (dx * wj.x + dy * wj.y + dz * wj.z), e3 = (double)c3;
  y0 += c5 * dx - e3 * wj.x;
  y1 += c5 * dy - e3 * wj.y;
  y2 += c5 * dz - e3 * wj.z;
}

// Local preconditioner, after a kernel that wrote w = alpha r: z = w + weight alpha (T_loc w) and
// the chunk partial r.z into part[slot * nchunk + c]. MODE 0 (after k_sg_init, every segment):
// also p = z and p's near-field records. MODE 1 (after k_sg_xrz) and MODE 2 (after k_sg_z):
// active segments, MODE 1 not in a halted step, MODE 2 not while a reliable update is pending.
// One thread per site walks its pair row in order (a handful of pairs: four in flight at a
// time), so every site of the chunk gathers at once; the sum over a row has a fixed order. Each
// pair's block is formed here from the two site records (pc_pair), not read from memory.
template <int MODE>
__global__ __launch_bounds__(256) void k_sg_prec(CgVecs v, int slot) {
  __shared__ double red[4];
  if ((MODE == 1 && step_halted(v.status)) || (MODE == 2 && v.status[ST_REL] > 0)) return;
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na;
  const ChunkIdx ci(v, c);
  double s = 0.0;
  if (MODE == 0 || v.si[I_ACTIVE * v.nseg + v.cseg[c]])
    for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x) {
      const int ia = i + ci.ia;  // site within its right-hand side
      double y0 = 0.0, y1 = 0.0, y2 = 0.0;
      const int q0 = v.pc_off[ia], q1 = v.pc_off[ia + 1];
      const double4 pi = v.pc_site[ia];
#pragma unroll 4
      for (int q = q0; q < q1; ++q) {
        const int j = v.pc_nj[q];
        pc_pair(pi, v.pc_site[j], v.pc_damping, v.w[j - ci.ia], y0, y1, y2);
      }
      double a0, a1, a2;
      z_of(v.a6, n, ia, y0, y1, y2, a0, a1, a2);
      const double4 wi = v.w[i];
      const size_t k0 = ci.b + i, k1 = k0 + n, k2 = k1 + n;
      const double z0 = wi.x + v.pc_weight * a0, z1 = wi.y + v.pc_weight * a1, z2 = wi.z + v.pc_weight * a2;
      v.z[k0] = z0;
      v.z[k1] = z1;
      v.z[k2] = z2;
      if (MODE == 0) {
        v.p[k0] = z0;
        v.p[k1] = z1;
        v.p[k2] = z2;
        put_record(v, ia, z0, z1, z2);
      }
      s += v.r[k0] * z0 + v.r[k1] * z1 + v.r[k2] * z2;
    }
  s = block_sum(s, red);
  if (threadIdx.x == 0) v.part[(size_t)slot * v.nchunk + c] = s;
}

// p = z + beta p (active segments)
__global__ __launch_bounds__(256) void k_sg_p(CgVecs v) {
  if (v.status[ST_REL] > 0) return;
  const int c = xcd_chunk(blockIdx.x, v.nchunk), n = v.na, g = v.cseg[c];
  if (!v.si[I_ACTIVE * v.nseg + g]) return;
  const ChunkIdx ci(v, c);
  const double be = v.sd[D_BETA * v.nseg + g];
  for (int i = v.clo[c] + threadIdx.x; i < v.clo[c + 1]; i += blockDim.x) {
    double pq[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const size_t k = ci.b + (size_t)q * n + i;
      pq[q]          = v.z[k] + be * v.p[k];
      v.p[k]         = pq[q];
    }
    put_record(v, i + ci.ia, pq[0], pq[1], pq[2]);
  }
}


struct CgFinal {
  int nseg, nchunk;
  const int* seg_chunk;  // nseg + 1 chunk offsets of the segments
  const double* part;
  double* sd;
  int* si;
  int* status;           // ST_COUNT ints
  int* hstat;            // host-mapped mirror of the status block
  const double* params;  // P_COUNT values
  double* hist;          // residual norms, iteration-major [it * nseg + g]
  int hist_cap;          // iterations held by hist
  int publish;           // copy the status block to the host mirror (0: steps inside a launch but its last)
};

// Per-segment finalisation of the chunk partials: one wavefront per segment (the block's four
// wavefronts take segments in turn); each lane sums its chunks (lane_chunk_sums), then a fixed
// shuffle tree combines the lanes -- deterministic for a given chunk split. Lane 0 then applies
// the segment's decision (cg_policy.hh); the status block is refreshed at the end.
// F_INIT with params[P_RESTART] set: the plain-form restart of unfinished segments after the
// mixed form stagnated (iteration count and history kept). Every mode also publishes the launch
// horizon ST_SAFE: the smallest number of steps before a segment asks for a replacement, capped by
// the largest number before every segment has left the active set.
// Up to kStageSeg segments, the segment tables are copied into LDS first (all threads, one load
// each) and written back at the end: the decisions read and write them one field after another.
constexpr int kStageSeg = 256;
template <int MODE, bool MIXED>
__global__ __launch_bounds__(256) void k_sg_final(CgFinal f) {
  __shared__ int cnt[7];
  __shared__ double stage_d[D_COUNT * kStageSeg];
  __shared__ int stage_i[I_COUNT * kStageSeg];
  const int seq0 = f.status[ST_SEQ];
  // F_BETA while a reliable update is pending and F_RNB of a halted step: no update, only the
  // publication below
  const bool skip = (MODE == F_BETA && f.status[ST_REL] > 0) || (MODE == F_RNB && step_halted(f.status));
  if (threadIdx.x < 7) cnt[threadIdx.x] = threadIdx.x == 5 ? kFarSteps : 0;
  const int ns      = skip ? 0 : f.nseg;
  const bool staged = ns <= kStageSeg;
  double* const sd  = staged ? stage_d : f.sd;
  int* const si     = staged ? stage_i : f.si;
  if (staged) {
    for (int k = threadIdx.x; k < D_COUNT * ns; k += blockDim.x) stage_d[k] = f.sd[k];
    for (int k = threadIdx.x; k < I_COUNT * ns; k += blockDim.x) stage_i[k] = f.si[k];
  }
  __syncthreads();
  const bool restart = MODE == F_INIT && f.params[P_RESTART] != 0.0;
  // iterations so far; a restart also counts the step taken before the stagnating replacement
  const int it = f.status[ST_IT] + (restart ? f.status[ST_STEPPED] : 0);
  const Refine pr{f.params[P_THRESH], f.params[P_DROP], f.params[P_PROGRESS], (int)f.params[P_STALL],
                  f.params[P_ETA], f.params[P_GAP]};
  const Hist h{f.hist, f.hist_cap};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int na = 0, nr = 0, nb = 0, np = 0, nst = 0, h_rel = kFarSteps, h_done = 0;
  for (int g = wv; g < ns; g += 4) {
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    constexpr bool TWO = MODE == F_INIT || MODE == F_RNB || MODE == F_CONF;
    // mixed steps with the gap bound also need p.Ap and Ap.Ap of the step (k_sg_pap's partials)
    const bool gap = MIXED && MODE == F_RNB && pr.eta_gap > 0.0;
    const int lo = f.seg_chunk[g], hi = f.seg_chunk[g + 1];
    const size_t nc = f.nchunk;
    if (gap) {  // (a mixed F_RNB step: TWO holds)
      const double* const pa[4] = {f.part, f.part + nc, f.part + 2 * nc, f.part + 3 * nc};
      double sum[4];
      lane_chunk_sums<4, 16>(pa, lo, hi, lane, sum);
      s1 = sum[0], s2 = sum[1], s3 = sum[2], s4 = sum[3];
    } else if (TWO) {
      const double* const pa[2] = {f.part, f.part + nc};
      double sum[2];
      lane_chunk_sums<2, 16>(pa, lo, hi, lane, sum);
      s1 = sum[0], s2 = sum[1];
    } else {
      const double* const pa[1] = {f.part};
      double sum[1];
      lane_chunk_sums<1, 16>(pa, lo, hi, lane, sum);
      s1 = sum[0];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      s1 += __shfl_xor(s1, off);
      if (TWO) s2 += __shfl_xor(s2, off);
      if (gap) {
        s3 += __shfl_xor(s3, off);
        s4 += __shfl_xor(s4, off);
      }
    }
    if (lane != 0) continue;

    const SegRef sg{sd, si, ns, g};
    if (MODE == F_INIT) {
      if (restart)
        seg_restart(sg, s1, s2, it, pr);
      else
        seg_init(sg, s1, s2, MIXED, pr);
    } else if (MODE == F_RNB) {
      sg.d(D_RZN) = s2;  // read back by this thread after the barrier
      seg_step(sg, s1, it, MIXED, pr, h, s3, s4);
    } else if (MODE == F_REL) {
      seg_true_residual(sg, s1, it, f.status[ST_STEPPED], pr, h);
    } else if (MODE == F_CONF) {
      seg_confirm32(sg, s1, s2, pr, h);
    } else {  // F_BETA
      seg_beta(sg, s1);
    }
    na += sg.i(I_ACTIVE);
    nr += sg.i(I_NEEDREL);
    nb += sg.i(I_BAD);
    np += sg.i(I_PENDING);
    nst += sg.i(I_STAG);
    int to_rel, to_done;
    seg_horizon(sg, MIXED, pr, to_rel, to_done);
    h_rel  = min(h_rel, to_rel);
    h_done = max(h_done, to_done);
  }
  if (na) atomicAdd(&cnt[0], na);
  if (nr) atomicAdd(&cnt[1], nr);
  if (nb) atomicAdd(&cnt[2], nb);
  if (np) atomicAdd(&cnt[3], np);
  if (nst) atomicAdd(&cnt[4], nst);
  if (h_rel < kFarSteps) atomicMin(&cnt[5], h_rel);
  if (h_done) atomicMax(&cnt[6], h_done);
  __syncthreads();
  const bool beta_now = MODE == F_RNB && cnt[1] == 0;  // no reliable update pending: direction now
  if (beta_now && lane == 0)
    for (int g = wv; g < ns; g += 4) {
      const SegRef sg{sd, si, ns, g};
      seg_beta(sg, sg.d(D_RZN));
    }
  if (threadIdx.x == 0 && !skip) {
    f.status[ST_ACTIVE] = cnt[0];
    f.status[ST_REL]    = cnt[1];
    f.status[ST_BAD]    = cnt[2];
    f.status[ST_PEND]   = cnt[3];
    f.status[ST_STAG]   = cnt[4];
    f.status[ST_SAFE]   = cnt[0] == 0 ? 0 : min(cnt[5], cnt[6]);
    if (MODE == F_INIT) {
      f.status[ST_IT]      = restart ? it : 0;
      f.status[ST_STEPPED] = 0;
    }
    if (MODE == F_RNB && !beta_now) f.status[ST_STEPPED] = 1;
    if (beta_now) {  // the step and its direction: one more iteration counted (as F_BETA does)
      f.status[ST_IT]      = it + 1;
      f.status[ST_STEPPED] = 0;
    }
    if (MODE == F_BETA) {  // counts iterations that took a step (not a resumption after confirmation)
      f.status[ST_IT]      = it + f.status[ST_STEPPED];
      f.status[ST_STEPPED] = 0;
    }
  }
  if (staged) {
    __syncthreads();
    for (int k = threadIdx.x; k < D_COUNT * ns; k += blockDim.x) f.sd[k] = stage_d[k];
    for (int k = threadIdx.x; k < I_COUNT * ns; k += blockDim.x) f.si[k] = stage_i[k];
  }
  if (threadIdx.x == 0) {  // publish: the fields, a system-scope fence, then the count
    const int seq    = seq0 + 1;
    f.status[ST_SEQ] = seq;
    if (f.publish) {
      for (int q = 0; q < ST_SEQ; ++q) f.hstat[q] = f.status[q];
      __threadfence_system();
      f.hstat[ST_SEQ] = seq;
    }
  }
}

// ---- host side -------------------------------------------------------------------------------
void CgState::release_graphs() {
  auto drop = [](hipGraphExec_t& g) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  };
  for (auto& form : g_step)
    for (auto& g : form) drop(g);
  drop(g_rel);
  drop(g_conf);
  drop(g_dir[0]);
  drop(g_dir[1]);
  drop(g_init[0]);
  drop(g_init[1]);
}

CgState::~CgState() {
  release_graphs();
  if (h_status) (void)hipHostFree(h_status);
  if (h_out) (void)hipHostFree(h_out);
}

// Chunks of <= kChunk sites inside each segment (site ranges seg_off, stacked numbering) and the
// per-segment scalar / status storage.
static void cg_layout(System& S, CgState& C, const std::vector<int>& seg_off) {
  const int ns = (int)seg_off.size() - 1;
  std::vector<int> clo, cseg, schunk(ns + 1, 0);
  for (int g = 0; g < ns; ++g) {
    schunk[g] = (int)cseg.size();
    for (int i = seg_off[g]; i < seg_off[g + 1]; i += kChunk) {
      clo.push_back(i);
      cseg.push_back(g);
    }
  }
  schunk[ns] = (int)cseg.size();
  clo.push_back(seg_off[ns]);
  C.release_graphs();
  C.nseg   = ns;
  C.nchunk = (int)cseg.size();
  auto up  = [&](dbuf<int>& b, const std::vector<int>& h) {
    b.alloc(h.size());
    CPPE_HIP_CHECK(hipMemcpyAsync(b.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, S.stream));
  };
  up(C.chunk_lo, clo);
  up(C.chunk_seg, cseg);
  up(C.seg_chunk, schunk);
  C.part.alloc(4 * (size_t)C.nchunk);  // [r.r or p.Ap | r.z | p.Ap of the step | Ap.Ap of the step]

  C.sd.alloc(D_COUNT * (size_t)ns);
  C.si.alloc(I_COUNT * (size_t)ns);
  C.status.alloc(ST_COUNT);
  C.params.alloc(P_COUNT);
  C.params_ok = false;
  C.hist_cap = 0;
  C.hist.release();
  if (!C.h_status) {
    CPPE_HIP_CHECK(hipHostMalloc((void**)&C.h_status, sizeof(int) * ST_COUNT,
                                 hipHostMallocMapped | hipHostMallocCoherent));
    CPPE_HIP_CHECK(hipHostGetDevicePointer((void**)&C.d_hstat, C.h_status, 0));
  }
  CPPE_HIP_CHECK(hipMemsetAsync(C.status.p, 0, sizeof(int) * ST_COUNT, S.stream));
  CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
  for (int q = 0; q < ST_COUNT; ++q) C.h_status[q] = 0;
  C.expect = 0;
}

void setup_segments(System& S, const std::vector<int>& seg_off, const std::vector<int>& seg_ncrit) {
  const int n = S.n, ns = (int)seg_off.size() - 1;
  S.nseg      = ns;
  S.h_seg_off = seg_off;
  std::vector<int> of(n);
  for (int g = 0; g < ns; ++g)
    for (int i = seg_off[g]; i < seg_off[g + 1]; ++i) of[i] = g;
  auto up = [&](dbuf<int>& b, const std::vector<int>& h) {
    b.alloc(h.size());
    CPPE_HIP_CHECK(hipMemcpyAsync(b.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, S.stream));
  };
  up(S.seg_off, seg_off);
  up(S.seg_ncrit, seg_ncrit);
  up(S.seg_of, of);
  S.cg.na   = n;
  S.cg.nrhs = 1;
  S.cg.spr  = ns;
  cg_layout(S, S.cg, seg_off);
}

void setup_multi_rhs(System& S, int nrhs) {
  if (S.nseg != 1) throw hip_error("several right-hand sides need a one-snapshot system");
  CgState& C = S.cgm;
  if (C.nrhs == nrhs && C.na == S.n && C.x) return;
  const int n = S.n;
  std::vector<int> off(nrhs + 1);
  for (int k = 0; k <= nrhs; ++k) off[k] = k * n;
  C.na   = n;
  C.nrhs = nrhs;
  C.spr  = 1;
  cg_layout(S, C, off);
  const size_t nv = 3 * (size_t)n * nrhs;
  C.store.alloc(7 * nv);
  double* b = C.store.p;
  C.x   = b;
  C.rhs = b + nv;
  C.r   = b + 2 * nv;
  C.z   = b + 3 * nv;
  C.p   = b + 4 * nv;
  C.Ap  = b + 5 * nv;
  C.t   = b + 6 * nv;
}

static void bind_single(System& S) {
  CgState& C = S.cg;
  C.x        = S.v_mu.p;
  C.rhs      = S.v_rhs.p;
  C.r        = S.v_r.p;
  C.z        = S.v_z.p;
  C.p        = S.v_p.p;
  C.Ap       = S.v_Ap.p;
  C.t        = S.v_t.p;
}

static CgVecs vecs(System& S, CgState& C) {
  CgVecs v;
  v.na     = C.na;
  v.spr    = C.spr;
  v.nseg   = C.nseg;
  v.nchunk = C.nchunk;
  v.clo    = C.chunk_lo.p;
  v.cseg   = C.chunk_seg.p;
  v.seg_chunk = C.seg_chunk.p;
  v.a6     = S.alpha6.p;
  v.rhs    = C.rhs;
  v.x      = C.x;
  v.r      = C.r;
  v.z      = C.z;
  v.p      = C.p;
  v.Ap     = C.Ap;
  v.t      = C.t;
  v.xrel   = C.xrel;
  v.rrel   = C.rrel;
  v.dv     = C.dv;
  v.part   = C.part.p;
  v.sd     = C.sd.p;
  v.si     = C.si.p;
  v.status = C.status.p;
  v.rec32  = C.rec32;
  v.rec64  = C.rec64;
  const bool pc = S.precond == 1 && S.pc;
  v.w         = pc ? reinterpret_cast<double4*>(C.wbuf.p) : nullptr;
  v.pc_off    = pc ? S.pc->off.p : nullptr;
  v.pc_nj     = pc ? S.pc->nj.p : nullptr;
  v.pc_site    = pc ? S.pc->site.p : nullptr;
  v.pc_weight  = S.precond_weight;
  v.pc_damping = pc ? S.pc->damping : 0.0;
  return v;
}

// z from w for the local preconditioner (MODE as in k_sg_prec); nothing for the Jacobi form
template <int MODE>
static void launch_prec(System& S, CgState& C, int slot) {
  if (S.precond != 1) return;
  hipLaunchKernelGGL(k_sg_prec<MODE>, dim3(C.nchunk), dim3(256), 0, S.stream, vecs(S, C), slot);
}

static CgFinal finals(CgState& C) {
  CgFinal f;
  f.nseg      = C.nseg;
  f.nchunk    = C.nchunk;
  f.seg_chunk = C.seg_chunk.p;
  f.part      = C.part.p;
  f.sd        = C.sd.p;
  f.si        = C.si.p;
  f.status    = C.status.p;
  f.hstat     = C.d_hstat;
  f.params    = C.params.p;
  f.hist      = C.hist.p;
  f.hist_cap  = C.hist_cap;
  f.publish   = 1;
  return f;
}

template <int MODE>
static void launch_final(System& S, CgState& C, bool mixed, bool publish = true) {
  CgFinal f  = finals(C);
  f.publish  = publish ? 1 : 0;
  if (mixed)
    hipLaunchKernelGGL((k_sg_final<MODE, true>), dim3(1), dim3(256), 0, S.stream, f);
  else
    hipLaunchKernelGGL((k_sg_final<MODE, false>), dim3(1), dim3(256), 0, S.stream, f);
}

static void apply(System& S, CgState& C, const double* x, double* out, bool fp32) {
  if (C.nrhs == 1)
    apply_b(S, x, out, fp32);
  else
    apply_b_multi(S, x, out, C.nrhs, fp32);
}

static void seq_reliable(System& S, CgState& C) {
  apply(S, C, C.x, C.t, false);
  hipLaunchKernelGGL(k_sg_resid, dim3(C.nchunk), dim3(256), 0, S.stream, vecs(S, C));
  launch_final<F_REL>(S, C, true);
}

// FP32-corrected confirmation of the pending segments (cg_policy.hh, seg_confirm32)
static void seq_confirm(System& S, CgState& C) {
  const dim3 g(C.nchunk), b(256);
  hipLaunchKernelGGL(k_sg_defect, g, b, 0, S.stream, vecs(S, C));
  apply(S, C, C.dv, C.t, true);
  hipLaunchKernelGGL(k_sg_confirm, g, b, 0, S.stream, vecs(S, C));
  launch_final<F_CONF>(S, C, true);
}

static void seq_direction(System& S, CgState& C, bool mixed) {
  const dim3 g(C.nchunk), b(256);
  hipLaunchKernelGGL(k_sg_z, g, b, 0, S.stream, vecs(S, C));
  launch_prec<2>(S, C, 0);
  launch_final<F_BETA>(S, C, mixed);
  hipLaunchKernelGGL(k_sg_p, g, b, 0, S.stream, vecs(S, C));
}

// One iteration: the step (apply, p.Ap, alpha, x / r / z update, residual norm) and -- unless a
// segment asked for a reliable update, which the final kernel and k_sg_p test on the device --
// beta and the new direction, so that one host round trip per iteration suffices.
static void seq_step_direction(System& S, CgState& C, bool mixed, bool publish) {
  const dim3 g(C.nchunk), b(256);
  if (C.rec32 || C.rec64) S.prepacked = C.p;  // records written by k_sg_init / k_sg_p
  // one segment of one right-hand side: its chunks are the 256-site blocks of the L2P pass, which
  // then writes the p.Ap and Ap.Ap partials itself (when the apply ends in one)
  if (C.nseg == 1 && C.nrhs == 1 && C.na == S.n) S.dots = System::StepDots{C.p, C.Ap, C.part.p, C.nchunk, false};
  apply(S, C, C.p, C.Ap, mixed);
  S.prepacked    = nullptr;
  const bool dots = S.dots.done;
  S.dots          = System::StepDots{};
  if (!dots) hipLaunchKernelGGL(k_sg_pap, g, b, 0, S.stream, vecs(S, C));
  hipLaunchKernelGGL(k_sg_xrz, g, b, 0, S.stream, vecs(S, C));
  launch_prec<1>(S, C, 1);
  launch_final<F_RNB>(S, C, mixed, publish);
  hipLaunchKernelGGL(k_sg_p, g, b, 0, S.stream, vecs(S, C));
}

void cg_graphs_release(System& S) {
  S.cg.release_graphs();
  S.cgm.release_graphs();
}

static void capture(System& S, hipGraphExec_t& exec, const std::function<void()>& body) {
  hipGraph_t g       = nullptr;
  hipEvent_t* prof   = S.prof;
  S.prof             = nullptr;  // no event records inside a graph
  CPPE_HIP_CHECK(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(S.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    S.prof = prof;
    throw;
  }
  CPPE_HIP_CHECK(hipStreamEndCapture(S.stream, &g));
  S.prof = prof;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CPPE_HIP_CHECK(e);
}

static void run_graph(System& S, hipGraphExec_t& exec, const std::function<void()>& body) {
  if (!S.use_graphs) {
    body();
    return;
  }
  if (!exec) capture(S, exec, body);
  CPPE_HIP_CHECK(hipGraphLaunch(exec, S.stream));
}

// Waits for the status block of the last final kernel launched (C.expect of them so far). The
// final kernels publish it to the host-mapped mirror, so the host can decide and enqueue the next
// launch sequence while the rest of the current one (the new direction) still runs. A mirror that
// has not caught up within 20 ms per step launched (a long apply, or mapped writes that are late)
// is replaced by a synchronous copy of the device block, which is authoritative.
static void read_status(System& S, CgState& C, int steps = 1) {
  volatile int* hs = C.h_status;
  const auto t0    = std::chrono::steady_clock::now();
  const auto limit = std::chrono::milliseconds(20 * steps);
  int spins        = 0;
  while (hs[ST_SEQ] != C.expect) {
    if (++spins % 1024 == 0 && std::chrono::steady_clock::now() - t0 > limit) {
      CPPE_HIP_CHECK(hipStreamSynchronize(S.stream));
      CPPE_HIP_CHECK(hipMemcpy(C.h_status, C.status.p, sizeof(int) * ST_COUNT, hipMemcpyDeviceToHost));
      C.expect = C.h_status[ST_SEQ];  // every launched final has run: the device count is exact
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

// The solve parameters into the device block, unless it already holds exactly these (repeated
// solves of a system reuse them).
static void upload_params(System& S, CgState& C, const double* params) {
  static_assert(P_COUNT <= 8, "host copy of the parameters");
  if (C.params_ok && std::equal(params, params + P_COUNT, C.h_params)) return;
  std::copy(params, params + P_COUNT, C.h_params);
  C.params_ok = true;
  CPPE_HIP_CHECK(hipMemcpyAsync(C.params.p, C.h_params, sizeof(double) * P_COUNT, hipMemcpyHostToDevice, S.stream));
}

// A solve's results for the host in one pass into host-mapped memory: the segment tables sd and
// si and the first nh rows of the residual history (doubles first, then the ints).
__global__ void k_sg_export(const double* __restrict__ sd, int nsd, const double* __restrict__ hist, long long nhist,
                            const int* __restrict__ si, int nsi, double* __restrict__ od, int* __restrict__ oi) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nsd) od[t] = sd[t];
  if (t < nhist) od[nsd + t] = hist[t];
  if (t < nsi) oi[t] = si[t];
  __threadfence_system();  // the host reads them once the stream has drained
}

// Solves B x = rhs for C.rhs into C.x (which holds the initial guess). Results per segment are
// left in the device state; res.norms holds every segment's residual norms, norms0 segment 0's.
int device_cg(System& S, CgState& C, double thresh, int maxiter, bool mixed, CgResult& res) {
  hipStream_t st = S.stream;
  if (&C == &S.cg) bind_single(S);
  const int ns = C.nseg;
  if (C.hist_cap < maxiter || !C.hist.p) {
    C.hist_cap = std::max(maxiter, 64);
    C.hist.alloc((size_t)C.hist_cap * ns);
    C.release_graphs();  // the graphs hold the history buffer
  }
  fmm_reserve(S);  // no allocation inside a graph
  if (C.nrhs > 1) fmm_reserve_multi(S);
  if (S.precond == 1) {
    precond_build(S);
    C.wbuf.alloc(4 * (size_t)C.na * C.nrhs);  // one double4 record per site and right-hand side
  }
  // single right-hand side: the direction kernels also write p's near-field dipole records in
  // the precision of the step's apply, which then skips its own packing
  C.rec32 = nullptr;
  C.rec64 = nullptr;
  if (C.nrhs == 1) {
    float4* r32;
    double4* r64;
    fmm_dipole_records(S, &r32, &r64);
    if (mixed)
      C.rec32 = r32;
    else
      C.rec64 = r64;
  }
  // mixed form: moments and residual of the last FP64 true residual, for FP32-corrected confirmations
  const bool conf = mixed && S.confirm_eta > 0.0 && (S.confirm_forms & 2);
  if (conf) {
    const size_t nv = 3 * (size_t)C.na * C.nrhs;
    if (C.rstore.n < 3 * nv) {
      C.rstore.alloc(3 * nv);
      C.release_graphs();  // the graphs hold the vectors
    }
    C.xrel = C.rstore.p;
    C.rrel = C.rstore.p + nv;
    C.dv   = C.rstore.p + 2 * nv;
  } else if (C.xrel) {
    C.xrel = C.rrel = C.dv = nullptr;
    C.release_graphs();
  }
  double params[P_COUNT];
  params[P_THRESH]   = thresh;
  params[P_DROP]     = S.reliable_drop;
  params[P_PROGRESS] = S.refine_min_progress;
  params[P_STALL]    = S.refine_stall;
  params[P_RESTART]  = 0.0;
  params[P_ETA]      = conf ? S.confirm_eta : 0.0;
  params[P_GAP]      = (S.confirm_forms & 1) ? S.confirm_eta : 0.0;
  upload_params(S, C, params);
  // r0 = rhs - B x0, z0 = alpha r0, p = z0. The mixed form takes r0 from an FP32 apply: it is
  // replaced by an FP64 true residual at the first reliable update, before any convergence test.
  // The first step goes in the same launch (its horizon is one step: no reduction rate is known
  // yet); it changes nothing when the initial decision stopped a segment or asked for a true
  // residual, as any step launched past an event.
  int m = mixed ? 1 : 0;
  const bool first = maxiter > 0;
  auto initial = [&] {
    apply(S, C, C.x, C.t, mixed);
    hipLaunchKernelGGL(k_sg_init, dim3(C.nchunk), dim3(256), 0, st, vecs(S, C));
    launch_prec<0>(S, C, 0);
    launch_final<F_INIT>(S, C, mixed, !first);
    if (first) seq_step_direction(S, C, mixed, true);
  };
  if (first)
    run_graph(S, C.g_init[m], initial);
  else
    initial();
  C.expect += first ? 2 : 1;  // every launch sequence below runs one final kernel per step
  read_status(S, C, first ? 2 : 1);
  int k = first ? C.h_status[ST_IT] + C.h_status[ST_STEPPED] : 0;
  int nrel = 0, nlaunch = first ? 1 : 0, nsteps = first ? 1 : 0, nread = 1, nconf = 0;
  res.fp64_fallback = -1;
  // Loop invariant at the top: every active segment has completed step, replacement and
  // direction of its last iteration. Pending segments (mixed form) are confirmed together: inside a
  // replacement that runs anyway, or once no segment is active.
  auto reliable = [&] {
    run_graph(S, C.g_rel, [&] { seq_reliable(S, C); });
    ++C.expect;
    read_status(S, C);
    ++nrel;
    ++nread;
  };
  // Pending segments with nothing else to do: first the FP32-corrected confirmation, then an FP64
  // true residual for those it did not confirm.
  auto confirm = [&] {
    if (mixed && conf && C.h_status[ST_REL] == 0) {
      run_graph(S, C.g_conf, [&] { seq_confirm(S, C); });
      ++C.expect;
      read_status(S, C);
      ++nconf;
      ++nread;
      if (C.h_status[ST_PEND] == 0) return;
    }
    reliable();
  };
  // as many steps as one launch as the device's horizon, the iterations left and the system's
  // limit allow (a graph per count); k = steps taken (steps that start after an event take none)
  auto steps = [&] {
    const int cnt = std::max(1, std::min({C.h_status[ST_SAFE], maxiter - k, S.graph_steps}));
    run_graph(S, C.g_step[m][cnt], [&] {
      // only the last step's final kernel publishes: the host waits for the launch as a whole
      for (int q = 0; q < cnt; ++q) seq_step_direction(S, C, mixed, q == cnt - 1);
    });
    C.expect += cnt;
    read_status(S, C, cnt);
    k = C.h_status[ST_IT] + C.h_status[ST_STEPPED];
    ++nlaunch;
    nsteps += cnt;
    ++nread;
  };
  // A stagnating refinement (a true residual that did not fall below min_progress times the
  // previous one) continues in the plain form: every unfinished segment restarts from its FP64
  // true residual, which the replacement apply has just left in t (beta = 0), and every later
  // apply runs in FP64. Returns whether it happened.
  auto fallback = [&]() -> bool {
    if (!mixed || C.h_status[ST_STAG] == 0) return false;
    mixed = false;
    m     = 0;
    if (C.nrhs == 1) {
      float4* r32;
      double4* r64;
      fmm_dipole_records(S, &r32, &r64);
      C.rec32 = nullptr;
      C.rec64 = r64;
    }
    params[P_RESTART] = 1.0;
    upload_params(S, C, params);
    hipLaunchKernelGGL(k_sg_init, dim3(C.nchunk), dim3(256), 0, st, vecs(S, C));
    launch_prec<0>(S, C, 0);
    launch_final<F_INIT>(S, C, false);
    params[P_RESTART] = 0.0;
    upload_params(S, C, params);
    ++C.expect;
    read_status(S, C);
    ++nread;
    res.fp64_fallback = k;
    return true;
  };
  if (C.h_status[ST_REL] > 0) {
    reliable();
    fallback();
  }
  while (k < maxiter) {
    if (C.h_status[ST_ACTIVE] == 0) {
      if (C.h_status[ST_PEND] == 0) break;
      confirm();
      if (fallback()) continue;
      if (C.h_status[ST_ACTIVE] == 0) break;
      run_graph(S, C.g_dir[m], [&] { seq_direction(S, C, mixed); });
      ++C.expect;
      continue;
    }
    steps();
    // The direction already ran unless a reliable update was asked for (it is then taken after
    // the update). Confirming pending segments when none is active leaves the direction pass a
    // no-op on the vectors; segments that resume then take theirs after the confirmation.
    if (C.h_status[ST_REL] > 0 || (C.h_status[ST_ACTIVE] == 0 && C.h_status[ST_PEND] > 0)) {
      if (C.h_status[ST_REL] > 0)
        reliable();
      else
        confirm();
      if (fallback()) continue;
      if (C.h_status[ST_ACTIVE] > 0) {
        run_graph(S, C.g_dir[m], [&] { seq_direction(S, C, mixed); });
        ++C.expect;
      }
    }
  }
  if (mixed && C.h_status[ST_PEND] > 0) reliable();  // maxiter reached with unconfirmed segments
  // per-segment results
  const int nh = std::min(k, C.hist_cap);
  const int nsd = D_COUNT * ns, nsi = I_COUNT * ns;
  const long long nhist = (long long)nh * ns;
  const size_t bytes = sizeof(double) * (nsd + (size_t)nhist) + sizeof(int) * (size_t)nsi;
  if (C.out_bytes < bytes) {
    if (C.h_out) CPPE_HIP_CHECK(hipHostFree(C.h_out));
    C.h_out     = nullptr;
    C.out_bytes = 0;
    const size_t cap = std::max<size_t>(bytes, 4096);
    CPPE_HIP_CHECK(hipHostMalloc((void**)&C.h_out, cap, hipHostMallocMapped | hipHostMallocCoherent));
    CPPE_HIP_CHECK(hipHostGetDevicePointer((void**)&C.d_out, C.h_out, 0));
    C.out_bytes = cap;
  }
  double* od = reinterpret_cast<double*>(C.d_out);
  int* oi    = reinterpret_cast<int*>(C.d_out + sizeof(double) * (nsd + (size_t)nhist));
  const long long nt = std::max<long long>({(long long)nsd, nhist, (long long)nsi});
  hipLaunchKernelGGL(k_sg_export, dim3(ceil_div(nt, 256)), dim3(256), 0, st, C.sd.p, nsd, C.hist.p, nhist, C.si.p, nsi,
                     od, oi);
  CPPE_HIP_CHECK(hipStreamSynchronize(st));
  const double* hd = reinterpret_cast<const double*>(C.h_out);
  const int* hi    = reinterpret_cast<const int*>(C.h_out + sizeof(double) * (nsd + (size_t)nhist));
  const std::vector<double> sd(hd, hd + nsd), hist(hd + nsd, hd + nsd + nhist);
  const std::vector<int> si(hi, hi + nsi);
  res.iters.assign(ns, 0);
  res.rnorm.assign(ns, 0.0);
  res.converged.assign(ns, 0);
  res.norms0.clear();
  res.norms.assign(ns, {});
  res.all_converged = true;
  for (int g = 0; g < ns; ++g) {
    const bool active = si[I_ACTIVE * ns + g] != 0 || si[I_PENDING * ns + g] != 0;
    res.iters[g]      = active ? k : si[I_ITERS * ns + g];
    res.rnorm[g]      = sd[D_RNORM * ns + g];
    res.converged[g]  = !active && !si[I_BAD * ns + g] && res.rnorm[g] < thresh;
    res.all_converged = res.all_converged && res.converged[g];
    for (int it = 0; it < std::min(res.iters[g], nh); ++it) res.norms[g].push_back(hist[(size_t)it * ns + g]);
  }
  res.norms0           = res.norms[0];
  res.loop_iters       = k;
  res.reliable_updates = nrel;
  res.confirmations    = nconf;
  res.step_launches    = nlaunch;
  res.steps_launched   = nsteps;
  res.host_reads       = nread;
  return k;
}

}  // namespace cppe_dev
