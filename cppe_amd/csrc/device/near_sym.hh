// Symmetric near field of the B apply: every unordered pair of near-field sites evaluated once.
//
// The dipole tensor is even in the displacement, T(d) = T(-d), and the Thole argument
// v = damping s_i s_j R is symmetric, so one evaluation of a pair's geometry (1/R, the Thole
// factors: about two thirds of the arithmetic of a pair) serves the field at both sites:
//   F_i += T(d) mu_j,   F_j += T(d) mu_i.
// The reference evaluates ordered pairs (fmm/calculate.cc:205-225, P2P over each target leaf's
// near list); the sums are the same up to rounding.
//
// Work split. The near lists are symmetric at the leaf level (B near A <=> A near B). A work item
// is a chunk of at most TS targets of leaf A; it walks A's "upper" list: the sites of A itself
// (ordered pairs, field at A's sites only), then every site of each near leaf B > A (both
// directions). One wavefront per item; lanes are sources (64 per stage), the item's targets are
// uniform across the wavefront (wavefront-private LDS, broadcast reads) and looped over, so
//   * the field at a target accumulates in per-lane partials f[t] (summed over lanes once per item),
//   * the field at the lane's source accumulates in three registers over the item's targets and is
//     stored once per stage: reverse slot (item, source site).
// Reverse slots are laid out per receiving leaf B: one block of 3 |B| values (x, y, z rows) per
// incoming chunk (leaf A < B with B near A, chunk q of A), blocks in increasing (A, q). The
// gather kernel adds, per site of B, its blocks in that order, then the L2P term: every sum has a
// fixed order, and no atomics are used, so applies are bit-reproducible.
//
// Exclusions: excluded pairs are evaluated like every other pair and taken out afterwards per
// site by k_fmm_excl_sub (the field at the lower leaf's site with the same arithmetic; at the other
// site to the rounding of the precision).
//
// Included by fmm_impl.hh inside namespace cppe_dev (uses NearArgs, the staged pair arithmetic
// and the FMM operators).
#pragma once

struct SymArgs {
  const int4* items;          // (leaf A, first target, targets, chunk index of A)
  int nitems;
  const long long* uoff;      // per cell: upper list range; the first count[A] entries are A's own sites
  const long long* usplit;    // per cell (FP32): first entry whose stages may use the undamped pair
  const int* usrc;            // upper list: source site
  const int2* udst;           // upper list: (reverse slot of chunk 0 / component x, |B|)
};

template <typename R>
struct alignas(4 * sizeof(R)) SymVec4 {
  R x, y, z, w;
};
// a target of a work item: position (FP32: offset from the leaf centre) and damping rate; dipole
template <typename R>
struct SymTarget {
  SymVec4<R> p, m;
};

// Geometry once, field at both ends: f += T(d) mu_s, r += T(d) mu_t.
template <typename R, bool DAMP, bool B2, bool FAR>
__device__ __forceinline__ void sym_pair(const SymTarget<R>& t, R xs, R ys, R zs, R ss, R mxs, R mys, R mzs, R& fx,
                                         R& fy, R& fz, R& rx, R& ry, R& rz) {
  const SymVec4<R> tp = t.p, tm = t.m;
  const R dx = tp.x - xs, dy = tp.y - ys, dz = tp.z - zs;
  if constexpr (FAR && sizeof(R) == 4) {
    float ri3, m3;
    geo_far_f32(dx, dy, dz, ri3, m3);
    acc_far_f32(ri3, m3, dx, dy, dz, mxs, mys, mzs, fx, fy, fz);
    acc_far_f32(ri3, m3, dx, dy, dz, tm.x, tm.y, tm.z, rx, ry, rz);
  } else {
    R c3, cf;
    staged_geo<R, DAMP, B2, FAR>(dx, dy, dz, tp.w * ss, c3, cf);
    acc_m1(c3, cf, dx, dy, dz, mxs, mys, mzs, fx, fy, fz);
    acc_m1(c3, cf, dx, dy, dz, tm.x, tm.y, tm.z, rx, ry, rz);
  }
}

// FP32 DPP sum over a wavefront; the total is uniform (lane 63's value, read into a scalar).
__device__ __forceinline__ float wave_sum_f32(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xa, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xc, 0xf, false));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <typename R>
__device__ __forceinline__ R sym_wave_sum(R v) {
  if constexpr (sizeof(R) == 4)
    return wave_sum_f32(v);
  else
    return wave_sum_f64(v);
}

// Symmetric near field (B apply, one right-hand side). Target records trec: FP32 offsets from
// the leaf centre (formed in FP64) with the damping rate d_t (B2: times log2 e), FP64 absolute
// positions with d_t. Source positions: FP32 offsets from the target leaf centre formed in FP64,
// as in the staged kernel. out[i] = (field from the item's pairs) + alpha^-1 mu for the item's
// targets; the reverse slots (rev) receive the field at the sources of upper leaves.
template <typename R, bool DAMP, bool B2, int TS, typename MU4, typename TREC>
__global__ __launch_bounds__(256) void k_fmm_near_sym(NearArgs a, SymArgs b, const TREC* __restrict__ trec,
                                                      const MU4* __restrict__ mu4, R* __restrict__ rev) {
  const int lane = threadIdx.x & 63;
  const int w    = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (w >= b.nitems) return;
  const int4 it  = b.items[w];
  const int A    = __builtin_amdgcn_readfirstlane(it.x);
  const int i0   = __builtin_amdgcn_readfirstlane(a.start[A] + it.y);
  const int T    = __builtin_amdgcn_readfirstlane(it.z);
  const int q    = __builtin_amdgcn_readfirstlane(it.w);
  const double cax = a.cx[A], cay = a.cy[A], caz = a.cz[A];
  constexpr R kFar = sizeof(R) == 4 ? R(1.0e4) : R(1.0e8);  // idle lanes: a source far away, zero dipole
  R f[TS][3];
#pragma unroll
  for (int t = 0; t < TS; ++t) f[t][0] = f[t][1] = f[t][2] = R(0);
  const long long e0 = b.uoff[A], e1 = b.uoff[A + 1];
  const long long es = e0 + a.count[A];
  const long long ef = (sizeof(R) == 4 && DAMP && b.usplit) ? b.usplit[A] : e1;
  auto convert = [&](const double4& pj, const MU4& mj, bool act, R& xs, R& ys, R& zs, R& ss, R& mx, R& my,
                     R& mz) {
    if (act) {
      xs = sizeof(R) == 8 ? (R)pj.x : (R)(pj.x - cax);
      ys = sizeof(R) == 8 ? (R)pj.y : (R)(pj.y - cay);
      zs = sizeof(R) == 8 ? (R)pj.z : (R)(pj.z - caz);
      ss = (R)pj.w;
      mx = (R)mj.x;
      my = (R)mj.y;
      mz = (R)mj.z;
    } else {
      xs = ys = zs = kFar;
      ss = R(0);
      mx = my = mz = R(0);
    }
  };
  // the item's targets in wavefront-private LDS (record, dipole), read with uniform addresses
  // (broadcast) inside the pair loops
  __shared__ SymTarget<R> tl[4][TS];
  SymTarget<R>* tg = tl[threadIdx.x >> 6];
  if (lane < T) {
    const TREC tr = trec[i0 + lane];
    const MU4 mt  = mu4[i0 + lane];
    tg[lane].p    = SymVec4<R>{(R)tr.x, (R)tr.y, (R)tr.z, (R)tr.w};
    tg[lane].m    = SymVec4<R>{(R)mt.x, (R)mt.y, (R)mt.z, R(0)};
  }
  wave_lds_barrier();
  // A's own sites: ordered pairs, field at the item's targets only (the self pair gives zero)
  for (long long q0 = e0; q0 < es; q0 += 64) {
    const bool act = q0 + lane < es;
    R xs, ys, zs, ss, mx, my, mz;
    const int j = a.start[A] + (int)(q0 - e0) + lane;
    convert(act ? a.pos4[j] : make_double4(0.0, 0.0, 0.0, 0.0), act ? mu4[j] : MU4{}, act, xs, ys, zs, ss, mx, my,
            mz);
#pragma unroll
    for (int t = 0; t < TS; ++t) {
      if (t < T) {
        const SymVec4<R> tp = tg[t].p;
        const R dx = tp.x - xs, dy = tp.y - ys, dz = tp.z - zs;
        R c3, cf;
        staged_geo<R, DAMP, B2, false>(dx, dy, dz, tp.w * ss, c3, cf);
        acc_m1(c3, cf, dx, dy, dz, mx, my, mz, f[t][0], f[t][1], f[t][2]);
      }
    }
  }
  // upper leaves: both directions. Stage pipeline: the source records of stage s + 1 and the
  // source index of stage s + 2 are loaded while stage s is evaluated.
  double4 pn = make_double4(0.0, 0.0, 0.0, 0.0);
  MU4 mn{};
  int2 dn = make_int2(0, 0), dnn = make_int2(0, 0);
  int jnn = 0;
  if (es + lane < e1) {
    const int j0 = b.usrc[es + lane];
    dn = b.udst[es + lane];
    pn = a.pos4[j0];
    mn = mu4[j0];
  }
  if (es + 64 + lane < e1) {
    jnn = b.usrc[es + 64 + lane];
    dnn = b.udst[es + 64 + lane];
  }
  for (long long q0 = es; q0 < e1; q0 += 64) {
    const bool act = q0 + lane < e1;
    R xs, ys, zs, ss, mx, my, mz;
    convert(pn, mn, act, xs, ys, zs, ss, mx, my, mz);
    const int2 dd = dn;
    if (q0 + 64 + lane < e1) {  // stage s + 1
      pn = a.pos4[jnn];
      mn = mu4[jnn];
      dn = dnn;
    }
    if (q0 + 128 + lane < e1) {  // stage s + 2: its index and reverse slot
      jnn = b.usrc[q0 + 128 + lane];
      dnn = b.udst[q0 + 128 + lane];
    }
    R rx = R(0), ry = R(0), rz = R(0);
    if (q0 >= ef) {  // every pair of the stage beyond the precision's damping cutoff (uniform)
#pragma unroll
      for (int t = 0; t < TS; ++t)
        if (t < T) sym_pair<R, DAMP, B2, true>(tg[t], xs, ys, zs, ss, mx, my, mz, f[t][0], f[t][1], f[t][2], rx, ry, rz);
    } else {
#pragma unroll
      for (int t = 0; t < TS; ++t)
        if (t < T) sym_pair<R, DAMP, B2, false>(tg[t], xs, ys, zs, ss, mx, my, mz, f[t][0], f[t][1], f[t][2], rx, ry, rz);
    }
    if (act) {
      const int o = dd.x + q * 3 * dd.y;
      rev[o]             = rx;
      rev[o + dd.y]      = ry;
      rev[o + 2 * dd.y]  = rz;
    }
  }
  // targets: lane t collects the wavefront sums of f[t]
  double F[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < TS; ++t) {
    if (t < T) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double v = (double)sym_wave_sum(f[t][c]);
        if (lane == t) F[c] = v;
      }
    }
  }
  if (lane >= T) return;
  const int i = i0 + lane, n = a.n;
  const double* ai = a.ainv;
  double a6[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) a6[c] = ai[(size_t)c * n + i];
  const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
  a.out[i]                 = F[0] + (a6[0] * m0 + a6[1] * m1 + a6[2] * m2);
  a.out[n + i]             = F[1] + (a6[1] * m0 + a6[3] * m1 + a6[4] * m2);
  a.out[2 * (size_t)n + i] = F[2] + (a6[2] * m0 + a6[4] * m1 + a6[5] * m2);
}

// out += (reverse blocks of the site's leaf, in block order) + L2P of the leaf expansion (when L).
template <int P, typename R>
__global__ __launch_bounds__(256) void k_fmm_sym_gather(int n, const int* __restrict__ leaf_of,
                                                        const int* __restrict__ start, const int* __restrict__ count,
                                                        const long long* __restrict__ rev_off,
                                                        const int* __restrict__ nin, const R* __restrict__ rev,
                                                        const double4* __restrict__ pos4, const double* __restrict__ cx,
                                                        const double* __restrict__ cy, const double* __restrict__ cz,
                                                        const R* __restrict__ L, double* __restrict__ out) {
  using O     = Ops<KIND_M1, P, R>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int B  = leaf_of[i];
  const int nb = count[B], u = i - start[B], nk = nin[B];
  const R* blk = rev + rev_off[B] + u;
  double g[3] = {0.0, 0.0, 0.0};
#pragma unroll 8
  for (int k = 0; k < nk; ++k, blk += 3 * nb) {  // in block order (the loads of 8 blocks issued together)
    g[0] += (double)blk[0];
    g[1] += (double)blk[nb];
    g[2] += (double)blk[2 * nb];
  }
  if (L) {
    const double4 pi = pos4[i];
    R lx = R(0), ly = R(0), lz = R(0);
    O::l2p((R)(pi.x - cx[B]), (R)(pi.y - cy[B]), (R)(pi.z - cz[B]), &L[(size_t)B * O::LS], lx, ly, lz);
    g[0] += (double)lx;
    g[1] += (double)ly;
    g[2] += (double)lz;
  }
  out[i] += g[0];
  out[n + i] += g[1];
  out[2 * (size_t)n + i] += g[2];
}


// ---- direct summation (one leaf holding every site) ---------------------------------------------
// All pairs are near, so the symmetric form tiles them: W wavefronts of a block hold TS targets
// each (a tile of TB = W TS consecutive sites); an item is a tile I against a segment of the
// sources above it (both directions), or against its own sites (ordered pairs, field at the tile
// only). Lanes are sources, 64 per stage, shared by the block's wavefronts; after each stage the
// wavefronts' fields at the stage's sources are summed through LDS in wavefront order and stored
// once per (tile, source): reverse block (J, I) of the receiving tile J, 3 TB values, at
// 3 TB (J (J - 1) / 2 + I). The forward sums of an item go to its own block of the tile
// (fwd_off[I] + k 3 TB, k = the item's index among the tile's items, its own sites first).
// k_direct_sym_gather adds, per site, its tile's forward blocks and then its reverse blocks in
// block order, with alpha^-1 mu: every sum in a fixed order, no atomics.
struct DSymArgs {
  const int4* items;  // (tile, first source, end source, index of the item among the tile's items)
  int nitems, n;
  const long long* fwd_off;  // per tile: first forward slot
  double cx, cy, cz;         // the leaf's centre (FP32 offsets)
};

__device__ __forceinline__ long long dsym_rev_block(int J, int I, int TB) {
  return 3ll * TB * ((long long)J * (J - 1) / 2 + I);
}

template <typename R, bool DAMP, bool B2, int TS, int W, typename MU4, typename TREC>
__global__ __launch_bounds__(64 * W) void k_direct_sym(NearArgs a, DSymArgs b, const TREC* __restrict__ trec,
                                                       const MU4* __restrict__ mu4, R* __restrict__ fwd,
                                                       R* __restrict__ rev) {
  constexpr int TB = W * TS;
  __shared__ SymTarget<R> tl[W][TS];
  __shared__ R red[2][W][3][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int4 it  = b.items[blockIdx.x];
  const int I = it.x, kb = it.w;
  const long long s0 = it.y, s1 = it.z;
  const bool diag = s0 == (long long)I * TB;  // the tile's own sites (uniform over the block)
  const int t0    = I * TB + wv * TS;
  const int T     = max(0, min(TS, b.n - t0));
  constexpr R kFar = sizeof(R) == 4 ? R(1.0e4) : R(1.0e8);
  SymTarget<R>* tg = tl[wv];
  if (lane < T) {
    const TREC tr = trec[t0 + lane];
    const MU4 mt  = mu4[t0 + lane];
    tg[lane].p    = SymVec4<R>{(R)tr.x, (R)tr.y, (R)tr.z, (R)tr.w};
    tg[lane].m    = SymVec4<R>{(R)mt.x, (R)mt.y, (R)mt.z, R(0)};
  }
  __syncthreads();
  R f[TS][3];
#pragma unroll
  for (int t = 0; t < TS; ++t) f[t][0] = f[t][1] = f[t][2] = R(0);
  int buf = 0;
  for (long long q0 = s0; q0 < s1; q0 += 64) {
    const long long j = q0 + lane;
    const bool act    = j < s1;
    R xs = kFar, ys = kFar, zs = kFar, ss = R(0), mx = R(0), my = R(0), mz = R(0);
    if (act) {
      const double4 pj = a.pos4[j];
      const MU4 mj     = mu4[j];
      xs = sizeof(R) == 8 ? (R)pj.x : (R)(pj.x - b.cx);
      ys = sizeof(R) == 8 ? (R)pj.y : (R)(pj.y - b.cy);
      zs = sizeof(R) == 8 ? (R)pj.z : (R)(pj.z - b.cz);
      ss = (R)pj.w;
      mx = (R)mj.x;
      my = (R)mj.y;
      mz = (R)mj.z;
    }
    if (diag) {  // ordered pairs inside the tile (the self pair gives zero)
#pragma unroll
      for (int t = 0; t < TS; ++t) {
        if (t < T) {
          const SymVec4<R> tp = tg[t].p;
          const R dx = tp.x - xs, dy = tp.y - ys, dz = tp.z - zs;
          R c3, cf;
          staged_geo<R, DAMP, B2, false>(dx, dy, dz, tp.w * ss, c3, cf);
          acc_m1(c3, cf, dx, dy, dz, mx, my, mz, f[t][0], f[t][1], f[t][2]);
        }
      }
      continue;
    }
    R rx = R(0), ry = R(0), rz = R(0);
#pragma unroll
    for (int t = 0; t < TS; ++t)
      if (t < T) sym_pair<R, DAMP, B2, false>(tg[t], xs, ys, zs, ss, mx, my, mz, f[t][0], f[t][1], f[t][2], rx, ry, rz);
    red[buf][wv][0][lane] = rx;
    red[buf][wv][1][lane] = ry;
    red[buf][wv][2][lane] = rz;
    __syncthreads();  // (also orders this buffer's reads two stages back before it is rewritten)
    if (threadIdx.x < 3 * 64) {
      const int c = threadIdx.x >> 6, l = threadIdx.x & 63;
      const long long jj = q0 + l;
      if (jj < s1) {
        R s = red[buf][0][c][l];
#pragma unroll
        for (int w = 1; w < W; ++w) s += red[buf][w][c][l];
        const int J = (int)(jj / TB), u = (int)(jj - (long long)J * TB);
        rev[dsym_rev_block(J, I, TB) + c * TB + u] = s;
      }
    }
    buf ^= 1;
  }
  // the item's forward sums of each target: wavefront sums, lane t stores target t
  R* blk = fwd + b.fwd_off[I] + 3ll * TB * kb + wv * TS;
#pragma unroll
  for (int t = 0; t < TS; ++t) {
    if (t < T) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R v = sym_wave_sum(f[t][c]);
        if (lane == t) blk[c * TB + t] = v;
      }
    }
  }
}

// out = (the tile's forward blocks) + (reverse blocks from the tiles below) + alpha^-1 mu
template <typename R>
__global__ __launch_bounds__(256) void k_direct_sym_gather(NearArgs a, int TB, const long long* __restrict__ fwd_off,
                                                           const R* __restrict__ fwd, const R* __restrict__ rev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.n;
  if (i >= n) return;
  const int J = i / TB, u = i - J * TB;
  const int nk = (int)((fwd_off[J + 1] - fwd_off[J]) / (3ll * TB));
  double g[3] = {0.0, 0.0, 0.0};
  const R* p = fwd + fwd_off[J] + u;
  for (int k = 0; k < nk; ++k, p += 3 * TB) {
    g[0] += (double)p[0];
    g[1] += (double)p[TB];
    g[2] += (double)p[2 * TB];
  }
  const R* q = rev + dsym_rev_block(J, 0, TB) + u;
#pragma unroll 8
  for (int I = 0; I < J; ++I, q += 3 * TB) {
    g[0] += (double)q[0];
    g[1] += (double)q[TB];
    g[2] += (double)q[2 * TB];
  }
  const double* ai = a.ainv;
  double a6[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) a6[c] = ai[(size_t)c * n + i];
  const double m0 = a.src[i], m1 = a.src[n + i], m2 = a.src[2 * (size_t)n + i];
  a.out[i]                 = g[0] + (a6[0] * m0 + a6[1] * m1 + a6[2] * m2);
  a.out[n + i]             = g[1] + (a6[1] * m0 + a6[3] * m1 + a6[4] * m2);
  a.out[2 * (size_t)n + i] = g[2] + (a6[2] * m0 + a6[4] * m1 + a6[5] * m2);
}
