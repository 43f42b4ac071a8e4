// Per-site MLP 64 -> 128 -> 128 -> 3 (SiLU) of the learned CG warm start, forward and backward
// on BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation. No reference counterpart.
//
// One block (256 threads = 4 wavefronts) owns a tile of 32 samples at a time:
//   layer 1  Z1 = X W1 + b1      MFMA: M = 32 samples, N = 128 (one 32-column slab per wave), K = 64
//   layer 2  Z2 = s(Z1) W2 + b2  MFMA: K = 128
//   layer 3  y  = s(Z2) W3 + b3  VALU (3 outputs)
// Training adds the loss of the equivariant guess mu_g = sum_a y_a v_a against the FP64 moment,
// L = sum_s w_s |mu_g - t|^2 / N (w_s = 1/|v1|^2), and its backward pass:
//   dW3, db3 (VALU), dZ2 = dy W3^T . s'(Z2), dW2 = s(Z1)^T dZ2 (MFMA, K = 32 samples),
//   dZ1 = dZ2 W2^T . s'(Z1) (MFMA, K = 128), dW1 = X^T dZ1 (MFMA).
// Weight gradients accumulate in registers across the tiles of a block and are written as
// per-block partials, summed by a second kernel in fixed order (deterministic).
// MFMA operand layout (32x32x16): lane l holds A[l%32][8(l/32)+e] and B[8(l/32)+e][l%32],
// e = 0..7; the accumulator element q of lane l is C[8(q/4) + 4(l/32) + q%4][l%32].
#include <cmath>

#include "ml.hh"

namespace cppe_dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TS  = 32;  // samples per tile
constexpr int LX  = kMlpIn + 8;    // padded row strides (bf16 elements), multiples of 8
constexpr int LH  = kMlpHid + 8;
constexpr int LT  = TS + 8;

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) {
  bf16x8 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

// acc += A[32 x K] * B[K x 32], A row-major (stride lda), B given as Bt (32 rows of K, stride ldb)
__device__ __forceinline__ void mfma_tile(const unsigned short* A, int lda, const unsigned short* Bt, int ldb,
                                          int K, f32x16& acc) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, kh = (l >> 5) * 8;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const bf16x8 a = ld8(A + r * lda + k0 + kh);
    const bf16x8 b = ld8(Bt + r * ldb + k0 + kh);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
}

__device__ __forceinline__ int acc_row(int q) {
  const int l = threadIdx.x & 63;
  return 8 * (q >> 2) + 4 * (l >> 5) + (q & 3);
}
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

struct MlpTileArgs {
  const float* w;                // FP32 master parameters
  const unsigned short *w1t, *w1, *w2t, *w2;  // bf16 operand copies
  const unsigned short* Xb;      // normalised bf16 features (ns x 64) or null (raw path)
  const float* Xraw;             // raw FP32 features (ns x 64), normalised here when Xb is null
  const float* mean;
  const float* inv_std;
  const float* V;                // 9 floats per sample
  const float* T;                // 3 floats per sample
  const float* Wt;               // per-sample loss weight
  long long ns;
  const int* tile_order;         // tile permutation (training), or null
  int ntiles;
  float inv_norm;                // 1 / (number of samples in the minibatch)
  float* y;                      // outputs (ns x 3), forward mode
  float* partials;               // per-block gradient partials (kMlpParams + 1 for the loss)
};

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_tile(MlpTileArgs a) {
  __shared__ __attribute__((aligned(16))) unsigned short Xs[TS * LX];
  __shared__ __attribute__((aligned(16))) unsigned short Xt[kMlpIn * LT];
  __shared__ __attribute__((aligned(16))) unsigned short H1s[TS * LH];
  __shared__ __attribute__((aligned(16))) unsigned short H1t[kMlpHid * LT];
  __shared__ __attribute__((aligned(16))) unsigned short D2s[TS * LH];
  __shared__ __attribute__((aligned(16))) unsigned short D2t[kMlpHid * LT];
  __shared__ __attribute__((aligned(16))) unsigned short D1t[kMlpHid * LT];
  __shared__ float Z1[TS * kMlpHid];
  __shared__ float Z2[TS * kMlpHid];
  __shared__ float ys[TS * 4], dys[TS * 4];
  __shared__ float lossw[4];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int col = acc_col();
  // persistent gradient accumulators (per lane): dW2 4 tiles, dW1 2 tiles
  f32x16 gW2[4], gW1[2];
  if (TRAIN) {
    for (int t = 0; t < 4; ++t) gW2[t] = f32x16{};
    for (int t = 0; t < 2; ++t) gW1[t] = f32x16{};
  }
  float gW3[2] = {0.f, 0.f};   // thread owns up to 2 of the 384 W3 entries
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
  for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
    const int tile = a.tile_order ? a.tile_order[it] : it;
    const long long s0 = (long long)tile * TS;
    // ---- load features (normalised bf16), both row-major and transposed
    for (int e = tid; e < TS * kMlpIn; e += 256) {
      const int s = e / kMlpIn, f = e % kMlpIn;
      unsigned short v = 0;
      if (s0 + s < a.ns) {
        if (a.Xb)
          v = a.Xb[(s0 + s) * kMlpIn + f];
        else
          v = f2bf((a.Xraw[(s0 + s) * kMlpIn + f] - a.mean[f]) * a.inv_std[f]);
      }
      Xs[s * LX + f] = v;
      Xt[f * LT + s] = v;
    }
    __syncthreads();
    // ---- layer 1: wave owns hidden columns [32 wave, 32 wave + 32)
    {
      f32x16 acc = f32x16{};
      mfma_tile(Xs, LX, a.w1t + (wave * 32) * kMlpIn, kMlpIn, kMlpIn, acc);
      const int n = wave * 32 + col;
      const float b = a.w[kOffB1 + n];
      for (int q = 0; q < 16; ++q) {
        const int s  = acc_row(q);
        const float z = acc[q] + b;
        Z1[s * kMlpHid + n] = z;
        const unsigned short h = f2bf(silu(z));
        H1s[s * LH + n] = h;
        H1t[n * LT + s] = h;
      }
    }
    __syncthreads();
    // ---- layer 2
    {
      f32x16 acc = f32x16{};
      mfma_tile(H1s, LH, a.w2t + (wave * 32) * kMlpHid, kMlpHid, kMlpHid, acc);
      const int n = wave * 32 + col;
      const float b = a.w[kOffB2 + n];
      for (int q = 0; q < 16; ++q) Z2[acc_row(q) * kMlpHid + n] = acc[q] + b;
    }
    __syncthreads();
    // ---- layer 3 (VALU): thread t < 96 computes y[s][o], s = t / 3, o = t % 3
    if (tid < TS * kMlpOut) {
      const int s = tid / 3, o = tid % 3;
      float acc = a.w[kOffB3 + o];
      for (int h = 0; h < kMlpHid; ++h) acc += silu(Z2[s * kMlpHid + h]) * a.w[kOffW3 + h * kMlpOut + o];
      ys[s * 4 + o] = acc;
      if (!TRAIN && s0 + s < a.ns) a.y[(s0 + s) * 3 + o] = acc;
    }
    if (!TRAIN) {
      __syncthreads();
      continue;
    }
    __syncthreads();
    // ---- loss and dy: thread s < TS
    if (tid < TS) {
      const int s = tid;
      float d0 = 0.f, d1 = 0.f, d2 = 0.f;
      if (s0 + s < a.ns) {
        const float* v = a.V + (s0 + s) * 9;
        const float* t = a.T + (s0 + s) * 3;
        const float c0 = ys[s * 4], c1 = ys[s * 4 + 1], c2 = ys[s * 4 + 2];
        const float ex = c0 * v[0] + c1 * v[3] + c2 * v[6] - t[0];
        const float ey = c0 * v[1] + c1 * v[4] + c2 * v[7] - t[1];
        const float ez = c0 * v[2] + c1 * v[5] + c2 * v[8] - t[2];
        const float w  = a.Wt[s0 + s] * a.inv_norm;
        lsum += w * (ex * ex + ey * ey + ez * ez);
        d0 = 2.f * w * (ex * v[0] + ey * v[1] + ez * v[2]);
        d1 = 2.f * w * (ex * v[3] + ey * v[4] + ez * v[5]);
        d2 = 2.f * w * (ex * v[6] + ey * v[7] + ez * v[8]);
      }
      dys[s * 4]     = d0;
      dys[s * 4 + 1] = d1;
      dys[s * 4 + 2] = d2;
    }
    __syncthreads();
    // ---- dW3 (384 entries: thread owns t and t + 256), db3; dZ2 = (dy W3^T) s'(Z2)
    for (int r = 0; r < 2; ++r) {
      const int e = tid + 256 * r;
      if (e < kMlpHid * kMlpOut) {
        const int h = e / kMlpOut, o = e % kMlpOut;
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g += silu(Z2[s * kMlpHid + h]) * dys[s * 4 + o];
        gW3[r] += g;
      }
    }
    if (tid < kMlpOut)
      for (int s = 0; s < TS; ++s) gb3 += dys[s * 4 + tid];
    __syncthreads();  // every dW3 read of Z2 precedes its overwrite with dZ2 below
    for (int e = tid; e < TS * kMlpHid; e += 256) {
      const int s = e / kMlpHid, h = e % kMlpHid;
      const float dh = dys[s * 4] * a.w[kOffW3 + h * 3] + dys[s * 4 + 1] * a.w[kOffW3 + h * 3 + 1] +
                       dys[s * 4 + 2] * a.w[kOffW3 + h * 3 + 2];
      const float dz = dh * dsilu(Z2[e]);
      Z2[e]          = dz;  // Z2 now holds dZ2 (FP32)
      const unsigned short b = f2bf(dz);
      D2s[s * LH + h] = b;
      D2t[h * LT + s] = b;
    }
    __syncthreads();
    // ---- db2; dW2[i][o] += sum_s H1[s][i] dZ2[s][o]: 16 tiles (4 per wave), K = TS
    if (tid < kMlpHid)
      for (int s = 0; s < TS; ++s) gb2 += Z2[s * kMlpHid + tid];
    for (int t = 0; t < 4; ++t) {
      const int ti = (wave * 4 + t) >> 2, to = (wave * 4 + t) & 3;  // i-tile, o-tile
      mfma_tile(H1t + (ti * 32) * LT, LT, D2t + (to * 32) * LT, LT, TS, gW2[t]);
    }
    // ---- dZ1 = (dZ2 W2^T) s'(Z1): wave owns columns [32 wave, +32) of dH1, K = 128
    {
      f32x16 acc = f32x16{};
      mfma_tile(D2s, LH, a.w2 + (wave * 32) * kMlpHid, kMlpHid, kMlpHid, acc);
      const int n = wave * 32 + col;
      for (int q = 0; q < 16; ++q) {
        const int s  = acc_row(q);
        const float dz = acc[q] * dsilu(Z1[s * kMlpHid + n]);
        Z1[s * kMlpHid + n] = dz;  // Z1 now holds dZ1
        D1t[n * LT + s] = f2bf(dz);
      }
    }
    __syncthreads();
    // ---- db1; dW1[f][h] += sum_s X[s][f] dZ1[s][h]: 8 tiles (2 per wave)
    if (tid < kMlpHid)
      for (int s = 0; s < TS; ++s) gb1 += Z1[s * kMlpHid + tid];
    for (int t = 0; t < 2; ++t) {
      const int tf = (wave * 2 + t) >> 2, th = (wave * 2 + t) & 3;
      mfma_tile(Xt + (tf * 32) * LT, LT, D1t + (th * 32) * LT, LT, TS, gW1[t]);
    }
    __syncthreads();
  }
  if (!TRAIN) return;
  // ---- write this block's partial gradient
  float* P = a.partials + (size_t)blockIdx.x * (kMlpParams + 1);
  for (int t = 0; t < 4; ++t) {
    const int ti = (wave * 4 + t) >> 2, to = (wave * 4 + t) & 3;
    for (int q = 0; q < 16; ++q) {
      const int i = ti * 32 + acc_row(q), o = to * 32 + col;
      P[kOffW2 + i * kMlpHid + o] = gW2[t][q];
    }
  }
  for (int t = 0; t < 2; ++t) {
    const int tf = (wave * 2 + t) >> 2, th = (wave * 2 + t) & 3;
    for (int q = 0; q < 16; ++q) {
      const int f = tf * 32 + acc_row(q), h = th * 32 + col;
      P[kOffW1 + f * kMlpHid + h] = gW1[t][q];
    }
  }
  for (int r = 0; r < 2; ++r) {
    const int e = tid + 256 * r;
    if (e < kMlpHid * kMlpOut) P[kOffW3 + e] = gW3[r];
  }
  if (tid < kMlpHid) {
    P[kOffB1 + tid] = gb1;
    P[kOffB2 + tid] = gb2;
  }
  if (tid < kMlpOut) P[kOffB3 + tid] = gb3;
  // loss: threads < TS hold partial sums
  if (tid < 64) {
    float v = lsum;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (tid == 0) lossw[0] = v;
  }
  __syncthreads();
  if (tid == 0) P[kMlpParams] = lossw[0];
}

__global__ __launch_bounds__(256) void k_sum_partials_f(int nblk, int len, const float* __restrict__ P,
                                                        float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += P[(size_t)b * len + e];
  out[e] = s;
}

__global__ void k_refresh_bf16(const float* __restrict__ w, unsigned short* w1t, unsigned short* w1,
                               unsigned short* w2t, unsigned short* w2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kMlpIn * kMlpHid) {
    const int f = e / kMlpHid, h = e % kMlpHid;
    const unsigned short b = f2bf(w[kOffW1 + e]);
    w1[e]               = b;
    w1t[h * kMlpIn + f] = b;
  }
  if (e < kMlpHid * kMlpHid) {
    const int i = e / kMlpHid, o = e % kMlpHid;
    const unsigned short b = f2bf(w[kOffW2 + e]);
    w2[e]                = b;
    w2t[o * kMlpHid + i] = b;
  }
}

// Adam (Kingma & Ba 2015), bias-corrected, on the FP32 master weights.
__global__ void k_adam(int np, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e]           = mi;
  v[e]           = vi;
  w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_normalise(long long ns, const float* __restrict__ X, const float* __restrict__ mean,
                            const float* __restrict__ inv_std, unsigned short* __restrict__ Xb) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ns * kMlpIn) return;
  const int f = (int)(e % kMlpIn);
  Xb[e]       = f2bf((X[e] - mean[f]) * inv_std[f]);
}

Mlp::~Mlp() {}

void mlp_refresh_bf16(Mlp& M) {
  M.w1t.alloc(kMlpIn * kMlpHid);
  M.w1.alloc(kMlpIn * kMlpHid);
  M.w2t.alloc(kMlpHid * kMlpHid);
  M.w2.alloc(kMlpHid * kMlpHid);
  hipLaunchKernelGGL(k_refresh_bf16, dim3(ceil_div(kMlpHid * kMlpHid, 256)), dim3(256), 0, M.stream, M.w.p,
                     M.w1t.p, M.w1.p, M.w2t.p, M.w2.p);
}

// Forward pass only (the warm-start guess): the arithmetic of k_mlp_tile<false> -- the same MFMA
// tiles in the same k order, the same roundings and layer-3 summation order, so the same outputs --
// without the training copies. Blocks are persistent over tiles: each wavefront loads its 32
// columns of W1 and W2 as MFMA operands into registers once, W3 and b3 sit in LDS, and s(Z2) is
// kept transposed (stride TS + 1) so that the 96 layer-3 threads read it without bank conflicts.
constexpr int LS2 = TS + 1;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_mlp_fwd(MlpTileArgs a) {
  __shared__ __attribute__((aligned(16))) unsigned short Xs[TS * LX];
  __shared__ __attribute__((aligned(16))) unsigned short H1s[TS * LH];
  __shared__ float S2t[kMlpHid * LS2];
  __shared__ float W3s[kMlpHid * kMlpOut + kMlpOut];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int col = acc_col(), r = l & 31, kh = (l >> 5) * 8;
  const int n = wave * 32 + col;
  bf16x8 w1v[kMlpIn / 16], w2v[kMlpHid / 16];
#pragma unroll
  for (int k = 0; k < kMlpIn / 16; ++k) w1v[k] = ld8(a.w1t + (wave * 32 + r) * kMlpIn + 16 * k + kh);
#pragma unroll
  for (int k = 0; k < kMlpHid / 16; ++k) w2v[k] = ld8(a.w2t + (wave * 32 + r) * kMlpHid + 16 * k + kh);
  const float bias1 = a.w[kOffB1 + n], bias2 = a.w[kOffB2 + n];
  for (int e = tid; e < kMlpHid * kMlpOut + kMlpOut; e += 256) W3s[e] = a.w[kOffW3 + e];  // W3 then b3
  // thread tid holds feature f = tid % 64 of samples tid / 64 + 4 k (k < 8) of a tile; the next
  // tile's are loaded while the current one is evaluated
  static_assert(TS * kMlpIn == 8 * 256, "eight features per thread and tile");
  const int f = tid & 63, sb = tid >> 6;
  const float mean_f = a.mean[f], istd_f = a.inv_std[f];
  float xr[8];
  auto fetch = [&](int tile) {
    const long long s0 = (long long)tile * TS;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const long long sk = s0 + sb + 4 * k;
      xr[k]              = tile < a.ntiles && sk < a.ns ? a.Xraw[sk * kMlpIn + f] : 0.0f;
    }
  };
  fetch(blockIdx.x);
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const long long s0 = (long long)tile * TS;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int s = sb + 4 * k;
      Xs[s * LX + f] = s0 + s < a.ns ? f2bf((xr[k] - mean_f) * istd_f) : (unsigned short)0;
    }
    __syncthreads();
    fetch(tile + gridDim.x);
    {
      f32x16 acc = f32x16{};
#pragma unroll
      for (int k = 0; k < kMlpIn / 16; ++k)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(Xs + r * LX + 16 * k + kh), w1v[k], acc, 0, 0, 0);
      for (int q = 0; q < 16; ++q) H1s[acc_row(q) * LH + n] = f2bf(silu(acc[q] + bias1));
    }
    __syncthreads();
    {
      f32x16 acc = f32x16{};
#pragma unroll
      for (int k = 0; k < kMlpHid / 16; ++k)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(H1s + r * LH + 16 * k + kh), w2v[k], acc, 0, 0, 0);
      for (int q = 0; q < 16; ++q) S2t[n * LS2 + acc_row(q)] = silu(acc[q] + bias2);
    }
    __syncthreads();
    if (tid < TS * kMlpOut) {
      const int s = tid / 3, o = tid % 3;
      float acc = W3s[kMlpHid * kMlpOut + o];
      for (int h = 0; h < kMlpHid; ++h) acc += S2t[h * LS2 + s] * W3s[h * kMlpOut + o];
      if (s0 + s < a.ns) a.y[(s0 + s) * 3 + o] = acc;
    }
    __syncthreads();
  }
}

void mlp_forward(Mlp& M, const float* X, int ns, float* y, hipStream_t st) {
  MlpTileArgs a{};
  a.w = M.w.p;
  a.w1t = M.w1t.p;
  a.w1 = M.w1.p;
  a.w2t = M.w2t.p;
  a.w2 = M.w2.p;
  a.Xraw = X;
  a.mean = M.mean.p;
  a.inv_std = M.inv_std.p;
  a.ns = ns;
  a.ntiles = ceil_div(ns, TS);
  a.y = y;
  const int grid = std::min(a.ntiles, 1024);  // persistent: four blocks per CU (register-bound)
  hipLaunchKernelGGL(k_mlp_fwd, dim3(std::max(grid, 1)), dim3(256), 0, st, a);
}

void mlp_loss_grad(Mlp& M, const unsigned short* Xb, const float* V, const float* T, const float* Wt, long long ns,
                   const int* tile_order, int ntiles, float inv_norm, double* loss_out, hipStream_t st) {
  MlpTileArgs a{};
  a.w = M.w.p;
  a.w1t = M.w1t.p;
  a.w1 = M.w1.p;
  a.w2t = M.w2t.p;
  a.w2 = M.w2.p;
  a.Xb = Xb;
  a.V = V;
  a.T = T;
  a.Wt = Wt;
  a.ns = ns;
  a.tile_order = tile_order;
  a.ntiles = ntiles >= 0 ? ntiles : ceil_div(ns, TS);
  a.inv_norm = inv_norm > 0.f ? inv_norm : 1.0f / (float)ns;
  const int grid = std::max(1, std::min(a.ntiles, 256));
  M.partials.alloc((size_t)grid * (kMlpParams + 1));
  a.partials = M.partials.p;
  hipLaunchKernelGGL(k_mlp_tile<true>, dim3(grid), dim3(256), 0, st, a);
  M.grad.alloc(kMlpParams + 1);
  hipLaunchKernelGGL(k_sum_partials_f, dim3(ceil_div(kMlpParams + 1, 256)), dim3(256), 0, st, grid,
                     kMlpParams + 1, M.partials.p, M.grad.p);
  if (loss_out) {
    float l = 0.f;
    CPPE_HIP_CHECK(hipMemcpyAsync(&l, M.grad.p + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, st));
    CPPE_HIP_CHECK(hipStreamSynchronize(st));
    *loss_out = l;
  }
}

void mlp_adam(Mlp& M, float lr, hipStream_t st) {
  M.step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)M.step), c2 = 1.f - std::pow(b2, (float)M.step);
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(kMlpParams, 256)), dim3(256), 0, st, kMlpParams, M.w.p, M.m.p, M.v.p,
                     M.grad.p, lr, b1, b2, c1, c2, 1e-8f);
  mlp_refresh_bf16(M);
}

void mlp_normalise(Mlp& M, const float* X, long long ns, unsigned short* Xb, hipStream_t st) {
  hipLaunchKernelGGL(k_normalise, dim3(ceil_div(ns * kMlpIn, 256)), dim3(256), 0, st, ns, X, M.mean.p,
                     M.inv_std.p, Xb);
}

}  // namespace cppe_dev
