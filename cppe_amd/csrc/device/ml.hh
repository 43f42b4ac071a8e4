// Learned CG warm start: shapes and device entry points (no counterpart in the reference).
#pragma once
#include "system.hh"

namespace cppe_dev {

constexpr int kMlpIn = 64, kMlpHid = 128, kMlpOut = 3;
// FP32 master parameter layout: W1[in][hid], b1[hid], W2[hid][hid], b2[hid], W3[hid][out], b3[out]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kOffB3 + kMlpOut;  // 25,219

// Device-side model: FP32 master weights, Adam moments, BF16 operand copies, feature scaling.
struct Mlp {
  dbuf<float> w, m, v, grad;  // kMlpParams each
  dbuf<float> mean, inv_std;  // kMlpIn each
  dbuf<unsigned short> w1t, w1, w2t, w2;  // bf16: W1^T [hid][in], W1 [in][hid], W2^T, W2
  dbuf<float> partials;                    // per-block gradient partials
  int step = 0;
  hipStream_t stream = nullptr;
  // training set resident in HBM
  dbuf<unsigned short> dX;  // bf16 normalised features, ns x 64
  dbuf<float> dV;           // 9 floats per sample (v1, v2, v3 as xyz triples)
  dbuf<float> dT;           // 3 floats per sample (target moment)
  dbuf<float> dS;           // per-sample loss weight 1/|v1|^2
  long long ns = 0, cap = 0;
  ~Mlp();
};

struct GridInfo {
  double x0, y0, z0;  // lower corner
  int nx, ny, nz;
};

// Neighbour grid of one system for the descriptors, built on first use (positions never change),
// and the guess's work buffers.
struct MlGrid {
  GridInfo g;
  dbuf<int> cstart, sorted;
  dbuf<unsigned> cellof;
  // neighbour list within rc (CSR over sites, device order): indices only, the pair geometry is
  // recomputed from the site records in each pass
  long long npairs = 0;
  dbuf<long long> noff;  // n + 1
  dbuf<int> nj;          // neighbour index, bit 31 set for excluded pairs (no tensor block)
  dbuf<double4> pos;     // per site: position and damping rate damping * alpha_iso^(-1/6)
  dbuf<float> iso;       // alpha_iso of each site
  // per-site records gathered by the list kernels: (F / |F|, alpha_iso), and v1 / v2 in FP32
  dbuf<float4> frec, w1, w2;
  dbuf<float> X, y;      // n x 64 features, n x 3 network outputs
  dbuf<double> V;        // 9n bases
};
MlGrid& ml_grid(System& S);

void ml_features(System& S, float* X, double* V);
void mlp_refresh_bf16(Mlp& M);
// y (ns x 3) = MLP(normalise(X)) ; X raw FP32 features (device), ns samples
void mlp_forward(Mlp& M, const float* X, int ns, float* y, hipStream_t st);
// gradient of the mean loss over the given samples (already normalised bf16 features)
// tile_order: ntiles tile indices (32 samples each) or null with ntiles = -1 for all tiles;
// inv_norm: 1 / (samples in the minibatch), 0 for 1/ns
void mlp_loss_grad(Mlp& M, const unsigned short* Xb, const float* V, const float* T, const float* Wt, long long ns,
                   const int* tile_order, int ntiles, float inv_norm, double* loss_out, hipStream_t st);
void mlp_adam(Mlp& M, float lr, hipStream_t st);
void mlp_normalise(Mlp& M, const float* X, long long ns, unsigned short* Xb, hipStream_t st);

}  // namespace cppe_dev
