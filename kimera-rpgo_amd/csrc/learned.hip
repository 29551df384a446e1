// Learned GNC initialisation (no counterpart in the reference): per-loop-closure
// features from the PCM state, and a 16 -> 128 -> 128 -> 2 MLP (ReLU) that
// predicts (initial-weight logit, log mu) per loop closure, trained on the
// weights the plain GNC converges to.  Forward, backward and the weight
// gradients run on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16: BF16
// operands, FP32 accumulation); the optimizer is Adam on FP32 master weights.
//
// MFMA 16x16x32 operand layout (one wavefront, lane l, q = l / 16, r = l % 16;
// checked on MI355X against a host product):
//   A (M x K): lane holds A[r][8q + j], j = 0..7
//   B (K x N): lane holds B[8q + j][r]
//   D (M x N): lane holds D[4q + j][r], j = 0..3
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "replay64.h"
#include "rpgo_internal.h"

namespace rpgo {

typedef short bf16x8 __attribute__((ext_vector_type(8)));  // BF16 bits
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int kF = RPGO_LC_FEATURES, kH = RPGO_MLP_HIDDEN, kO = RPGO_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kF * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kO, kNP = oB3 + kO;
static_assert(kNP == RPGO_MLP_NPARAMS, "parameter layout");

struct MlpState {
  DevBuf P, G, M, V;          // FP32 parameters, gradients, Adam moments
  DevBuf W1t, W2t, W2b, W3t;  // BF16 copies (see k_pack)
  DevBuf norm;                // feature mean[16], inv_std[16]
  DevBuf X, T, SW, perm;      // training set (device resident)
  DevBuf xin, Y, Xb, H1, H2, dY, dYb, dH1, dH2, loss;
  int steps = 0;
  bool have_params = false;
};

__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f32x4 mfma32(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c,
                                                 0, 0, 0);
}

__device__ __forceinline__ double wave_sum_l(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------- features
// One thread per accepted loop closure a (compacted order of the last PCM run):
//  0 log(1 + odometry-chain Mahalanobis distance)          (Pcm.h:604-629)
//  1 degree / (m - 1) in the consistency graph             (row statistics)
//  2 log(1 + degree)
//  3 consistent clique members / k
//  4 clique membership
//  5 log(1 + chi^2 of the loop closure at the chain poses)
//  6 log(1 + r / barc), 7 min(r / barc, 10): residual r = 1/2 e^T Lambda e of
//    the loop closure after GNC's first solve (all weights 1), when given
//  8 log(1 + |t| of the measurement), 9 rotation angle of the measurement
// 10 |i - j| / N, 11 log(1 + |i - j|)
// 12 log m, 13 log N
// 14 log(1 + distance of the two chain poses)
// 15 log(1 + translation residual at the chain poses)
__global__ void __launch_bounds__(128) k_lc_features(int m, int n_poses, int W, const double* __restrict__ traj,
                              const int* __restrict__ c_from, const int* __restrict__ c_to,
                              const double* __restrict__ c_lcm, const int* __restrict__ accepted,
                              const double* __restrict__ odom_dist, const uint64_t* __restrict__ adj,
                              const uint64_t* __restrict__ cmask, int k_clique,
                              const double* __restrict__ resid, double barc, float* __restrict__ feat) {
  using G = R64<3>;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= m) return;
  const int i = c_from[a], j = c_to[a];
  const double* Ti = traj + (size_t)i * G::PWC;
  const double* Tj = traj + (size_t)j * G::PWC;
  const double* Z = c_lcm + (size_t)a * G::PWC;
  int deg = 0, inc = 0;
  for (int w = 0; w < W; ++w) {
    const uint64_t row = adj[(size_t)a * W + w];
    deg += __popcll(row);
    inc += __popcll(row & cmask[w]);
  }
  const bool member = (cmask[a >> 6] >> (a & 63)) & 1ull;
  double Tii[12], Tij[12], Zi[12], E[12], e[6];
  G::pose_inverse(Ti, Tii);
  G::pose_compose(Tii, Tj, Tij);
  G::pose_inverse(Z, Zi);
  G::pose_compose(Zi, Tij, E);
  G::logmap(E, e);
  double Linv[36];
  lu_inverse64<6>(Z + 12, Linv);
  double chi2 = 0.0;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += Linv[r * 6 + c] * e[c];
    chi2 += e[r] * s;
  }
  const double rb = resid ? fmax(resid[a], 0.0) / fmax(barc, 1e-300) : 0.0;
  const double tn = sqrt(Z[9] * Z[9] + Z[10] * Z[10] + Z[11] * Z[11]);
  const double ang = acos(fmin(1.0, fmax(-1.0, 0.5 * (Z[0] + Z[4] + Z[8] - 1.0))));
  const double gap = fabs((double)(i - j));
  const double dpos = sqrt(Tij[9] * Tij[9] + Tij[10] * Tij[10] + Tij[11] * Tij[11]);
  const double vres = sqrt(e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
  float f[kF];
  f[0] = (float)log1p(fmax(odom_dist[accepted[a]], 0.0));
  f[1] = (float)(deg / fmax(1.0, (double)(m - 1)));
  f[2] = (float)log1p((double)deg);
  f[3] = (float)(inc / fmax(1.0, (double)k_clique));
  f[4] = member ? 1.f : 0.f;
  f[5] = (float)log1p(fmax(chi2, 0.0));
  f[6] = (float)log1p(rb);
  f[7] = (float)fmin(rb, 10.0);
  f[8] = (float)log1p(tn);
  f[9] = (float)ang;
  f[10] = (float)(gap / fmax(1.0, (double)n_poses));
  f[11] = (float)log1p(gap);
  f[12] = (float)log(fmax(1.0, (double)m));
  f[13] = (float)log(fmax(1.0, (double)n_poses));
  f[14] = (float)log1p(dpos);
  f[15] = (float)log1p(vres);
#pragma unroll
  for (int c = 0; c < kF; ++c) feat[(size_t)a * kF + c] = isfinite(f[c]) ? f[c] : 0.f;
}

// ---------------------------------------------------------------- MLP kernels
// BF16 copies: W1t[n][k] (n < 128, k < 16), W2t[n][k] and W2b[k][n] (128 x 128),
// W3t[o][k] (16 rows, o >= 2 zero), so that every lane reads 4 consecutive
// BF16 of its MFMA operand.
__global__ void k_pack(const float* __restrict__ P, bf16* __restrict__ W1t, bf16* __restrict__ W2t,
                       bf16* __restrict__ W2b, bf16* __restrict__ W3t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kF * kH) {
    const int k = t / kH, n = t % kH;
    W1t[n * kF + k] = f2bf(P[oW1 + t]);
  }
  if (t < kH * kH) {
    const int k = t / kH, n = t % kH;
    const bf16 v = f2bf(P[oW2 + t]);
    W2t[n * kH + k] = v;
    W2b[k * kH + n] = v;
  }
  if (t < 16 * kH) {
    const int o = t / kH, k = t % kH;
    W3t[o * kH + k] = o < kO ? f2bf(P[oW3 + k * kO + o]) : (bf16)0;
  }
}

// Forward pass, 64 rows per workgroup (one wavefront per 16 rows).  Rows are
// gathered through idx when given.  Saves the BF16 layer inputs (Xb, H1, H2)
// for the backward pass when the pointers are non-null.
constexpr int kLdA = kH + 8;  // LDS row pitch of the activations (BF16)

__global__ void __launch_bounds__(256) k_mlp_forward(int n, const float* __restrict__ X, const int* __restrict__ idx,
                                                     const float* __restrict__ norm, const bf16* __restrict__ W1t,
                                                     const bf16* __restrict__ W2t, const bf16* __restrict__ W3t,
                                                     const float* __restrict__ P, float* __restrict__ Y,
                                                     bf16* __restrict__ Xb, bf16* __restrict__ H1,
                                                     bf16* __restrict__ H2) {
  __shared__ bf16 sW2[kH * kH];
  __shared__ bf16 act1[4][16][kLdA];
  __shared__ bf16 act2[4][16][kLdA];
  for (int e = threadIdx.x; e < kH * kH / 4; e += blockDim.x)
    reinterpret_cast<uint2*>(sW2)[e] = reinterpret_cast<const uint2*>(W2t)[e];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int row0 = blockIdx.x * 64 + w * 16;
  const float* b1 = P + oB1;
  const float* b2 = P + oB2;
  const float* b3 = P + oB3;
  // layer 1: A = normalised features (16 x 16, K zero-padded to 32), B = W1 (16 x 128)
  bf16x8 a;
  {
    const int row = row0 + r;
    const bool ok = row < n && q < 2;
    const int src = ok ? (idx ? idx[row] : row) : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = (8 * q + j) & (kF - 1);
      const float x = ok ? (X[(size_t)src * kF + c] - norm[c]) * norm[kF + c] : 0.f;
      const bf16 h = f2bf(x);
      a[j] = (short)h;
      if (ok && Xb) Xb[(size_t)row * kF + c] = h;
    }
  }
#pragma unroll
  for (int nt = 0; nt < kH / 16; ++nt) {
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = q < 2 ? (short)W1t[(nt * 16 + r) * kF + 8 * q + j] : (short)0;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma32(a, b, acc);
    const int col = nt * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bf16 h = f2bf(fmaxf(acc[j] + b1[col], 0.f));
      act1[w][4 * q + j][col] = h;
      const int row = row0 + 4 * q + j;
      if (H1 && row < n) H1[(size_t)row * kH + col] = h;
    }
  }
  __syncthreads();
  // layer 2: A = H1 (16 x 128), B = W2 (128 x 128)
#pragma unroll
  for (int nt = 0; nt < kH / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < kH / 32; ++kt) {
      bf16x8 aa, bb;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        aa[j] = (short)act1[w][r][kt * 32 + 8 * q + j];
        bb[j] = (short)sW2[(nt * 16 + r) * kH + kt * 32 + 8 * q + j];
      }
      acc = mfma32(aa, bb, acc);
    }
    const int col = nt * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bf16 h = f2bf(fmaxf(acc[j] + b2[col], 0.f));
      act2[w][4 * q + j][col] = h;
      const int row = row0 + 4 * q + j;
      if (H2 && row < n) H2[(size_t)row * kH + col] = h;
    }
  }
  __syncthreads();
  // layer 3: A = H2 (16 x 128), B = W3 padded to 128 x 16
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < kH / 32; ++kt) {
    bf16x8 aa, bb;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      aa[j] = (short)act2[w][r][kt * 32 + 8 * q + j];
      bb[j] = (short)W3t[r * kH + kt * 32 + 8 * q + j];
    }
    acc = mfma32(aa, bb, acc);
  }
  if (r < kO) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = row0 + 4 * q + j;
      if (row < n) Y[(size_t)row * kO + r] = acc[j] + b3[r];
    }
  }
}

// Loss and output gradient, one thread per row (mean over the batch):
// BCE-with-logits on the weight target, (mu_w / 2)(y1 - t1)^2 on log mu,
// both scaled by the sample weight.
__global__ void k_mlp_loss(int n, const int* __restrict__ idx, const float* __restrict__ Y,
                           const float* __restrict__ T, const float* __restrict__ SW, float mu_w,
                           float* __restrict__ dY, bf16* __restrict__ dYb, double* __restrict__ loss) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  double lsum = 0.0;
  if (row < n) {
    const int src = idx ? idx[row] : row;
    const float sw = SW ? SW[src] : 1.f;
    const float inv = 1.f / (float)n;
    const float z = Y[(size_t)row * kO], t0 = T[(size_t)src * kO];
    const float sp = z > 0.f ? z + log1pf(expf(-z)) : log1pf(expf(z));
    const float sig = 1.f / (1.f + expf(-z));
    const float m = Y[(size_t)row * kO + 1], t1 = T[(size_t)src * kO + 1];
    lsum = (double)(sw * (sp - t0 * z) + sw * 0.5f * mu_w * (m - t1) * (m - t1)) * inv;
    const float g0 = sw * (sig - t0) * inv, g1 = sw * mu_w * (m - t1) * inv;
    dY[(size_t)row * kO] = g0;
    dY[(size_t)row * kO + 1] = g1;
#pragma unroll
    for (int c = 0; c < 16; ++c) dYb[(size_t)row * 16 + c] = c == 0 ? f2bf(g0) : c == 1 ? f2bf(g1) : (bf16)0;
  }
  lsum = wave_sum_l(lsum);
  if ((threadIdx.x & 63) == 0 && lsum != 0.0) atomicAdd(loss, lsum);
}

// dH2 = relu'(H2) * (dY W3^T), one thread per (row, unit); W3 as its BF16 copy
__global__ void k_mlp_back_h2(int n, const float* __restrict__ dY, const bf16* __restrict__ W3t,
                              const bf16* __restrict__ H2, bf16* __restrict__ dH2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * kH) return;
  const int row = (int)(t / kH), c = (int)(t % kH);
  float g = 0.f;
#pragma unroll
  for (int o = 0; o < kO; ++o) g += dY[(size_t)row * kO + o] * bf2f(W3t[o * kH + c]);
  dH2[t] = bf2f(H2[t]) > 0.f ? f2bf(g) : (bf16)0;
}

// dH1 = relu'(H1) * (dH2 W2^T) on the matrix cores, 64 rows per workgroup
__global__ void __launch_bounds__(256) k_mlp_back_h1(int n, const bf16* __restrict__ dH2,
                                                     const bf16* __restrict__ W2b, const bf16* __restrict__ H1,
                                                     bf16* __restrict__ dH1) {
  __shared__ bf16 sW2[kH * kH];
  for (int e = threadIdx.x; e < kH * kH / 4; e += blockDim.x)
    reinterpret_cast<uint2*>(sW2)[e] = reinterpret_cast<const uint2*>(W2b)[e];
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int row0 = blockIdx.x * 64 + w * 16;
  const int arow = row0 + r;
#pragma unroll
  for (int nt = 0; nt < kH / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < kH / 32; ++kt) {
      bf16x8 aa, bb;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        aa[j] = arow < n ? (short)dH2[(size_t)arow * kH + kt * 32 + 8 * q + j] : (short)0;
        bb[j] = (short)sW2[(nt * 16 + r) * kH + kt * 32 + 8 * q + j];
      }
      acc = mfma32(aa, bb, acc);
    }
    const int col = nt * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = row0 + 4 * q + j;
      if (row < n) dH1[(size_t)row * kH + col] = bf2f(H1[(size_t)row * kH + col]) > 0.f ? f2bf(acc[j]) : (bf16)0;
    }
  }
}

// dW[m0 + i][n0 + c] += sum_rows A[row][m0 + i] * B[row][n0 + c] on the matrix
// cores: one 16 x 16 output tile per workgroup (grid.x, grid.y), the rows split
// over grid.z and the 4 wavefronts; columns >= ncols are not stored.
__global__ void __launch_bounds__(256) k_mlp_wgrad(int n, const bf16* __restrict__ A, int lda,
                                                   const bf16* __restrict__ B, int ldb, float* __restrict__ dW,
                                                   int ldw, int ncols) {
  __shared__ float part[4][16][17];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, r = l & 15;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  const int chunk = (n + gridDim.z - 1) / gridDim.z;
  const int rb = blockIdx.z * chunk, re = min(n, rb + chunk);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int r0 = rb + w * 32; r0 < re; r0 += 128) {
    bf16x8 aa, bb;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int row = r0 + 8 * q + j;
      const bool ok = row < re;
      aa[j] = ok ? (short)A[(size_t)row * lda + m0 + r] : (short)0;
      bb[j] = ok ? (short)B[(size_t)row * ldb + n0 + r] : (short)0;
    }
    acc = mfma32(aa, bb, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) part[w][4 * q + j][r] = acc[j];
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = 4 * q + j;
      const float v = part[0][i][r] + part[1][i][r] + part[2][i][r] + part[3][i][r];
      if (n0 + r < ncols) atomicAdd(&dW[(size_t)(m0 + i) * ldw + n0 + r], v);
    }
  }
}

// bias gradients: column sums of dH1, dH2 (BF16) and dY (FP32), one workgroup per column
__global__ void __launch_bounds__(256) k_mlp_bgrad(int n, const bf16* __restrict__ dH1,
                                                   const bf16* __restrict__ dH2, const float* __restrict__ dY,
                                                   float* __restrict__ G) {
  __shared__ double red[4];
  const int col = blockIdx.x;  // 0..127: b1, 128..255: b2, 256..257: b3
  double s = 0.0;
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    if (col < kH) s += bf2f(dH1[(size_t)row * kH + col]);
    else if (col < 2 * kH) s += bf2f(dH2[(size_t)row * kH + col - kH]);
    else s += dY[(size_t)row * kO + col - 2 * kH];
  }
  s = wave_sum_l(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = red[0] + red[1] + red[2] + red[3];
    const int o = col < kH ? oB1 + col : col < 2 * kH ? oB2 + col - kH : oB3 + col - 2 * kH;
    G[o] += (float)t;
  }
}

__global__ void k_adam(int np, float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const float g = G[t];
  const float m = b1 * M[t] + (1.f - b1) * g;
  const float v = b2 * V[t] + (1.f - b2) * g * g;
  M[t] = m;
  V[t] = v;
  P[t] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ---------------------------------------------------------------- host side
static MlpState& mlp_state(rpgo_ctx* ctx) {
  if (!ctx->mlp) ctx->mlp = new MlpState();
  return *ctx->mlp;
}

void mlp_destroy(rpgo_ctx* ctx) {
  if (!ctx->mlp) return;
  MlpState& s = *ctx->mlp;
  for (DevBuf* b : {&s.P, &s.G, &s.M, &s.V, &s.W1t, &s.W2t, &s.W2b, &s.W3t, &s.norm, &s.X, &s.T, &s.SW, &s.perm,
                    &s.xin, &s.Y, &s.Xb, &s.H1, &s.H2, &s.dY, &s.dYb, &s.dH1, &s.dH2, &s.loss})
    b->release();
  delete ctx->mlp;
  ctx->mlp = nullptr;
}

static int mlp_batch_buffers(MlpState& s, int n) {
  int rc;
  if ((rc = s.Y.reserve(sizeof(float) * kO * (size_t)n))) return rc;
  if ((rc = s.Xb.reserve(sizeof(bf16) * kF * (size_t)n))) return rc;
  if ((rc = s.H1.reserve(sizeof(bf16) * kH * (size_t)n))) return rc;
  if ((rc = s.H2.reserve(sizeof(bf16) * kH * (size_t)n))) return rc;
  if ((rc = s.dY.reserve(sizeof(float) * kO * (size_t)n))) return rc;
  if ((rc = s.dYb.reserve(sizeof(bf16) * 16 * (size_t)n))) return rc;
  if ((rc = s.dH1.reserve(sizeof(bf16) * kH * (size_t)n))) return rc;
  if ((rc = s.dH2.reserve(sizeof(bf16) * kH * (size_t)n))) return rc;
  if ((rc = s.loss.reserve(sizeof(double) * 4096))) return rc;
  return RPGO_OK;
}

static void launch_pack(MlpState& s, hipStream_t st) {
  k_pack<<<(kH * kH + 255) / 256, 256, 0, st>>>(s.P.as<float>(), s.W1t.as<bf16>(), s.W2t.as<bf16>(),
                                                 s.W2b.as<bf16>(), s.W3t.as<bf16>());
}

// forward (+ saved activations), loss, backward and gradients for n rows
// gathered through idx (device) from the data set X/T/SW (device); G is
// accumulated (cleared by the caller); the loss adds into *loss_dev.
static void mlp_grad_pass(MlpState& s, hipStream_t st, int n, const float* X, const int* idx, const float* T,
                          const float* SW, float mu_w, double* loss_dev) {
  const unsigned blocks = (unsigned)((n + 63) / 64);
  k_mlp_forward<<<blocks, 256, 0, st>>>(n, X, idx, s.norm.as<float>(), s.W1t.as<bf16>(), s.W2t.as<bf16>(),
                                        s.W3t.as<bf16>(), s.P.as<float>(), s.Y.as<float>(), s.Xb.as<bf16>(),
                                        s.H1.as<bf16>(), s.H2.as<bf16>());
  k_mlp_loss<<<(n + 255) / 256, 256, 0, st>>>(n, idx, s.Y.as<float>(), T, SW, mu_w, s.dY.as<float>(),
                                              s.dYb.as<bf16>(), loss_dev);
  k_mlp_back_h2<<<(unsigned)(((long long)n * kH + 255) / 256), 256, 0, st>>>(n, s.dY.as<float>(), s.W3t.as<bf16>(),
                                                                          s.H2.as<bf16>(), s.dH2.as<bf16>());
  k_mlp_back_h1<<<blocks, 256, 0, st>>>(n, s.dH2.as<bf16>(), s.W2b.as<bf16>(), s.H1.as<bf16>(), s.dH1.as<bf16>());
  const int zs = std::max(1, std::min(64, n / 1024));
  float* G = s.G.as<float>();
  k_mlp_wgrad<<<dim3(kF / 16, kH / 16, zs), 256, 0, st>>>(n, s.Xb.as<bf16>(), kF, s.dH1.as<bf16>(), kH, G + oW1,
                                                         kH, kH);
  k_mlp_wgrad<<<dim3(kH / 16, kH / 16, zs), 256, 0, st>>>(n, s.H1.as<bf16>(), kH, s.dH2.as<bf16>(), kH, G + oW2,
                                                         kH, kH);
  k_mlp_wgrad<<<dim3(kH / 16, 1, zs), 256, 0, st>>>(n, s.H2.as<bf16>(), kH, s.dYb.as<bf16>(), 16, G + oW3, kO,
                                                   kO);
  k_mlp_bgrad<<<2 * kH + kO, 256, 0, st>>>(n, s.dH1.as<bf16>(), s.dH2.as<bf16>(), s.dY.as<float>(), G);
}

static int mlp_require(rpgo_ctx* ctx) {
  if (!ctx->mlp || !ctx->mlp->have_params) {
    set_error("MLP parameters not set (rpgo_mlp_set_params)");
    return RPGO_ERR_STATE;
  }
  return RPGO_OK;
}

}  // namespace rpgo

using namespace rpgo;

extern "C" {

int rpgo_lc_features(rpgo_ctx* ctx, const int32_t* clique_idx, int32_t k, const double* lc_resid, double barc,
                     float* feat) {
  if (!ctx || !feat || k < 0 || (k > 0 && !clique_idx)) {
    set_error("rpgo_lc_features: bad argument");
    return RPGO_ERR_ARG;
  }
  PcmState& s = ctx->pcm;
  if (!s.ran) {
    set_error("rpgo_lc_features: no PCM run in this context");
    return RPGO_ERR_STATE;
  }
  if (s.dim != 3) {
    set_error("rpgo_lc_features: SE(3) graphs only");
    return RPGO_ERR_ARG;
  }
  const int m = s.m;
  if (m == 0) return RPGO_OK;
  const int W = (m + 63) / 64;
  std::vector<uint64_t> mask(W, 0ull);
  for (int t = 0; t < k; ++t) {
    if (clique_idx[t] < 0 || clique_idx[t] >= m) {
      set_error("rpgo_lc_features: clique index %d outside [0, %d)", clique_idx[t], m);
      return RPGO_ERR_ARG;
    }
    mask[clique_idx[t] >> 6] |= 1ull << (clique_idx[t] & 63);
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  DevBuf dmask, dfeat, dres;
  int rc;
  if ((rc = dmask.reserve(sizeof(uint64_t) * W)) || (rc = dfeat.reserve(sizeof(float) * kF * (size_t)m))) return rc;
  if (lc_resid && (rc = dres.reserve(sizeof(double) * (size_t)m))) return rc;
  hipStream_t st = ctx->stream;
  RPGO_HIP_CHECK(hipMemcpyAsync(dmask.ptr, mask.data(), sizeof(uint64_t) * W, hipMemcpyHostToDevice, st));
  if (lc_resid)
    RPGO_HIP_CHECK(hipMemcpyAsync(dres.ptr, lc_resid, sizeof(double) * (size_t)m, hipMemcpyHostToDevice, st));
  k_lc_features<<<(m + 127) / 128, 128, 0, st>>>(m, s.n_poses, W, s.traj.as<double>(), s.c_from.as<int>(),
                                                 s.c_to.as<int>(), s.c_lcm.as<double>(), s.accepted.as<int>(),
                                                 s.odom_dist.as<double>(), s.adj.as<uint64_t>(),
                                                 dmask.as<uint64_t>(), k, lc_resid ? dres.as<double>() : nullptr,
                                                 barc, dfeat.as<float>());
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipMemcpyAsync(feat, dfeat.ptr, sizeof(float) * kF * (size_t)m, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  dmask.release();
  dfeat.release();
  dres.release();
  return RPGO_OK;
}

int rpgo_mlp_set_params(rpgo_ctx* ctx, const float* params, const float* norm) {
  if (!ctx || !params || !norm) {
    set_error("rpgo_mlp_set_params: NULL argument");
    return RPGO_ERR_ARG;
  }
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  MlpState& s = mlp_state(ctx);
  int rc;
  for (DevBuf* b : {&s.P, &s.G, &s.M, &s.V})
    if ((rc = b->reserve(sizeof(float) * kNP))) return rc;
  if ((rc = s.W1t.reserve(sizeof(bf16) * kF * kH)) || (rc = s.W2t.reserve(sizeof(bf16) * kH * kH)) ||
      (rc = s.W2b.reserve(sizeof(bf16) * kH * kH)) || (rc = s.W3t.reserve(sizeof(bf16) * 16 * kH)) ||
      (rc = s.norm.reserve(sizeof(float) * 2 * kF)))
    return rc;
  hipStream_t st = ctx->stream;
  RPGO_HIP_CHECK(hipMemcpyAsync(s.P.ptr, params, sizeof(float) * kNP, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemcpyAsync(s.norm.ptr, norm, sizeof(float) * 2 * kF, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s.M.ptr, 0, sizeof(float) * kNP, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s.V.ptr, 0, sizeof(float) * kNP, st));
  launch_pack(s, st);
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  s.steps = 0;
  s.have_params = true;
  return RPGO_OK;
}

int rpgo_mlp_get_params(rpgo_ctx* ctx, float* params) {
  if (!ctx || !params) {
    set_error("rpgo_mlp_get_params: NULL argument");
    return RPGO_ERR_ARG;
  }
  int rc;
  if ((rc = mlp_require(ctx))) return rc;
  RPGO_HIP_CHECK(hipMemcpy(params, ctx->mlp->P.ptr, sizeof(float) * kNP, hipMemcpyDeviceToHost));
  return RPGO_OK;
}

int rpgo_mlp_forward(rpgo_ctx* ctx, int32_t n, const float* x, float* y) {
  if (!ctx || n < 0 || (n > 0 && (!x || !y))) {
    set_error("rpgo_mlp_forward: bad argument");
    return RPGO_ERR_ARG;
  }
  int rc;
  if ((rc = mlp_require(ctx))) return rc;
  if (n == 0) return RPGO_OK;
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  MlpState& s = *ctx->mlp;
  if ((rc = s.xin.reserve(sizeof(float) * kF * (size_t)n)) || (rc = s.Y.reserve(sizeof(float) * kO * (size_t)n)))
    return rc;
  hipStream_t st = ctx->stream;
  RPGO_HIP_CHECK(hipMemcpyAsync(s.xin.ptr, x, sizeof(float) * kF * (size_t)n, hipMemcpyHostToDevice, st));
  k_mlp_forward<<<(n + 63) / 64, 256, 0, st>>>(n, s.xin.as<float>(), nullptr, s.norm.as<float>(), s.W1t.as<bf16>(),
                                               s.W2t.as<bf16>(), s.W3t.as<bf16>(), s.P.as<float>(), s.Y.as<float>(),
                                               nullptr, nullptr, nullptr);
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipMemcpyAsync(y, s.Y.ptr, sizeof(float) * kO * (size_t)n, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  return RPGO_OK;
}

int rpgo_mlp_gradients(rpgo_ctx* ctx, int32_t n, const float* x, const float* target, const float* sample_w,
                       float mu_weight, float* grad, double* loss) {
  if (!ctx || n <= 0 || !x || !target || !grad) {
    set_error("rpgo_mlp_gradients: bad argument");
    return RPGO_ERR_ARG;
  }
  int rc;
  if ((rc = mlp_require(ctx))) return rc;
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  MlpState& s = *ctx->mlp;
  if ((rc = mlp_batch_buffers(s, n)) || (rc = s.X.reserve(sizeof(float) * kF * (size_t)n)) ||
      (rc = s.T.reserve(sizeof(float) * kO * (size_t)n)) || (rc = s.SW.reserve(sizeof(float) * (size_t)n)))
    return rc;
  hipStream_t st = ctx->stream;
  RPGO_HIP_CHECK(hipMemcpyAsync(s.X.ptr, x, sizeof(float) * kF * (size_t)n, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemcpyAsync(s.T.ptr, target, sizeof(float) * kO * (size_t)n, hipMemcpyHostToDevice, st));
  if (sample_w)
    RPGO_HIP_CHECK(hipMemcpyAsync(s.SW.ptr, sample_w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s.G.ptr, 0, sizeof(float) * kNP, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s.loss.ptr, 0, sizeof(double), st));
  mlp_grad_pass(s, st, n, s.X.as<float>(), nullptr, s.T.as<float>(), sample_w ? s.SW.as<float>() : nullptr,
                mu_weight, s.loss.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  RPGO_HIP_CHECK(hipMemcpyAsync(grad, s.G.ptr, sizeof(float) * kNP, hipMemcpyDeviceToHost, st));
  double l = 0.0;
  RPGO_HIP_CHECK(hipMemcpyAsync(&l, s.loss.ptr, sizeof(double), hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (loss) *loss = l;
  return RPGO_OK;
}

int rpgo_mlp_train(rpgo_ctx* ctx, int32_t n, const float* x, const float* target, const float* sample_w,
                   float mu_weight, int32_t batch, int32_t epochs, float lr, uint64_t seed, double* loss_hist) {
  if (!ctx || n <= 0 || !x || !target || batch <= 0 || epochs < 0 || !(lr > 0.f)) {
    set_error("rpgo_mlp_train: bad argument");
    return RPGO_ERR_ARG;
  }
  if (epochs > 4096) {
    set_error("rpgo_mlp_train: at most 4096 epochs per call");
    return RPGO_ERR_ARG;
  }
  int rc;
  if ((rc = mlp_require(ctx))) return rc;
  RPGO_HIP_CHECK(hipSetDevice(ctx->device));
  MlpState& s = *ctx->mlp;
  const int B = std::min(batch, n);
  if ((rc = mlp_batch_buffers(s, B)) || (rc = s.X.reserve(sizeof(float) * kF * (size_t)n)) ||
      (rc = s.T.reserve(sizeof(float) * kO * (size_t)n)) || (rc = s.SW.reserve(sizeof(float) * (size_t)n)) ||
      (rc = s.perm.reserve(sizeof(int) * (size_t)n)))
    return rc;
  hipStream_t st = ctx->stream;
  RPGO_HIP_CHECK(hipMemcpyAsync(s.X.ptr, x, sizeof(float) * kF * (size_t)n, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemcpyAsync(s.T.ptr, target, sizeof(float) * kO * (size_t)n, hipMemcpyHostToDevice, st));
  if (sample_w)
    RPGO_HIP_CHECK(hipMemcpyAsync(s.SW.ptr, sample_w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s.loss.ptr, 0, sizeof(double) * (size_t)epochs, st));
  std::mt19937_64 rng(seed);
  std::vector<int> perm(n);
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    // the upload must land before the previous epoch's kernels stop reading perm
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    RPGO_HIP_CHECK(hipMemcpyAsync(s.perm.ptr, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
    const int nb = n / B;  // full batches only
    for (int b = 0; b < std::max(nb, 1); ++b) {
      const int rows = nb ? B : n;
      RPGO_HIP_CHECK(hipMemsetAsync(s.G.ptr, 0, sizeof(float) * kNP, st));
      mlp_grad_pass(s, st, rows, s.X.as<float>(), s.perm.as<int>() + (size_t)b * B, s.T.as<float>(),
                    sample_w ? s.SW.as<float>() : nullptr, mu_weight, s.loss.as<double>() + e);
      ++s.steps;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)s.steps), c2 = 1.f - std::pow(b2, (float)s.steps);
      k_adam<<<(kNP + 255) / 256, 256, 0, st>>>(kNP, s.P.as<float>(), s.G.as<float>(), s.M.as<float>(),
                                               s.V.as<float>(), lr, b1, b2, 1e-8f, c1, c2);
      launch_pack(s, st);
    }
    RPGO_HIP_CHECK(hipGetLastError());
  }
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  if (loss_hist && epochs > 0) {
    std::vector<double> h(epochs);
    RPGO_HIP_CHECK(hipMemcpy(h.data(), s.loss.ptr, sizeof(double) * (size_t)epochs, hipMemcpyDeviceToHost));
    const int nb = std::max(n / B, 1);
    for (int e = 0; e < epochs; ++e) loss_hist[e] = h[e] / nb;
  }
  return RPGO_OK;
}

}  // extern "C"
