// Robust pose-graph solve on gfx950: GNC-TLS over Levenberg-Marquardt or
// Gauss-Newton, with block-Jacobi PCG for every linear step.
//
// Semantics follow src/RobustSolver.cpp:115-306 (GNC wiring, known inliers,
// thresholds, bias_odom initial weights) and the GTSAM algorithms it drives:
// GncOptimizer (TLS weights, mu init/update, convergence tests, every inner
// solve restarted from the initial values), LevenbergMarquardtOptimizer
// (diagonal damping, lambda schedule, model-fidelity acceptance) and
// BetweenFactor<Pose3> with SLOW_BUT_CORRECT_BETWEENFACTOR under the EXPMAP
// retraction: e = Log(z^-1 T_i^-1 T_j), J_j = Jr^-1(e),
// J_i = -Jr^-1(e) Ad((T_i^-1 T_j)^-1); PriorFactor e = -Log(x^-1 prior), J = I.
//
// Precision split: residuals, Jacobians, gradient and costs in FP64 (so the
// fixed point of the Gauss-Newton iteration is the FP64 one: each outer step
// is an iterative-refinement step), the 6x6 Hessian blocks and the PCG
// vectors in FP32, PCG dot products and norms accumulated in FP64.
//
// Layout in HBM: poses as FP64 (R row-major, t); factors SoA-by-record; the
// Gauss-Newton system in block form: one 6x6 FP32 diagonal block per pose and
// one off-diagonal block per between factor (its transpose implied), with a
// pose->incident-factor index (sorted by factor id) for deterministic
// row-parallel assembly and SpMV.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "replay64.h"
#include "rpgo_internal.h"
#include "rpgo_solver.h"

namespace rpgo {

using R3 = R64<3>;

// ------------------------------------------------------------------ FP64 SE(3)
__device__ __forceinline__ void skew3(const double* w, double* W) {
  W[0] = 0; W[1] = -w[2]; W[2] = w[1];
  W[3] = w[2]; W[4] = 0; W[5] = -w[0];
  W[6] = -w[1]; W[7] = w[0]; W[8] = 0;
}

__device__ __forceinline__ void mm3(const double* A, const double* B, double* C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      C[i * 3 + j] = A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j];
}

// Rodrigues + Pose3::Expmap closed form (t = (w x v - R (w x v) + w w.v) / |w|^2)
__device__ void se3_exp64(const double* xi, double* T) {
  const double* w = xi;
  const double* v = xi + 3;
  const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double th = sqrt(th2);
  double A, B;
  if (th2 < 1e-20) { A = 1.0 - th2 / 6.0; B = 0.5 - th2 / 24.0; }
  else { A = sin(th) / th; B = (1.0 - cos(th)) / th2; }
  double W[9], W2[9];
  skew3(w, W);
  mm3(W, W, W2);
#pragma unroll
  for (int e = 0; e < 9; ++e) T[e] = ((e % 4 == 0) ? 1.0 : 0.0) + A * W[e] + B * W2[e];
  if (th2 > 2.220446049250313e-16) {
    const double wv = w[0] * v[0] + w[1] * v[1] + w[2] * v[2];
    const double c[3] = {w[1] * v[2] - w[2] * v[1], w[2] * v[0] - w[0] * v[2], w[0] * v[1] - w[1] * v[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double Rc = T[i * 3] * c[0] + T[i * 3 + 1] * c[1] + T[i * 3 + 2] * c[2];
      T[9 + i] = (c[i] - Rc + w[i] * wv) / th2;
    }
  } else {
    T[9] = v[0]; T[10] = v[1]; T[11] = v[2];
  }
}

// SE(3) LogmapDerivative (Jr^-1): [[Jw, 0], [-Jw Q Jw, Jw]], Q of Barfoot14tro eq. 102
__device__ void se3_logmap_derivative64(const double* xi, double* J) {
  const double* w = xi;
  const double* v = xi + 3;
  const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double W[9], W2[9], Jw[9];
  skew3(w, W);
  mm3(W, W, W2);
  if (th2 <= 2.220446049250313e-16) {
#pragma unroll
    for (int e = 0; e < 9; ++e) Jw[e] = (e % 4 == 0) ? 1.0 : 0.0;
  } else {
    const double th = sqrt(th2);
    const double coef = 1.0 / th2 - (1.0 + cos(th)) / (2.0 * th * sin(th));
#pragma unroll
    for (int e = 0; e < 9; ++e) Jw[e] = ((e % 4 == 0) ? 1.0 : 0.0) + 0.5 * W[e] + coef * W2[e];
  }
  double V[9], WV[9], VW[9], WVW[9], WWV[9], VWW[9], WVWW[9], WWVW[9];
  skew3(v, V);
  mm3(W, V, WV);
  mm3(V, W, VW);
  mm3(WV, W, WVW);
  mm3(W2, V, WWV);
  mm3(V, W2, VWW);
  mm3(WVW, W, WVWW);
  mm3(W, WVW, WWVW);
  const double phi = sqrt(th2);
  double a, b, c;
  if (phi > 1e-5) {
    const double s = sin(phi), co = cos(phi);
    const double p2 = phi * phi, p3 = p2 * phi, p4 = p3 * phi, p5 = p4 * phi;
    a = (phi - s) / p3;
    b = (1 - p2 / 2 - co) / p4;
    c = -0.5 * ((1 - p2 / 2 - co) / p4 - 3 * (phi - s - p3 / 6.0) / p5);
  } else {
    a = 1.0 / 6.0;
    b = -1.0 / 24.0;
    c = 1.0 / 120.0;
  }
  double Q[9];
#pragma unroll
  for (int e = 0; e < 9; ++e)
    Q[e] = -0.5 * V[e] + a * (WV[e] + VW[e] - WVW[e]) + b * (WWV[e] + VWW[e] - 3 * WVW[e]) +
           c * (WVWW[e] + WWVW[e]);
  double JQ[9], Q2[9];
  mm3(Jw, Q, JQ);
  mm3(JQ, Jw, Q2);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      J[i * 6 + j] = Jw[i * 3 + j];
      J[i * 6 + j + 3] = 0.0;
      J[(i + 3) * 6 + j] = -Q2[i * 3 + j];
      J[(i + 3) * 6 + j + 3] = Jw[i * 3 + j];
    }
}

// ------------------------------------------------------------------ factors
// Whitened-cost and Jacobians of one factor at poses x.
// kind 0: between(i, j), kind 1: prior(i).
struct FacLin {
  double e[6];
  double Ji[36];
  double Jj[36];
};

// Pose2 graphs are embedded in the plane (theta -> omega_z, x -> v_x,
// y -> v_y).  With chart = 1 their factors keep GTSAM's default Pose2 chart
// (Pose2::ChartAtOrigin, SLOW_BUT_CORRECT_EXPMAP undefined): the in-plane
// error is the (x, y, theta) of the relative pose instead of its Logmap, and
// values move by x * Pose2(dx).  The out-of-plane components (omega_x,
// omega_y, v_z) keep the SE(3) terms, which vanish and decouple for planar
// poses.
constexpr int kPlane[3] = {3, 4, 2};  // tangent index of Pose2 x, y, theta

__device__ __forceinline__ void chart_in_plane(const double* E, double sign, double* e) {
  e[kPlane[0]] = sign * E[9];
  e[kPlane[1]] = sign * E[10];
  e[kPlane[2]] = sign * atan2(E[3], E[0]);
}

__device__ void factor_error(int kind, const double* Ti, const double* Tj, const double* z,
                            double* e, int chart) {
  if (kind == 0) {
    double Tij[12], zi[12], E[12];
    double Tii[12];
    R3::pose_inverse(Ti, Tii);
    R3::pose_compose(Tii, Tj, Tij);
    R3::pose_inverse(z, zi);
    R3::pose_compose(zi, Tij, E);
    R3::logmap(E, e);
    if (chart) chart_in_plane(E, 1.0, e);  // Local(z, hx) = chart(z^-1 hx)
  } else {
    double Tii[12], E[12];
    R3::pose_inverse(Ti, Tii);
    R3::pose_compose(Tii, z, E);
    R3::logmap(E, e);
#pragma unroll
    for (int k = 0; k < 6; ++k) e[k] = -e[k];
    if (chart) chart_in_plane(E, -1.0, e);  // -Local(x, prior) = -chart(x^-1 prior)
  }
}

__device__ void factor_linearize(int kind, const double* Ti, const double* Tj, const double* z,
                                 FacLin& L, int chart) {
  if (kind == 0) {
    double Tii[12], Tij[12], zi[12], E[12];
    R3::pose_inverse(Ti, Tii);
    R3::pose_compose(Tii, Tj, Tij);
    R3::pose_inverse(z, zi);
    R3::pose_compose(zi, Tij, E);
    R3::logmap(E, L.e);
    double Jr[36], Tji[12], Ad[36];
    se3_logmap_derivative64(L.e, Jr);
    R3::pose_inverse(Tij, Tji);
    R3::adjoint(Tji, Ad);
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += Jr[r * 6 + k] * Ad[k * 6 + c];
        L.Ji[r * 6 + c] = -s;
        L.Jj[r * 6 + c] = Jr[r * 6 + c];
      }
    if (chart) {
      // Pose2::between's first Jacobian -Ad(between(Tj, Ti)) (hx = Tij: its
      // rotation (c, s), dt the world-frame translation Tj - Ti, (c2, s2) the
      // rotation of Tj) and the chart's Jacobian H_r = [[R(r), 0], [0, 1]] at
      // r = z^-1 hx: Ji = H_r Hb1, Jj = H_r (BetweenFactor with
      // LieGroup::localCoordinates)
      const double c = Tij[0], sn = Tij[3];
      const double c2 = Tj[0], s2 = Tj[3];
      const double dxw = Tj[9] - Ti[9], dyw = Tj[10] - Ti[10];
      const double Hb[9] = {-c, -sn, -s2 * dxw + c2 * dyw, sn, -c, -c2 * dxw - s2 * dyw, 0.0, 0.0, -1.0};
      const double cr = E[0], sr = E[3];
      const double Hr[9] = {cr, -sr, 0.0, sr, cr, 0.0, 0.0, 0.0, 1.0};
      chart_in_plane(E, 1.0, L.e);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // in-plane rows: only in-plane columns
#pragma unroll
        for (int c6 = 0; c6 < 6; ++c6) {
          L.Ji[kPlane[a] * 6 + c6] = 0.0;
          L.Jj[kPlane[a] * 6 + c6] = 0.0;
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          double v = 0.0;
#pragma unroll
          for (int m = 0; m < 3; ++m) v += Hr[a * 3 + m] * Hb[m * 3 + b];
          L.Ji[kPlane[a] * 6 + kPlane[b]] = v;
          L.Jj[kPlane[a] * 6 + kPlane[b]] = Hr[a * 3 + b];
        }
      }
      // out-of-plane rows: no in-plane columns (zero for planar poses)
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        if (r == kPlane[0] || r == kPlane[1] || r == kPlane[2]) continue;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          L.Ji[r * 6 + kPlane[b]] = 0.0;
          L.Jj[r * 6 + kPlane[b]] = 0.0;
        }
      }
    }
  } else {
    factor_error(1, Ti, Tj, z, L.e, chart);
#pragma unroll
    for (int r = 0; r < 36; ++r) {
      L.Ji[r] = (r % 7 == 0) ? 1.0 : 0.0;
      L.Jj[r] = 0.0;
    }
  }
}

__device__ __forceinline__ double quad6(const double* e, const double* L) {
  double q = 0.0;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += L[r * 6 + c] * e[c];
    q += e[r] * s;
  }
  return q;
}

// 0.5 e^T Lambda e per factor (unweighted) at poses x
__global__ void k_factor_errors(int F, const int* __restrict__ kind, const int* __restrict__ fi,
                                const int* __restrict__ fj, const double* __restrict__ z,
                                const double* __restrict__ info, const double* __restrict__ x,
                                double* __restrict__ err, int chart) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  double e[6];
  const int k = kind[f];
  factor_error(k, x + (size_t)fi[f] * 12, x + (size_t)(k == 0 ? fj[f] : fi[f]) * 12,
               z + (size_t)f * 12, e, chart);
  err[f] = 0.5 * quad6(e, info + (size_t)f * 36);
}

// per-factor Gauss-Newton contributions: blocks, gradients and cost (FP64)
// 64-thread workgroups (the launch shape): the register budget of one
// wavefront per SIMD.  The products are formed in an order that keeps at
// most Ji, Jj and one w Lambda J product live (216 VGPRs), without scratch.
__global__ void __launch_bounds__(64) k_linearize(int F, const int* __restrict__ kind, const int* __restrict__ fi,
                                                  const int* __restrict__ fj, const double* __restrict__ z,
                                                  const double* __restrict__ info, const double* __restrict__ w,
                                                  const double* __restrict__ x, double* __restrict__ Hii,
                                                  double* __restrict__ Hjj, double* __restrict__ Hij,
                                                  double* __restrict__ gi, double* __restrict__ gj,
                                                  double* __restrict__ cost, int chart) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int k = kind[f];
  FacLin L;
  factor_linearize(k, x + (size_t)fi[f] * 12, x + (size_t)(k == 0 ? fj[f] : fi[f]) * 12,
                   z + (size_t)f * 12, L, chart);
  const double wf = w[f];
  const double* I = info + (size_t)f * 36;
  // Le = w Lambda e, cost, gradients
  double Le[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double se = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) se += I[r * 6 + c] * L.e[c];
    Le[r] = wf * se;
  }
  double q = 0.0;
#pragma unroll
  for (int r = 0; r < 6; ++r) q += L.e[r] * Le[r];
  cost[f] = 0.5 * q;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      a += L.Ji[m * 6 + r] * Le[m];
      b += L.Jj[m * 6 + r] * Le[m];
    }
    gi[(size_t)f * 6 + r] = a;
    gj[(size_t)f * 6 + r] = b;
  }
  // LJ = w Lambda Jj: Hjj = Jj^T LJ, Hij = Ji^T LJ
  {
    double LJ[36];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double sj = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) sj += I[r * 6 + m] * L.Jj[m * 6 + c];
        LJ[r * 6 + c] = wf * sj;
      }
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double sjj = 0.0, sij = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          sjj += L.Jj[m * 6 + r] * LJ[m * 6 + c];
          sij += L.Ji[m * 6 + r] * LJ[m * 6 + c];
        }
        Hjj[(size_t)f * 36 + r * 6 + c] = sjj;
        Hij[(size_t)f * 36 + r * 6 + c] = sij;
      }
  }
  // LJ = w Lambda Ji: Hii = Ji^T LJ
  {
    double LJ[36];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double si = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) si += I[r * 6 + m] * L.Ji[m * 6 + c];
        LJ[r * 6 + c] = wf * si;
      }
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double sii = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) sii += L.Ji[m * 6 + r] * LJ[m * 6 + c];
        Hii[(size_t)f * 36 + r * 6 + c] = sii;
      }
  }
}

// Incidence: for pose p, entries inc[off[p]..off[p+1]) = 2 f + side (side 0: f's
// first key, side 1: second key), sorted by f.
__global__ void k_inc_count(int F, const int* __restrict__ kind, const int* __restrict__ fi,
                            const int* __restrict__ fj, int* __restrict__ cnt) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  atomicAdd(&cnt[fi[f]], 1);
  if (kind[f] == 0) atomicAdd(&cnt[fj[f]], 1);
}

__global__ void __launch_bounds__(1024) k_exclusive_scan(int n, const int* __restrict__ in,
                                                         int* __restrict__ out) {
  __shared__ int wsum[16];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int start = 0; start < n; start += 1024) {
    const int k = start + threadIdx.x;
    const int v = k < n ? in[k] : 0;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int pre = base;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
    if (k < n) out[k] = pre + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < 16; ++w) t += wsum[w];
      base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = base;
}

__global__ void k_inc_fill(int F, const int* __restrict__ kind, const int* __restrict__ fi,
                           const int* __restrict__ fj, const int* __restrict__ off,
                           int* __restrict__ cursor, int* __restrict__ inc) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  inc[off[fi[f]] + atomicAdd(&cursor[fi[f]], 1)] = 2 * f;
  if (kind[f] == 0) inc[off[fj[f]] + atomicAdd(&cursor[fj[f]], 1)] = 2 * f + 1;
}

// insertion sort of each pose's (short) incidence list, for deterministic sums
__global__ void k_inc_sort(int N, const int* __restrict__ off, int* __restrict__ inc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int a = off[p], b = off[p + 1];
  for (int k = a + 1; k < b; ++k) {
    const int v = inc[k];
    int m = k - 1;
    while (m >= a && inc[m] > v) {
      inc[m + 1] = inc[m];
      --m;
    }
    inc[m + 1] = v;
  }
}

// the other pose of each incidence, -1 for a factor with no off-diagonal
// block (kind != 0), so the row SpMV reads it together with the incidence
__global__ void k_inc_other(int N, const int* __restrict__ off, const int* __restrict__ inc,
                            const int* __restrict__ kind, const int* __restrict__ fi,
                            const int* __restrict__ fj, int* __restrict__ inc_q) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  for (int k = off[p]; k < off[p + 1]; ++k) {
    const int t = inc[k], f = t >> 1;
    inc_q[k] = kind[f] != 0 ? -1 : ((t & 1) ? fi[f] : fj[f]);
  }
}

// Diagonal blocks, gradient rhs b = -g and the Hessian diagonal, one thread
// per pose, summing the incident factors in factor order (deterministic).
__global__ void k_assemble(int N, const int* __restrict__ off, const int* __restrict__ inc,
                           const double* __restrict__ Hii, const double* __restrict__ Hjj,
                           const double* __restrict__ gi, const double* __restrict__ gj,
                           double* __restrict__ D, double* __restrict__ bvec,
                           double* __restrict__ hdiag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double acc[36], g[6];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
#pragma unroll
  for (int e = 0; e < 6; ++e) g[e] = 0.0;
  for (int k = off[p]; k < off[p + 1]; ++k) {
    const int t = inc[k], f = t >> 1;
    const double* B = (t & 1) ? Hjj + (size_t)f * 36 : Hii + (size_t)f * 36;
    const double* gg = (t & 1) ? gj + (size_t)f * 6 : gi + (size_t)f * 6;
#pragma unroll
    for (int e = 0; e < 36; ++e) acc[e] += B[e];
#pragma unroll
    for (int e = 0; e < 6; ++e) g[e] += gg[e];
  }
#pragma unroll
  for (int e = 0; e < 36; ++e) D[(size_t)p * 36 + e] = acc[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    bvec[(size_t)p * 6 + e] = -g[e];
    hdiag[(size_t)p * 6 + e] = acc[e * 7];
  }
}

// ---------------------------------------------------------------------------
// Segment block-tridiagonal preconditioner.  The odometry chain k -> k+1 makes
// the Gauss-Newton matrix block-tridiagonal up to the loop-closure blocks; the
// preconditioner solves that chain exactly inside segments of S consecutive
// poses (block LDL^T, one thread per segment), dropping the loop-closure
// blocks and the coupling across segment borders.  S = 1 is block-Jacobi.
// ---------------------------------------------------------------------------

// DPP row rotation of an FP64 value (both halves), all lanes of the row valid
template <int N>
__device__ __forceinline__ double row_ror_f64(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, 0x120 + N, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(0, lo, 0x120 + N, 0xF, 0xF, false));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a converged wavefront, in every lane: rotations
// within each row of 16 lanes by DPP (VALU, a few cycles each), then the four
// row sums by v_readlane.  A six-step shuffle reduction went through the LDS
// crossbar (ds_bpermute) at every step, on the critical path of each PCG
// launch.
__device__ __forceinline__ double wave_sum_pre(double v) {
  v += row_ror_f64<8>(v);
  v += row_ror_f64<4>(v);
  v += row_ror_f64<2>(v);
  v += row_ror_f64<1>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// block-wide FP64 sum (all threads call it; the result is valid in thread 0)
__device__ __forceinline__ double block_sum(double v, double* lds16) {
  v = wave_sum_pre(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds16[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < 64) {
    t = threadIdx.x < (blockDim.x >> 6) ? lds16[threadIdx.x] : 0.0;
    t = wave_sum_pre(t);
  }
  return t;  // valid in thread 0
}

// chain[k] = 2 f + flip for the first odometry factor joining k and k+1
// (flip = 1 when the factor runs k+1 -> k), -1 when none
__global__ void k_chain_init(int n, int* __restrict__ chain) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) chain[k] = INT_MAX;
}

__global__ void k_chain_find(int F, const int* __restrict__ kind, const int* __restrict__ fi,
                             const int* __restrict__ fj, int* __restrict__ chain) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F || kind[f] != 0) return;
  const int a = fi[f], b = fj[f];
  if (b == a + 1) atomicMin(&chain[a], 2 * f);
  else if (a == b + 1) atomicMin(&chain[b], 2 * f + 1);
}

__device__ __forceinline__ void damped_block(const double* __restrict__ D, const double* __restrict__ hdiag,
                                             int p, double lambda, int diag_damping, double* A) {
#pragma unroll
  for (int e = 0; e < 36; ++e) A[e] = D[(size_t)p * 36 + e];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double d = diag_damping ? fmin(fmax(hdiag[(size_t)p * 6 + i], 1e-6), 1e32) : 1.0;
    A[i * 7] += lambda * d;
  }
}

// in-place inverse of an SPD 6x6 (Cholesky; tiny pivots floored)
__device__ void spd_inverse6(const double* A, double* Ainv) {
  double L[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) L[e] = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double x = A[k * 7];
#pragma unroll
    for (int j = 0; j < k; ++j) x -= L[k * 6 + j] * L[k * 6 + j];
    x = sqrt(fmax(x, 1e-300));
    L[k * 7] = x;
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      double v = A[i * 6 + k];
#pragma unroll
      for (int j = 0; j < k; ++j) v -= L[i * 6 + j] * L[k * 6 + j];
      L[i * 6 + k] = v / x;
    }
  }
  double Li[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) Li[e] = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      if (i < c) continue;
      double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j >= c && j < i) v -= L[i * 6 + j] * Li[j * 6 + c];
      Li[i * 6 + c] = v / L[i * 7];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) s += Li[k * 6 + i] * Li[k * 6 + j];
      Ainv[i * 6 + j] = s;
    }
}

// Factor: Dp_k = A_kk - U_{k-1}^T Dp_{k-1}^-1 U_{k-1}; stores Dp_k^-1 (Minv)
// and U_k = H_{k,k+1} (Useg, zero at a segment end or without odometry).
__global__ void __launch_bounds__(64) k_seg_factor(int N, int S, const int* __restrict__ chain, const double* __restrict__ D,
                             const double* __restrict__ hdiag, const double* __restrict__ Hij,
                             double lambda, int diag_damping, float* __restrict__ Minv,
                             float* __restrict__ Useg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int k0 = s * S;
  if (k0 >= N) return;
  const int k1 = min(k0 + S, N);
  double Dinv_prev[36], U_prev[36];
  bool have_prev = false;
  for (int k = k0; k < k1; ++k) {
    double A[36];
    damped_block(D, hdiag, k, lambda, diag_damping, A);
    if (have_prev) {
      // A -= U^T (Dinv U)
      double T[36];
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          double v = 0.0;
#pragma unroll
          for (int m = 0; m < 6; ++m) v += Dinv_prev[i * 6 + m] * U_prev[m * 6 + j];
          T[i * 6 + j] = v;
        }
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          double v = 0.0;
#pragma unroll
          for (int m = 0; m < 6; ++m) v += U_prev[m * 6 + i] * T[m * 6 + j];
          A[i * 6 + j] -= v;
        }
    }
    double Ainv[36];
    spd_inverse6(A, Ainv);
#pragma unroll
    for (int e = 0; e < 36; ++e) {
      Minv[(size_t)k * 36 + e] = (float)Ainv[e];
      Dinv_prev[e] = Ainv[e];
    }
    // U_k = H_{k,k+1}
    const int c = (k + 1 < k1) ? chain[k] : INT_MAX;
    if (c != INT_MAX) {
      const int f = c >> 1;
      const bool flip = c & 1;
      const double* B = Hij + (size_t)f * 36;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          const double v = flip ? B[j * 6 + i] : B[i * 6 + j];
          U_prev[i * 6 + j] = v;
          Useg[(size_t)k * 36 + i * 6 + j] = (float)v;
        }
      have_prev = true;
    } else {
#pragma unroll
      for (int e = 0; e < 36; ++e) Useg[(size_t)k * 36 + e] = 0.f;
      have_prev = false;
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse level of the two-level preconditioner.  Every coarse segment of Sc
// consecutive poses moves as one rigid body: delta_k = Ad(T_k^-1 T_a) c_s with
// a the segment's first pose, so intra-segment between factors see no change
// under the coarse modes (the low-curvature modes block methods converge on
// slowly).  With P that prolongation, Ac = P^T (H + lambda damp) P
// (nc = 6 nseg, dense FP64) is assembled per linear solve, factored (dense
// Cholesky) and inverted explicitly; each PCG iteration applies the additive
// preconditioner  z = B_fine^-1 r + P Ac^-1 P^T r.
// ---------------------------------------------------------------------------
__global__ void k_coarse_basis(int N, int Sc, const double* __restrict__ x, float* __restrict__ Pb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int a = (k / Sc) * Sc;
  double ki[12], T[12], Ad[36];
  R3::pose_inverse(x + (size_t)k * 12, ki);
  R3::pose_compose(ki, x + (size_t)a * 12, T);
  R3::adjoint(T, Ad);
#pragma unroll
  for (int e = 0; e < 36; ++e) Pb[(size_t)k * 36 + e] = (float)Ad[e];
}

// out = P1^T M P2 for 6x6 blocks (P in FP32, accumulation in FP64)
__device__ __forceinline__ void ptmp(const float* P1, const double* M, const float* P2, double* out) {
  double T[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) s += M[i * 6 + m] * (double)P2[m * 6 + j];
      T[i * 6 + j] = s;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) s += (double)P1[m * 6 + i] * T[m * 6 + j];
      out[i * 6 + j] = s;
    }
}

// Ac(s,s) += P_k^T (D_k + lambda damp_k) P_k
__global__ void __launch_bounds__(128) k_coarse_diag(int N, int Sc, int nc, const float* __restrict__ Pb,
                              const double* __restrict__ D, const double* __restrict__ hdiag,
                              double lambda, int diag_damping, double* __restrict__ Ac) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double B[36], M[36];
  damped_block(D, hdiag, k, lambda, diag_damping, B);
  const float* P = Pb + (size_t)k * 36;
  ptmp(P, B, P, M);
  const int s = k / Sc;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) atomicAdd(&Ac[(size_t)(6 * s + i) * nc + 6 * s + j], M[i * 6 + j]);
}

// Ac(s,t) += P_i^T H_ij P_j and Ac(t,s) += its transpose, per between factor
__global__ void __launch_bounds__(128) k_coarse_offdiag(int F, int Sc, int nc, const int* __restrict__ kind,
                                 const int* __restrict__ fi, const int* __restrict__ fj,
                                 const double* __restrict__ Hij, const float* __restrict__ Pb,
                                 double* __restrict__ Ac) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F || kind[f] != 0) return;
  const int i = fi[f], j = fj[f];
  double H[36], M[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) H[e] = Hij[(size_t)f * 36 + e];
  ptmp(Pb + (size_t)i * 36, H, Pb + (size_t)j * 36, M);
  const int s = i / Sc, t = j / Sc;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      atomicAdd(&Ac[(size_t)(6 * s + a) * nc + 6 * t + b], M[a * 6 + b]);
      atomicAdd(&Ac[(size_t)(6 * t + b) * nc + 6 * s + a], M[a * 6 + b]);
    }
}

constexpr int kCholNb = 16;

// Blocked right-looking Cholesky of the coarse matrix A (n x n, row-major,
// lower triangle used), panels of 16 columns.  The factor L is written only
// transposed, into AT (row j of AT = column j of L), which is what the
// inverse reads; A keeps the trailing matrix.  Pivots are floored at a tiny
// positive value so a gauge-free (semi-definite) problem keeps a finite
// preconditioner.
//
// Panel step 1: the 16 x 16 diagonal block at (k0, k0), factored by one
// wavefront (four entries per lane, row e >> 4, column e & 15), then L_kk
// into AT.  The rows below the block are independent triangular solves and
// are done by the trailing update's workgroups (step 2); with the whole
// panel (up to 768 rows) on one workgroup, a panel took 13-28 us.
__global__ void __launch_bounds__(64) k_chol_panel(int n, int k0, const double* __restrict__ A,
                                                   double* __restrict__ AT) {
  static_assert(kCholNb == 16, "entry e = 16 row + column");
  __shared__ double blk[kCholNb * kCholNb];
  const int nb = min(kCholNb, n - k0);
  const int lane = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = lane + 64 * u, r = e >> 4, j = e & 15;
    blk[e] = (r < nb && j < nb) ? A[(size_t)(k0 + r) * n + k0 + j] : 0.0;
  }
  __syncthreads();
  // unscaled right-looking elimination: step c updates columns > c
  for (int c = 0; c < nb; ++c) {
    const double inv = 1.0 / fmax(blk[c * kCholNb + c], 1e-300);
    double upd[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = lane + 64 * u, r = e >> 4, j = e & 15;
      upd[u] = blk[e] - blk[r * kCholNb + c] * blk[j * kCholNb + c] * inv;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = lane + 64 * u, r = e >> 4, j = e & 15;
      if (r < nb && j > c && j <= r) blk[e] = upd[u];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = lane + 64 * u, r = e >> 4, j = e & 15;
    if (r < nb && j <= r) {
      const double piv = sqrt(fmax(blk[j * kCholNb + j], 1e-300));
      AT[(size_t)(k0 + j) * n + k0 + r] = r == j ? piv : blk[e] / piv;
    }
  }
}

// Panel step 2, over the rows below the diagonal block (r >= k0 + 16): one
// 32 x 32 lower tile (ti, tj) of the trailing matrix per workgroup.  The
// workgroup solves its two row tiles of the panel, l_r L_kk^T = a_r (forward
// substitution, one thread per row), writes them into AT when ti == tj (each
// row tile has one diagonal tile), and updates A_ij -= sum_c l_ic l_jc.
__global__ void __launch_bounds__(256) k_chol_update(int n, int k0, double* __restrict__ A,
                                                     double* __restrict__ AT) {
  const int nb = min(kCholNb, n - k0);
  const int r0 = k0 + nb;
  // lower-triangle tile index -> (ti, tj), tj <= ti
  const int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  __shared__ double Lkk[kCholNb][kCholNb + 1];
  __shared__ double Lr[64][kCholNb + 1];  // rows 0-31: tile ti, 32-63: tile tj
  for (int e = threadIdx.x; e < kCholNb * kCholNb; e += blockDim.x) {
    const int c = e / kCholNb, r = e % kCholNb;  // AT row c = column c of L
    Lkk[r][c] = (r < nb && c < nb && c <= r) ? AT[(size_t)(k0 + c) * n + k0 + r] : 0.0;
  }
  for (int e = threadIdx.x; e < 64 * kCholNb; e += blockDim.x) {
    const int rr = e / kCholNb, cc = e % kCholNb;
    const int g = r0 + (rr < 32 ? ti * 32 + rr : tj * 32 + rr - 32);
    Lr[rr][cc] = (g < n && cc < nb) ? A[(size_t)g * n + k0 + cc] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int rr = threadIdx.x;
    for (int c = 0; c < nb; ++c) {
      double v = Lr[rr][c];
      for (int k = 0; k < c; ++k) v -= Lr[rr][k] * Lkk[c][k];
      Lr[rr][c] = v / Lkk[c][c];
    }
  }
  __syncthreads();
  if (ti == tj)
    for (int e = threadIdx.x; e < 32 * kCholNb; e += blockDim.x) {
      const int cc = e >> 5, rr = e & 31;  // AT row k0 + cc, rows coalesced
      const int g = r0 + ti * 32 + rr;
      if (g < n && cc < nb) AT[(size_t)(k0 + cc) * n + g] = Lr[rr][cc];
    }
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int rr = e >> 5, cc = e & 31;
    const int gi = r0 + ti * 32 + rr, gj = r0 + tj * 32 + cc;
    if (gi >= n || gj > gi) continue;
    double sum = 0.0;
#pragma unroll 8
    for (int m = 0; m < kCholNb; ++m) sum += Lr[rr][m] * Lr[32 + cc][m];
    A[(size_t)gi * n + gj] -= sum;
  }
}

// W = L^-1 (lower): one wave per column c, column-oriented substitution:
// x_i = (delta_ic - s_i) / L_ii, then s_k += L_ki x_i for every k > i.  Lane
// l keeps the partial sums s_k of the rows k = l (mod 64) in registers, x_i
// reaches every lane by readlane from the owner of s_i, and column i of L is
// row i of LT (coalesced), prefetched PF rows ahead.  The row-oriented form
// (w_i from a dot product over j < i) needed a wave reduction per row:
// 691 us at 630 unknowns, 5k poses.
template <int PL>  // rows per lane: n <= 64 PL
__global__ void __launch_bounds__(256) k_tri_inv(int n, const double* __restrict__ LT, double* __restrict__ W) {
  constexpr int PF = 4;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wv;
  if (c >= n) return;
  double sk[PL];
  double buf[PF][PL];
#pragma unroll
  for (int q = 0; q < PL; ++q) sk[q] = 0.0;
#pragma unroll
  for (int u = 0; u < PF; ++u)
#pragma unroll
    for (int q = 0; q < PL; ++q) {
      const int k = lane + 64 * q, row = min(c + u, n - 1);
      buf[u][q] = LT[(size_t)row * n + min(k, n - 1)];
    }
  for (int i0 = c; i0 < n; i0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int i = i0 + u;
      if (i < n) {
        const int qi = i >> 6, li = i & 63;
        double si = 0.0, di = 0.0;
#pragma unroll
        for (int q = 0; q < PL; ++q)
          if (q == qi) {
            si = sk[q];
            di = buf[u][q];
          }
        si = readlane_f64(si, li);
        di = readlane_f64(di, li);  // L_ii
        const double x = (((i == c) ? 1.0 : 0.0) - si) / di;
        if (lane == 0) W[(size_t)i * n + c] = x;
#pragma unroll
        for (int q = 0; q < PL; ++q) {
          const int k = lane + 64 * q;
          const double upd = sk[q] + buf[u][q] * x;
          sk[q] = k > i ? upd : sk[q];
        }
        const int row = min(i + PF, n - 1);
#pragma unroll
        for (int q = 0; q < PL; ++q) buf[u][q] = LT[(size_t)row * n + min(lane + 64 * q, n - 1)];
      }
    }
  }
  for (int i = lane; i < c; i += 64) W[(size_t)i * n + c] = 0.0;
}

// Ainv = Linv^T Linv (symmetric), Linv lower triangular (row-major, zeros
// above the diagonal).  One 32x32 tile (I, J), I <= J, per workgroup, 2x2
// entries per thread, k from the tile's first nonzero row in chunks of 32
// staged through LDS; the tile and its mirror are written (a product and its
// transpose are the same sums in the same order, so Ainv is exactly
// symmetric).  The former one-thread-per-entry form read every operand from
// global memory for each entry: 442 us at 630 unknowns, 5k poses.
__global__ void __launch_bounds__(256) k_gram_tiles(int n, int T, const double* __restrict__ Li,
                                                   double* __restrict__ Ainv) {
  __shared__ double sa[32][33], sb[32][33];
  int b = blockIdx.x, I = 0;
  while (b >= T - I) {  // upper-triangle tiles numbered row by row
    b -= T - I;
    ++I;
  }
  const int J = I + b;
  const int i0 = I * 32, j0 = J * 32;
  const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
  const int lr = tid >> 3, lc = (tid & 7) * 4;  // loader: row lr, columns lc .. lc + 3
  double acc00 = 0.0, acc01 = 0.0, acc10 = 0.0, acc11 = 0.0;
  for (int k0 = j0; k0 < n; k0 += 32) {
    const int k = k0 + lr;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sa[lr][lc + e] = (k < n && i0 + lc + e < n) ? Li[(size_t)k * n + i0 + lc + e] : 0.0;
      sb[lr][lc + e] = (k < n && j0 + lc + e < n) ? Li[(size_t)k * n + j0 + lc + e] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) {
      const double a0 = sa[kk][2 * ti], a1 = sa[kk][2 * ti + 1];
      const double b0 = sb[kk][2 * tj], b1 = sb[kk][2 * tj + 1];
      acc00 += a0 * b0;
      acc01 += a0 * b1;
      acc10 += a1 * b0;
      acc11 += a1 * b1;
    }
    __syncthreads();
  }
  const int ia = i0 + 2 * ti, ja = j0 + 2 * tj;
  const double v[2][2] = {{acc00, acc01}, {acc10, acc11}};
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int i = ia + x, j = ja + y;
      if (i < n && j < n) {
        Ainv[(size_t)i * n + j] = v[x][y];
        Ainv[(size_t)j * n + i] = v[x][y];
      }
    }
}

// y = (H + lambda damp) x : one thread per pose row (deterministic order)
__global__ void k_spmv(int N, const int* __restrict__ off, const int* __restrict__ inc,
                       const int* __restrict__ fi, const int* __restrict__ fj,
                       const int* __restrict__ kind, const double* __restrict__ D,
                       const double* __restrict__ Hij, const double* __restrict__ hdiag,
                       const double* __restrict__ lam_ptr, double lambda_imm,
                       int diag_damping, double min_diag, double max_diag,
                       const double* __restrict__ x, double* __restrict__ y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double lambda = lam_ptr ? *lam_ptr : lambda_imm;
  double xp[6], acc[6];
#pragma unroll
  for (int e = 0; e < 6; ++e) xp[e] = x[(size_t)p * 6 + e];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += D[(size_t)p * 36 + r * 6 + c] * xp[c];
    const double d = diag_damping ? fmin(fmax(hdiag[(size_t)p * 6 + r], min_diag), max_diag) : 1.0;
    acc[r] = s + lambda * d * xp[r];
  }
  for (int k = off[p]; k < off[p + 1]; ++k) {
    const int t = inc[k], f = t >> 1;
    if (kind[f] != 0) continue;
    const double* B = Hij + (size_t)f * 36;
    if ((t & 1) == 0) {  // row i: + Hij x_j
      const double* xq = x + (size_t)fj[f] * 6;
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s += B[r * 6 + c] * xq[c];
        acc[r] += s;
      }
    } else {  // row j: + Hij^T x_i
      const double* xq = x + (size_t)fi[f] * 6;
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s += B[c * 6 + r] * xq[c];
        acc[r] += s;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) y[(size_t)p * 6 + r] = acc[r];
}

// FP64 dot, block-reduced then atomically accumulated
__global__ void k_dot(long long n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    s += a[k] * b[k];
  s = block_sum(s, red);
  if (threadIdx.x == 0) atomicAdd(out, s);
}

// ---------------------------------------------------------------------------
// Fused PCG iteration (4 launches, 3 without the coarse level).  Scalars live
// in parity slots so that no launch has to reset what a concurrent block of
// the same launch still reads: for iteration g with a = g & 1,
//   R[a] = r_g . z_g, R[a^1] accumulates r_{g+1} . z_{g+1}; PQ[a] = p_g . A p_g;
//   cc[a] accumulates the coarse correction; the kHistSlots slots of
//   hist[cnt * kHistRow] sum to |r_{g+1}|^2.
// Each launch clears only slots that no block of the same launch touches.
// ---------------------------------------------------------------------------
constexpr int kScCnt = 32, kScLam = 30;  // iteration counter, LM lambda
// The PCG scalars r.z (R) and p.q (PQ) of each parity a are summed over
// workgroups in kRedSlots slots 128 bytes apart (sc + kScSlots): a workgroup
// adds its partial to slot (block id mod kRedSlots), and a reader sums the
// slots.  Hundreds of workgroups adding to one address serialised in L2 and
// cost the SpMV and coarse kernels about 5 and 4.5 us per PCG iteration at
// 50,000 poses.
constexpr int kScSlots = 64, kRedSlots = 32, kRedStride = 16;
constexpr int kScTotal = kScSlots + 4 * kRedSlots * kRedStride;  // doubles in sc
constexpr int kRedR = 0, kRedPQ = 1;
// the residual history |r|^2 of each iteration in kHistSlots slots of one
// 128-byte line each (hist + it * kHistRow), summed on the host
constexpr int kHistSlots = 8, kHistRow = kHistSlots * kRedStride;

__device__ __forceinline__ double* red_slot(double* sc, int which, int a, int slot) {
  return sc + kScSlots + ((which * 2 + a) * kRedSlots + slot) * kRedStride;
}

// add this workgroup's partial (one thread per workgroup calls it)
__device__ __forceinline__ void red_add(double* sc, int which, int a, double v) {
  const int blk = blockIdx.x + blockIdx.y * gridDim.x;
  atomicAdd(red_slot(sc, which, a, blk % kRedSlots), v);
}

// the sum over the slots, in every lane of a converged wavefront
__device__ __forceinline__ double red_sum(double* sc, int which, int a) {
  const int lane = threadIdx.x & 63;
  const double v = lane < kRedSlots ? *red_slot(sc, which, a, lane) : 0.0;
  return wave_sum_pre(v);
}

// clear the slots of (which, a): lanes < kRedSlots of one wavefront
__device__ __forceinline__ void red_clear(double* sc, int which, int a) {
  const int lane = threadIdx.x & 63;
  if (lane < kRedSlots) *red_slot(sc, which, a, lane) = 0.0;
}
constexpr int kSpmvPoseMin = 16384;  // poses from which the SpMV runs one thread per pose
constexpr int kSpmvBatch = 4;        // incidences per pose whose loads the row SpMV issues together
constexpr bool kSpmvChain = true;    // PCG SpMV on the entry-major chain layout (k_pcg_spmv_chain)

// ---------------------------------------------------------------------------
// Chain-structured SpMV.  The odometry chain makes the Gauss-Newton matrix
// block-tridiagonal along the pose order plus a few loop-closure blocks:
//   y_p = (D_p + lambda damp_p) x_p + C_p x_{p+1} + C_{p-1}^T x_{p-1}
//         + sum over p's other between factors (Hij or Hij^T) x_q,
// with C_p = H_{p,p+1} the block of the chain link p -> p+1 (chain[p]).  The
// diagonal and link blocks are kept entry-major (entry e of pose p at
// e * N + p): a wavefront of 64 consecutive poses reads each entry as 512
// contiguous bytes, where the per-factor layout made every lane fetch its
// own 288-byte block through a chain of dependent index loads.
// ---------------------------------------------------------------------------

// extras: per pose, the between factors that are not its chain links (loop
// closures, repeated odometry), in incidence (factor) order
__device__ __forceinline__ bool is_link(const int* __restrict__ chain, int N, int p, int f) {
  const int c0 = chain[p];
  if (c0 != INT_MAX && (c0 >> 1) == f) return true;
  if (p > 0) {
    const int c1 = chain[p - 1];
    if (c1 != INT_MAX && (c1 >> 1) == f) return true;
  }
  (void)N;
  return false;
}

__global__ void k_ext_count(int N, const int* __restrict__ off, const int* __restrict__ inc,
                            const int* __restrict__ kind, const int* __restrict__ chain,
                            int* __restrict__ cnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  int c = 0;
  for (int k = off[p]; k < off[p + 1]; ++k) {
    const int f = inc[k] >> 1;
    if (kind[f] == 0 && !is_link(chain, N, p, f)) ++c;
  }
  cnt[p] = c;
}

__global__ void k_ext_fill(int N, const int* __restrict__ off, const int* __restrict__ inc,
                           const int* __restrict__ incq, const int* __restrict__ kind,
                           const int* __restrict__ chain, const int* __restrict__ eoff,
                           int* __restrict__ ext, int* __restrict__ extq) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  int o = eoff[p];
  for (int k = off[p]; k < off[p + 1]; ++k) {
    const int t = inc[k], f = t >> 1;
    if (kind[f] == 0 && !is_link(chain, N, p, f)) {
      ext[o] = t;
      extq[o] = incq[k];
      ++o;
    }
  }
}

// entry-major diagonal blocks Ds and link blocks Cs (zero where pose p has no
// link to p + 1); one thread per (entry, pose)
__global__ void k_chain_blocks(int N, const int* __restrict__ chain, const double* __restrict__ D,
                               const double* __restrict__ Hij, double* __restrict__ Ds,
                               double* __restrict__ Cs) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36ll * N) return;
  const int e = (int)(t / N), p = (int)(t - (long long)e * N);
  Ds[t] = D[(size_t)p * 36 + e];
  double c = 0.0;
  const int l = (p + 1 < N) ? chain[p] : INT_MAX;
  if (l != INT_MAX) {
    const int f = l >> 1, r = e / 6, col = e % 6;
    // flip: the factor runs p+1 -> p, its Hij is H_{p+1,p} = C_p^T
    c = (l & 1) ? Hij[(size_t)f * 36 + col * 6 + r] : Hij[(size_t)f * 36 + e];
  }
  Cs[t] = c;
}

// K1 on the chain layout: a workgroup of six wavefronts covers 64 poses,
// wavefront r computing output row r of each (its lanes 64 consecutive
// poses).  Row sums in a fixed order: diagonal, damping, link to p+1, link
// from p-1, extras in factor order.
__global__ void __launch_bounds__(384) k_pcg_spmv_chain(int N, int a, const int* __restrict__ chain,
                                                  const double* __restrict__ Ds, const double* __restrict__ Cs,
                                                  const int* __restrict__ eoff, const int* __restrict__ ext,
                                                  const int* __restrict__ extq, const double* __restrict__ Hij,
                                                  const double* __restrict__ hdiag, int diag_damping,
                                                  double* __restrict__ sc, const double* __restrict__ x,
                                                  double* __restrict__ y) {
  __shared__ double red[16];
  const int r = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + (threadIdx.x & 63);
  const double lambda = sc[kScLam];
  double pq = 0.0;
  if (p < N) {
    const double* xp = x + (size_t)p * 6;
    const double* xn = x + (size_t)(p + 1 < N ? p + 1 : p) * 6;  // Cs is zero at p = N - 1
    const double* xb = x + (size_t)(p > 0 ? p - 1 : p) * 6;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += Ds[(size_t)(r * 6 + c) * N + p] * xp[c];
    const double d = diag_damping ? fmin(fmax(hdiag[(size_t)p * 6 + r], 1e-6), 1e32) : 1.0;
    double acc = s + lambda * d * xp[r];
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) v += Cs[(size_t)(r * 6 + c) * N + p] * xn[c];
    acc += v;
    if (p > 0) {
      v = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) v += Cs[(size_t)(c * 6 + r) * N + p - 1] * xb[c];
      acc += v;
    }
    for (int k = eoff[p]; k < eoff[p + 1]; ++k) {
      const int t = ext[k], q = extq[k];
      const double* B = Hij + (size_t)(t >> 1) * 36;
      const double* xq = x + (size_t)q * 6;
      v = 0.0;
      if (t & 1) {
#pragma unroll
        for (int c = 0; c < 6; ++c) v += B[c * 6 + r] * xq[c];
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) v += B[r * 6 + c] * xq[c];
      }
      acc += v;
    }
    y[(size_t)p * 6 + r] = acc;
    pq = xp[r] * acc;
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) {
    red_add(sc, kRedPQ, a, pq);
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    red_clear(sc, kRedR, a ^ 1);
    red_clear(sc, kRedPQ, a ^ 1);
  }
  (void)chain;
}

// K1: q = (H + lambda damp) p ; PQ[a] += p . q ; clear R[a^1], PQ[a^1].
// Eight lanes per pose, lane r < 6 owning output row r: block rows (or
// columns, for the transposed half) are read as 288 contiguous bytes per
// group, and each row sums its products in the same order as a per-pose loop.
__global__ void __launch_bounds__(256) k_pcg_spmv_rows(int N, int a, const int* __restrict__ off,
                                                 const int* __restrict__ inc, const int* __restrict__ incq,
                                                 const double* __restrict__ D, const double* __restrict__ Hij,
                                                 const double* __restrict__ hdiag, int diag_damping,
                                                 double* __restrict__ sc, const double* __restrict__ x,
                                                 double* __restrict__ y) {
  __shared__ double red[16];
  const int lane = threadIdx.x & 63;
  const int r = lane & 7, base = lane & ~7;
  const int rr = r < 6 ? r : 0;  // idle rows 6, 7 shadow row 0 (never stored)
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const double lambda = sc[kScLam];
  double pq = 0.0;
  if (p < N) {
    const double xr = x[(size_t)p * 6 + rr];
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += D[(size_t)p * 36 + rr * 6 + c] * __shfl(xr, base + c);
    const double d = diag_damping ? fmin(fmax(hdiag[(size_t)p * 6 + rr], 1e-6), 1e32) : 1.0;
    double acc = s + lambda * d * xr;
    // incident factors in order; the first kSpmvBatch with every load of the
    // batch independent of the others (clamped indices, results selected): a
    // pose's usual 2-4 factors cost three dependent round trips in all, where
    // one at a time they cost about four each
    const int k0 = off[p], k1 = off[p + 1];
    if (k1 > k0) {
      int tt[kSpmvBatch], qq[kSpmvBatch];
#pragma unroll
      for (int u = 0; u < kSpmvBatch; ++u) {
        const int k = min(k0 + u, k1 - 1);
        tt[u] = inc[k];
        qq[u] = incq[k];
      }
#pragma unroll
      for (int u = 0; u < kSpmvBatch; ++u) {
        const bool use = k0 + u < k1 && qq[u] >= 0;
        const bool tr = (tt[u] & 1) != 0;
        const double* B = Hij + (size_t)(tt[u] >> 1) * 36;
        const double xq = x[(size_t)(qq[u] >= 0 ? qq[u] : p) * 6 + rr];
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) v += B[tr ? c * 6 + rr : rr * 6 + c] * __shfl(xq, base + c);
        acc = use ? acc + v : acc;
      }
      for (int k = k0 + kSpmvBatch; k < k1; ++k) {
        const int t = inc[k], q = incq[k];
        if (q < 0) continue;
        const double* B = Hij + (size_t)(t >> 1) * 36;
        const bool tr = (t & 1) != 0;
        const double xq = x[(size_t)q * 6 + rr];
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) v += B[tr ? c * 6 + rr : rr * 6 + c] * __shfl(xq, base + c);
        acc += v;
      }
    }
    if (r < 6) {
      y[(size_t)p * 6 + r] = acc;
      pq = xr * acc;
    }
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) {
    red_add(sc, kRedPQ, a, pq);
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    red_clear(sc, kRedR, a ^ 1);
    red_clear(sc, kRedPQ, a ^ 1);
  }
}

// K1, one thread per pose: for large graphs, where the pose count alone fills
// the chip and each thread streams its rows with many loads in flight.
__global__ void __launch_bounds__(128) k_pcg_spmv_pose(int N, int a, const int* __restrict__ off,
                                                 const int* __restrict__ inc, const int* __restrict__ fi,
                                                 const int* __restrict__ fj, const int* __restrict__ kind,
                                                 const double* __restrict__ D, const double* __restrict__ Hij,
                                                 const double* __restrict__ hdiag, int diag_damping,
                                                 double* __restrict__ sc, const double* __restrict__ x,
                                                 double* __restrict__ y) {
  __shared__ double red[16];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const double lambda = sc[kScLam];
  double pq = 0.0;
  if (p < N) {
    double xp[6], acc[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) xp[e] = x[(size_t)p * 6 + e];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s += D[(size_t)p * 36 + r * 6 + c] * xp[c];
      const double d = diag_damping ? fmin(fmax(hdiag[(size_t)p * 6 + r], 1e-6), 1e32) : 1.0;
      acc[r] = s + lambda * d * xp[r];
    }
    for (int k = off[p]; k < off[p + 1]; ++k) {
      const int t = inc[k], f = t >> 1;
      if (kind[f] != 0) continue;
      const double* B = Hij + (size_t)f * 36;
      if ((t & 1) == 0) {
        const double* xq = x + (size_t)fj[f] * 6;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s += B[r * 6 + c] * xq[c];
          acc[r] += s;
        }
      } else {
        const double* xq = x + (size_t)fi[f] * 6;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) s += B[c * 6 + r] * xq[c];
          acc[r] += s;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      y[(size_t)p * 6 + r] = acc[r];
      pq += xp[r] * acc[r];
    }
  }
  pq = block_sum(pq, red);
  if (threadIdx.x == 0) {
    red_add(sc, kRedPQ, a, pq);
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    red_clear(sc, kRedR, a ^ 1);
    red_clear(sc, kRedPQ, a ^ 1);
  }
}

// End of K2: the coarse restriction rc += P^T r (rows summed per coarse
// segment), |r|^2 into the residual history and r . z into R[a^1].
__device__ __forceinline__ void pcg_update_tail(int lane, int i, bool row, int cs_mine, double racc, double rr,
                                                double rz, int first, int a, double* __restrict__ sc,
                                                bool coarse, double* __restrict__ rc,
                                                double* __restrict__ hist, double* red) {
  if (coarse) {
    // a wavefront whose segments share one coarse segment sums each row over
    // its groups and adds once; otherwise every row adds its own
    const int cs0 = __shfl(cs_mine, 0), cs1 = __shfl(cs_mine, 56);
    const bool uniform = cs0 == cs1 && cs0 >= 0;
    if (uniform) {
      double v = row ? racc : 0.0;
      v += __shfl_xor(v, 8);
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      if (lane < 6) atomicAdd(&rc[6 * cs0 + lane], v);
    } else if (cs_mine >= 0) {
      atomicAdd(&rc[6 * cs_mine + i], racc);
    }
  }
  rr = block_sum(rr, red);
  rz = block_sum(rz, red);
  if (threadIdx.x == 0) {
    if (!first) {
      const int blk = blockIdx.x + blockIdx.y * gridDim.x;
      atomicAdd(&hist[(size_t)(int)sc[kScCnt] * kHistRow + (blk % kHistSlots) * kRedStride], rr);
    }
    red_add(sc, kRedR, a ^ 1, rz);
  }
}

// K2: alpha = R[a] / PQ[a]; x += alpha p; r -= alpha q (skipped when first);
// z = B_fine^-1 r; R[a^1] += r . z; rc += P^T r (coarse segments are whole
// multiples of the fine segment S); hist[cnt] += |r|^2.
// Eight lanes per fine segment, lane i < 6 owning row i of every 6-vector:
// the block-tridiagonal LDL^T solve is sequential along the segment, so its
// parallelism is the six rows of each 6x6 product (operands exchanged by
// shuffles within the group).  One thread per segment left most of the chip
// idle (1,250 segments at 5k poses).
__global__ void __launch_bounds__(256) k_pcg_update(int N, int S, int first, int a,
                                                   double* __restrict__ sc, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   const float* __restrict__ Minv,
                                                   const float* __restrict__ Useg, int Sc,
                                                   const float* __restrict__ Pb, double* __restrict__ rc,
                                                   double* __restrict__ hist) {
  __shared__ double red[16];
  const int lane = threadIdx.x & 63;
  const int i = lane & 7;          // row of this lane (6, 7: idle rows)
  const int base = lane & ~7;      // first lane of the group
  const int seg = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const int k0 = seg * S;
  const bool row = i < 6;
  const int ir = row ? i : 0;      // idle rows shadow row 0 (never stored)
  double rr = 0.0, rz = 0.0;
  int cs_mine = -1;
  double racc = 0.0;
  const double pq = first ? 0.0 : red_sum(sc, kRedPQ, a);
  const double alpha = pq != 0.0 ? red_sum(sc, kRedR, a) / pq : 0.0;
  if (k0 < N) {
    const int k1 = min(k0 + S, N);
    if (!first) {
      for (int k = k0; k < k1; ++k) {
        const size_t o = (size_t)k * 6 + ir;
        const double xv = x[o] + alpha * p[o];
        const double rv = r[o] - alpha * q[o];
        if (row) {
          x[o] = xv;
          r[o] = rv;
          rr += rv * rv;
        }
      }
    }
    // forward: y_k = r_k - U_{k-1}^T t_{k-1}, t_k = Minv_k y_k
    float t = 0.f;
    for (int k = k0; k < k1; ++k) {
      float y = (float)r[(size_t)k * 6 + ir];
      if (k > k0) {
        const float* U = Useg + (size_t)(k - 1) * 36;
#pragma unroll
        for (int m = 0; m < 6; ++m) y -= U[m * 6 + ir] * __shfl(t, base + m);
      }
      const float* Di = Minv + (size_t)k * 36;
      float v = 0.f;
#pragma unroll
      for (int m = 0; m < 6; ++m) v += Di[ir * 6 + m] * __shfl(y, base + m);
      t = v;
      if (row) z[(size_t)k * 6 + i] = t;
    }
    // backward: z_k -= Minv_k U_k z_{k+1}
    float zn = t;
    for (int k = k1 - 2; k >= k0; --k) {
      const float* U = Useg + (size_t)k * 36;
      const float* Di = Minv + (size_t)k * 36;
      float u = 0.f;
#pragma unroll
      for (int m = 0; m < 6; ++m) u += U[ir * 6 + m] * __shfl(zn, base + m);
      float v = (float)z[(size_t)k * 6 + ir];
#pragma unroll
      for (int m = 0; m < 6; ++m) v -= Di[ir * 6 + m] * __shfl(u, base + m);
      zn = v;
      if (row) z[(size_t)k * 6 + i] = v;
    }
    for (int k = k0; k < k1; ++k) {
      const double rv = r[(size_t)k * 6 + ir];
      if (row) rz += rv * z[(size_t)k * 6 + i];
      if (Pb) {
        // coarse restriction, row i: sum_m P[m][i] r_m
        const float* P = Pb + (size_t)k * 36;
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) v += (double)P[m * 6 + ir] * __shfl(rv, base + m);
        racc += v;
      }
    }
    if (Pb && row) cs_mine = k0 / Sc;
  }
  pcg_update_tail(lane, i, row, cs_mine, racc, rr, rz, first, a, sc, Pb != nullptr, rc, hist, red);
}

// The six row values of an 8-lane group (lanes base .. base + 5) on every
// lane of the group, by DPP row broadcasts: a 16-lane row holds two groups, so
// banks 0-1 (lanes 0-7 of the row) take lane m and banks 2-3 lane 8 + m.  VALU
// moves, where a lane shuffle is an LDS round trip.
template <int M>
__device__ __forceinline__ float grp_bcast(float v) {
  const int iv = __float_as_int(v);
  const int lo = __builtin_amdgcn_update_dpp(0, iv, 0x150 + M, 0xf, 0x3, false);  // row_newbcast:M
  return __int_as_float(__builtin_amdgcn_update_dpp(lo, iv, 0x158 + M, 0xf, 0xc, false));
}
__device__ __forceinline__ void grp_bcast6(float v, float (&o)[6]) {
  o[0] = grp_bcast<0>(v);
  o[1] = grp_bcast<1>(v);
  o[2] = grp_bcast<2>(v);
  o[3] = grp_bcast<3>(v);
  o[4] = grp_bcast<4>(v);
  o[5] = grp_bcast<5>(v);
}

// K2 with the segment length S fixed at compile time: a 256-thread workgroup
// over 8 fine segments (8 S consecutive poses).  The generic kernel walks each
// of its four loops one pose at a time and waits on a global load in every
// step (4 S round trips per PCG iteration), with groups that put the few
// thousand segments of a large graph on a few dozen CUs.  Here the four
// wavefronts stage the poses' Minv and U blocks into LDS and update x and r
// entry-parallel (independent coalesced loads, all in flight at once); then
// wavefront 0 runs the sequential block-tridiagonal sweeps of the 8 segments
// from LDS (group operands exchanged by DPP row broadcasts) while wavefront 1
// forms the coarse restriction, and all four write z.  z, x, r and the restriction are the generic kernel's to the bit; |r|^2
// and r . z are the same sums in another order.
template <int S>
__global__ void __launch_bounds__(256) k_pcg_update_w(int N, int first, int a, double* __restrict__ sc,
                                                     const double* __restrict__ p,
                                                     const double* __restrict__ q, double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     const float* __restrict__ Minv,
                                                     const float* __restrict__ Useg, int Sc,
                                                     const float* __restrict__ Pb, double* __restrict__ rc,
                                                     double* __restrict__ hist) {
  constexpr int NP = 8 * S;    // poses of the workgroup
  constexpr int NV4 = NP * 9;  // float4 per 6x6 block array
  constexpr int NE = NP * 6;   // vector entries
  constexpr int NIT = (NE + 255) / 256;  // entries per thread
  constexpr int CH = NIT < 6 ? NIT : 6;  // entries per thread per load chunk
  static_assert(NIT % CH == 0, "segment length");
  __shared__ float4 sM[NV4], sU[NV4];
  __shared__ double sr[NE];
  __shared__ float sz[NE];
  __shared__ double red[16];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int kb = blockIdx.x * NP;
  const int np = min(NP, N - kb);  // >= 1 by the grid
  const int ne = 6 * np, nv = 9 * np;
  const size_t e0 = (size_t)kb * 6;
  // stage the blocks; slots past np take copies of the last block, never read
  {
    const float4* gM = reinterpret_cast<const float4*>(Minv) + (size_t)kb * 9;
    const float4* gU = reinterpret_cast<const float4*>(Useg) + (size_t)kb * 9;
#pragma unroll
    for (int it = 0; it < (NV4 + 255) / 256; ++it) {
      const int t = min(tid + 256 * it, NV4 - 1);
      const int g = min(t, nv - 1);
      sM[t] = gM[g];
      sU[t] = gU[g];
    }
  }
  double rr = 0.0;
  if (!first) {
    const double pq = red_sum(sc, kRedPQ, a);
    const double alpha = pq != 0.0 ? red_sum(sc, kRedR, a) / pq : 0.0;
#pragma unroll
    for (int c0 = 0; c0 < NIT; c0 += CH) {
      double xv[CH], pv[CH], rv[CH], qv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const size_t o = e0 + min(tid + 256 * (c0 + u), ne - 1);
        xv[u] = x[o];
        pv[u] = p[o];
        rv[u] = r[o];
        qv[u] = q[o];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int e = tid + 256 * (c0 + u);
        const double xn = xv[u] + alpha * pv[u];
        const double rn = rv[u] - alpha * qv[u];
        if (e < NE) sr[e] = rn;
        if (e < ne) {
          x[e0 + e] = xn;
          r[e0 + e] = rn;
          rr += rn * rn;
        }
      }
    }
  } else {
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = tid + 256 * it;
      const double v = r[e0 + min(e, ne - 1)];
      if (e < NE) sr[e] = v;
    }
  }
  __syncthreads();
  // eight lanes per segment, lane i < 6 owning row i
  const int i = lane & 7;
  const bool row = i < 6;
  const int ir = row ? i : 0;
  const int k0l = (lane >> 3) * S;  // first pose of the segment in the workgroup
  const int n = np - k0l;           // >= S: whole segment
  double racc = 0.0;
  int cs_mine = -1;
  if (wave == 0) {
    const float* fM = reinterpret_cast<const float*>(sM);
    const float* fU = reinterpret_cast<const float*>(sU);
    // forward: y_k = r_k - U_{k-1}^T t_{k-1}, t_k = Minv_k y_k
    float t = 0.f;
#pragma unroll
    for (int kk = 0; kk < S; ++kk) {
      if (kk < n) {
        const int kl = k0l + kk;
        float y = (float)sr[kl * 6 + ir];
        float g[6];
        if (kk > 0) {
          const float* U = fU + (kl - 1) * 36;
          grp_bcast6(t, g);
#pragma unroll
          for (int m = 0; m < 6; ++m) y -= U[m * 6 + ir] * g[m];
        }
        const float* Di = fM + kl * 36;
        grp_bcast6(y, g);
        float v = 0.f;
#pragma unroll
        for (int m = 0; m < 6; ++m) v += Di[ir * 6 + m] * g[m];
        t = v;
        if (row) sz[kl * 6 + i] = t;
      }
    }
    // backward: z_k -= Minv_k U_k z_{k+1}
    float zn = t;
#pragma unroll
    for (int kk = S - 2; kk >= 0; --kk) {
      if (kk < n - 1) {
        const int kl = k0l + kk;
        const float* U = fU + kl * 36;
        const float* Di = fM + kl * 36;
        float g[6];
        grp_bcast6(zn, g);
        float u = 0.f;
#pragma unroll
        for (int m = 0; m < 6; ++m) u += U[ir * 6 + m] * g[m];
        grp_bcast6(u, g);
        float v = sz[kl * 6 + ir];
#pragma unroll
        for (int m = 0; m < 6; ++m) v -= Di[ir * 6 + m] * g[m];
        zn = v;
        if (row) sz[kl * 6 + i] = v;
      }
    }
  } else if (wave == 1 && Pb) {
    // coarse restriction, row i of each segment: sum_k sum_m P_k[m][i] r_k[m]
    // (loads clamped into the workgroup's poses, poses past np add nothing)
#pragma unroll
    for (int kk = 0; kk < S; ++kk) {
      const int kl = min(k0l + kk, np - 1);
      const float* P = Pb + (size_t)(kb + kl) * 36;
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) v += (double)P[m * 6 + ir] * sr[kl * 6 + m];
      racc += kk < n ? v : 0.0;
    }
    if (row && n > 0) cs_mine = (kb + k0l) / Sc;
  }
  __syncthreads();
  double rz = 0.0;
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int e = tid + 256 * it;
    if (e < ne) {
      const float zv = sz[e];
      z[e0 + e] = zv;
      rz += sr[e] * (double)zv;
    }
  }
  pcg_update_tail(lane, i, row, cs_mine, racc, rr, rz, first, a, sc, Pb != nullptr && wave == 1, rc, hist,
                  red);
}

// K3: cc[a] += Ainv rc over a (64-row, 64-column) tile per workgroup;
// R[a^1] += rc . (Ainv rc) (the coarse part of r . z, since r . P c = P^T r . c).
__global__ void __launch_bounds__(256) k_pcg_coarse(int nc, int a, const double* __restrict__ Ainv,
                                                   const double* __restrict__ rc, double* __restrict__ cc,
                                                   double* __restrict__ sc) {
  __shared__ double part[4][64];
  __shared__ double red[16];
  const int row = blockIdx.x * 64 + (threadIdx.x & 63);
  const int qd = threadIdx.x >> 6;
  const int j0 = blockIdx.y * 64 + qd * 16, j1 = min(j0 + 16, nc);
  double s = 0.0;
  if (row < nc)
    for (int j = j0; j < j1; ++j) s += Ainv[(size_t)j * nc + row] * rc[j];  // Ainv symmetric
  part[qd][threadIdx.x & 63] = s;
  __syncthreads();
  double contrib = 0.0;
  if (qd == 0 && row < nc) {
    const double v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
    atomicAdd(&cc[row], v);
    contrib = rc[row] * v;
  }
  contrib = block_sum(contrib, red);
  if (threadIdx.x == 0) red_add(sc, kRedR, a ^ 1, contrib);
}

// K4: beta = R[a^1] / R[a] (0 when first); z += P cc[a]; p = z + beta p;
// clear rc and cc[a^1]; advance the iteration counter.  One thread per
// vector entry (pose k = e / 6, row e % 6).
__global__ void __launch_bounds__(256) k_pcg_direction(int N, int first, int a, double* __restrict__ sc,
                                                      double* __restrict__ z, double* __restrict__ p,
                                                      int Sc, const float* __restrict__ Pb,
                                                      const double* __restrict__ cc_a,
                                                      double* __restrict__ cc_other,
                                                      double* __restrict__ rc, int nc) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double rz = red_sum(sc, kRedR, a);
  const double beta = (!first && rz != 0.0) ? red_sum(sc, kRedR, a ^ 1) / rz : 0.0;
  if (e < 6ll * N) {
    const int k = (int)(e / 6), i = (int)(e % 6);
    double v = z[e];
    if (Pb) {
      const float* P = Pb + (size_t)k * 36 + i * 6;
      const double* c = cc_a + 6 * (k / Sc);
#pragma unroll
      for (int m = 0; m < 6; ++m) v += (double)P[m] * c[m];
    }
    p[e] = first ? v : v + beta * p[e];
  }
  if (e < nc) {
    rc[e] = 0.0;
    cc_other[e] = 0.0;
  }
  if (e == 0 && !first) sc[kScCnt] += 1.0;
}

__global__ void k_copy_d(long long n, const double* __restrict__ a, double* __restrict__ b) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    b[k] = a[k];
}

// x_new = x * Exp(delta) per pose
__global__ void k_retract(int N, const double* __restrict__ x, const double* __restrict__ dx,
                          double* __restrict__ xn, int chart) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double d[6], E[12], o[12];
#pragma unroll
  for (int e = 0; e < 6; ++e) d[e] = dx[(size_t)p * 6 + e];
  if (chart) {
    // x * Exp(out-of-plane part) * Pose2(dx_x, dx_y, dx_theta)
    double dout[6] = {d[0], d[1], 0.0, 0.0, 0.0, d[5]};
    double Eo[12], P2[12] = {cos(d[2]), -sin(d[2]), 0.0, sin(d[2]), cos(d[2]), 0.0, 0.0, 0.0, 1.0,
                             d[3], d[4], 0.0};
    se3_exp64(dout, Eo);
    R3::pose_compose(Eo, P2, E);
  } else {
    se3_exp64(d, E);
  }
  R3::pose_compose(x + (size_t)p * 12, E, o);
#pragma unroll
  for (int e = 0; e < 12; ++e) xn[(size_t)p * 12 + e] = o[e];
}

__global__ void k_sum(long long n, const double* __restrict__ a, const double* __restrict__ w,
                      double* __restrict__ out) {
  double s = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    s += w ? a[k] * w[k] : a[k];
  __shared__ double red[16];
  s = block_sum(s, red);
  if (threadIdx.x == 0) atomicAdd(out, s);
}

// b . dx in FP64 (b kept in FP64)
__global__ void k_tls_weights(int F, const double* __restrict__ err, const double* __restrict__ barc,
                              const unsigned char* __restrict__ known, double mu,
                              double* __restrict__ w) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  if (known[f]) {
    w[f] = 1.0;
    return;
  }
  const double u2 = err[f], b = barc[f];
  const double ub = (mu + 1.0) / mu * b, lb = mu / (mu + 1.0) * b;
  double wk = sqrt(b * mu * (mu + 1.0) / u2) - mu;
  if (u2 >= ub || wk < 0.0) wk = 0.0;
  else if (u2 <= lb || wk > 1.0) wk = 1.0;
  w[f] = wk;
}

// candidate mu per factor: barc / (2 r - barc) when positive, +inf otherwise
__global__ void k_mu_candidates(int F, const double* __restrict__ err,
                                const double* __restrict__ barc, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const double den = 2.0 * err[f] - barc[f];
  out[f] = den > 0.0 ? barc[f] / den : INFINITY;
}

__global__ void k_min(long long n, const double* __restrict__ a, unsigned long long* __restrict__ out) {
  double m = INFINITY;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    m = fmin(m, a[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  // non-negative doubles order like their bit patterns
  if ((threadIdx.x & 63) == 0) atomicMin(out, (unsigned long long)__double_as_longlong(m));
}

// weights binary within tol: counts the violations
// unknown factors whose final weight disagrees with the TLS decision at the
// final poses (inlier iff 0.5 e^T Lambda e < barc^2): 0 at a GNC fixed point
__global__ void k_tls_check(int F, const double* __restrict__ err, const double* __restrict__ barc,
                            const unsigned char* __restrict__ known, const double* __restrict__ w,
                            unsigned int* __restrict__ bad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F || known[f]) return;
  if ((err[f] < barc[f]) != (w[f] > 0.5)) atomicAdd(bad, 1u);
}

__global__ void k_weights_binary(int F, const double* __restrict__ w, double tol,
                                 unsigned int* __restrict__ bad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  if (fabs(w[f] - rint(w[f])) > tol) atomicAdd(bad, 1u);
}

// ------------------------------------------------------------------ host driver
static inline unsigned grid_for(long long n, int block = 256, long long cap = 2048) {
  long long g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int SolverRun::alloc(int N_, int F_) {
  N = N_;
  F = F_;
  int rc;
  // the captured PCG chunk holds buffer addresses: capture it again per solve
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  graph_exec = nullptr;
  graph_key = GraphKey{};
  const size_t n6 = (size_t)N * 6;
  if ((rc = x0.reserve(sizeof(double) * 12 * N))) return rc;
  if ((rc = x.reserve(sizeof(double) * 12 * N))) return rc;
  if ((rc = xn.reserve(sizeof(double) * 12 * N))) return rc;
  if ((rc = kind.reserve(sizeof(int) * F))) return rc;
  if ((rc = fi.reserve(sizeof(int) * F))) return rc;
  if ((rc = fj.reserve(sizeof(int) * F))) return rc;
  if ((rc = z.reserve(sizeof(double) * 12 * F))) return rc;
  if ((rc = info.reserve(sizeof(double) * 36 * F))) return rc;
  if ((rc = w.reserve(sizeof(double) * F))) return rc;
  if ((rc = barc.reserve(sizeof(double) * F))) return rc;
  if ((rc = known.reserve(F))) return rc;
  if ((rc = err.reserve(sizeof(double) * F))) return rc;
  if ((rc = cost.reserve(sizeof(double) * F))) return rc;
  if ((rc = Hii.reserve(sizeof(double) * 36 * F))) return rc;
  if ((rc = Hjj.reserve(sizeof(double) * 36 * F))) return rc;
  if ((rc = Hij.reserve(sizeof(double) * 36 * F))) return rc;
  if ((rc = gi.reserve(sizeof(double) * 6 * F))) return rc;
  if ((rc = gj.reserve(sizeof(double) * 6 * F))) return rc;
  if ((rc = inc_off.reserve(sizeof(int) * (N + 1)))) return rc;
  if ((rc = inc_cnt.reserve(sizeof(int) * (N + 1)))) return rc;
  if ((rc = inc.reserve(sizeof(int) * 2 * F))) return rc;
  if ((rc = inc_q.reserve(sizeof(int) * 2 * F))) return rc;
  if ((rc = D.reserve(sizeof(double) * 36 * N))) return rc;
  if ((rc = Minv.reserve(sizeof(float) * 36 * N))) return rc;
  if ((rc = Useg.reserve(sizeof(float) * 36 * N))) return rc;
  if ((rc = chain.reserve(sizeof(int) * (N + 1)))) return rc;
  if ((rc = ext_off.reserve(sizeof(int) * (N + 1)))) return rc;
  if ((rc = ext.reserve(sizeof(int) * 2 * F))) return rc;
  if ((rc = ext_q.reserve(sizeof(int) * 2 * F))) return rc;
  if ((rc = Ds.reserve(sizeof(double) * 36 * N))) return rc;
  if ((rc = Cs.reserve(sizeof(double) * 36 * N))) return rc;
  if ((rc = hdiag.reserve(sizeof(double) * n6))) return rc;
  for (DevBuf* v : {&b64, &dx, &r, &zv, &p, &q, &hx})
    if ((rc = v->reserve(sizeof(double) * n6))) return rc;
  if ((rc = sc.reserve(sizeof(double) * kScTotal))) return rc;
  return RPGO_OK;
}

// pinned host scratch for scalar readbacks
double* SolverRun::pinned_scalars() {
  if (!h_scalars &&
      hipHostMalloc((void**)&h_scalars, sizeof(double) * 64, hipHostMallocDefault) != hipSuccess)
    h_scalars = nullptr;
  return h_scalars;
}

// pinned host scratch for the residual history of one PCG chunk
double* SolverRun::pinned_history() {
  if (!h_history && hipHostMalloc((void**)&h_history, sizeof(double) * kHistRow * kPcgChunk,
                                  hipHostMallocDefault) != hipSuccess)
    h_history = nullptr;
  return h_history;
}

int SolverRun::read_scalars(hipStream_t st, int count, double* out) {
  double* h = pinned_scalars();
  if (!h) {
    set_error("hipHostMalloc of the scalar readback failed");
    return RPGO_ERR_NOMEM;
  }
  RPGO_HIP_CHECK(hipMemcpyAsync(h, sc.ptr, sizeof(double) * count, hipMemcpyDeviceToHost, st));
  RPGO_HIP_CHECK(hipStreamSynchronize(st));
  for (int k = 0; k < count; ++k) out[k] = h[k];
  return RPGO_OK;
}

int SolverRun::build_incidence(hipStream_t st) {
  RPGO_HIP_CHECK(hipMemsetAsync(inc_cnt.ptr, 0, sizeof(int) * (N + 1), st));
  k_inc_count<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, kind.as<int>(), fi.as<int>(), fj.as<int>(),
                                                        inc_cnt.as<int>());
  k_exclusive_scan<<<1, 1024, 0, st>>>(N, inc_cnt.as<int>(), inc_off.as<int>());
  RPGO_HIP_CHECK(hipMemsetAsync(inc_cnt.ptr, 0, sizeof(int) * (N + 1), st));
  k_inc_fill<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, kind.as<int>(), fi.as<int>(), fj.as<int>(),
                                                       inc_off.as<int>(), inc_cnt.as<int>(),
                                                       inc.as<int>());
  k_inc_sort<<<grid_for(N, 256, 1 << 20), 256, 0, st>>>(N, inc_off.as<int>(), inc.as<int>());
  k_inc_other<<<grid_for(N, 256, 1 << 20), 256, 0, st>>>(N, inc_off.as<int>(), inc.as<int>(), kind.as<int>(),
                                                         fi.as<int>(), fj.as<int>(), inc_q.as<int>());
  k_chain_init<<<grid_for(N + 1, 256, 1 << 20), 256, 0, st>>>(N + 1, chain.as<int>());
  k_chain_find<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, kind.as<int>(), fi.as<int>(), fj.as<int>(),
                                                          chain.as<int>());
  // extras of the chain SpMV: every between incidence that is not a chain link
  k_ext_count<<<grid_for(N, 256, 1 << 20), 256, 0, st>>>(N, inc_off.as<int>(), inc.as<int>(), kind.as<int>(),
                                                         chain.as<int>(), inc_cnt.as<int>());
  k_exclusive_scan<<<1, 1024, 0, st>>>(N, inc_cnt.as<int>(), ext_off.as<int>());
  k_ext_fill<<<grid_for(N, 256, 1 << 20), 256, 0, st>>>(N, inc_off.as<int>(), inc.as<int>(), inc_q.as<int>(),
                                                        kind.as<int>(), chain.as<int>(), ext_off.as<int>(),
                                                        ext.as<int>(), ext_q.as<int>());
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

// total weighted cost at poses xp (FP64)
int SolverRun::total_cost(hipStream_t st, const double* xp, double* out) {
  k_factor_errors<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, kind.as<int>(), fi.as<int>(), fj.as<int>(),
                                                            z.as<double>(), info.as<double>(), xp,
                                                            err.as<double>(), prm.pose2_chart);
  RPGO_HIP_CHECK(hipMemsetAsync(sc.as<double>() + 8, 0, sizeof(double), st));
  k_sum<<<grid_for(F), 256, 0, st>>>(F, err.as<double>(), w.as<double>(), sc.as<double>() + 8);
  RPGO_HIP_CHECK(hipGetLastError());
  double h[9];
  int rc = read_scalars(st, 9, h);
  if (rc) return rc;
  *out = h[8];
  return RPGO_OK;
}

int SolverRun::linearize(hipStream_t st) {
  k_linearize<<<grid_for(F, 64, 1 << 22), 64, 0, st>>>(
      F, kind.as<int>(), fi.as<int>(), fj.as<int>(), z.as<double>(), info.as<double>(), w.as<double>(),
      x.as<double>(), Hii.as<double>(), Hjj.as<double>(), Hij.as<double>(), gi.as<double>(),
      gj.as<double>(), cost.as<double>(), prm.pose2_chart);
  k_assemble<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(
      N, inc_off.as<int>(), inc.as<int>(), Hii.as<double>(), Hjj.as<double>(), gi.as<double>(),
      gj.as<double>(), D.as<double>(), b64.as<double>(), hdiag.as<double>());
  k_chain_blocks<<<grid_for(36ll * N, 256, 1 << 22), 256, 0, st>>>(N, chain.as<int>(), D.as<double>(),
                                                                   Hij.as<double>(), Ds.as<double>(),
                                                                   Cs.as<double>());
  RPGO_HIP_CHECK(hipGetLastError());
  ++stats.linearizations;
  return RPGO_OK;
}

// y = (H + lambda damp) x
void SolverRun::spmv(hipStream_t st, double lambda, const double* xin, double* yout) {
  k_spmv<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(
      N, inc_off.as<int>(), inc.as<int>(), fi.as<int>(), fj.as<int>(), kind.as<int>(), D.as<double>(),
      Hij.as<double>(), hdiag.as<double>(), nullptr, lambda, prm.lm_diagonal_damping, 1e-6, 1e32, xin,
      yout);
}

// same, lambda read from device memory (for the captured PCG body)
void SolverRun::spmv_dev(hipStream_t st, const double* lam, const double* xin, double* yout) {
  k_spmv<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(
      N, inc_off.as<int>(), inc.as<int>(), fi.as<int>(), fj.as<int>(), kind.as<int>(), D.as<double>(),
      Hij.as<double>(), hdiag.as<double>(), lam, 0.0, prm.lm_diagonal_damping, 1e-6, 1e32, xin, yout);
}

// Poses per fine preconditioner segment.  Automatic (pcg_segment <= 0): 16
// below 10,000 poses, 32 from there.  Longer segments hold more of the
// odometry chain (fewer PCG iterations) but make the coarse segments, whole
// multiples of S, coarser; on SE(3) helices with 50 % outliers, plain GNC:
// 5k poses S = 16 / 32 -> 1.73 / 1.74 s, 20k 8.23 / 7.64 s, 50k (LM on the
// outlier-free graph) 1.54 / 1.42 s.
static int fine_segment(const rpgo_solve_params& prm, int N) {
  if (prm.pcg_segment > 0) return prm.pcg_segment;
  return N < 10000 ? 16 : 32;
}

// Coarse level set-up for one linear solve: prolongation blocks at the current
// poses, Ac = P^T (H + lambda damp) P, its Cholesky factor, L^-1 and the FP32
// explicit inverse.  *nseg = 0 when the coarse level is off.
int SolverRun::coarse_setup(hipStream_t st, double lambda, int* Sc_out, int* nseg_out) {
  *Sc_out = 0;
  *nseg_out = 0;
  if (prm.pcg_coarse < 0 || N < 2) return RPGO_OK;
  // more rigid segments cut PCG iterations but cost nc^2 per iteration and
  // nc^3 per set-up: on SE(3) helices with 50 % outliers, plain GNC with
  // 128 / 256 segments took 1.44 / 1.48 s at 5k poses, 7.02 / 6.68 s at 20k
  // and 42.4 / 39.7 s at 50k
  const int segs_max = prm.pcg_coarse > 0 ? prm.pcg_coarse : (N < kCoarseLargeN ? kCoarseSegSmall : kCoarseSegMax);
  const int Sfine = fine_segment(prm, N);
  int Sc = std::max(Sfine, (N + segs_max - 1) / segs_max);
  Sc = (Sc + Sfine - 1) / Sfine * Sfine;  // whole fine segments per coarse segment
  const int nseg = (N + Sc - 1) / Sc;
  if (nseg < 2) return RPGO_OK;  // a single rigid segment is the gauge mode only
  const int nc = 6 * nseg;
  int rc_;
  if ((rc_ = Pb.reserve(sizeof(float) * 36 * (size_t)N))) return rc_;
  if ((rc_ = Ac.reserve(sizeof(double) * (size_t)nc * nc))) return rc_;
  if ((rc_ = Lci.reserve(sizeof(double) * (size_t)nc * nc))) return rc_;
  if ((rc_ = LcT.reserve(sizeof(double) * (size_t)nc * nc))) return rc_;
  if ((rc_ = Ainv.reserve(sizeof(double) * (size_t)nc * nc))) return rc_;
  k_coarse_basis<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(N, Sc, x.as<double>(), Pb.as<float>());
  RPGO_HIP_CHECK(hipMemsetAsync(Ac.ptr, 0, sizeof(double) * (size_t)nc * nc, st));
  k_coarse_diag<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(N, Sc, nc, Pb.as<float>(), D.as<double>(),
                                                          hdiag.as<double>(), lambda,
                                                          prm.lm_diagonal_damping, Ac.as<double>());
  k_coarse_offdiag<<<grid_for(F, 128, 1 << 22), 128, 0, st>>>(F, Sc, nc, kind.as<int>(), fi.as<int>(),
                                                             fj.as<int>(), Hij.as<double>(),
                                                             Pb.as<float>(), Ac.as<double>());
  for (int k0 = 0; k0 < nc; k0 += kCholNb) {
    k_chol_panel<<<1, 64, 0, st>>>(nc, k0, Ac.as<double>(), LcT.as<double>());
    const int rest = nc - k0 - kCholNb;
    if (rest > 0) {
      const int T = (rest + 31) / 32;
      k_chol_update<<<T * (T + 1) / 2, 256, 0, st>>>(nc, k0, Ac.as<double>(), LcT.as<double>());
    }
  }
  if (nc <= 64 * 12)
    k_tri_inv<12><<<(nc + 3) / 4, 256, 0, st>>>(nc, LcT.as<double>(), Lci.as<double>());
  else
    k_tri_inv<kCoarseMax / 64><<<(nc + 3) / 4, 256, 0, st>>>(nc, LcT.as<double>(), Lci.as<double>());
  {
    const int T = (nc + 31) / 32;
    k_gram_tiles<<<T * (T + 1) / 2, 256, 0, st>>>(nc, T, Lci.as<double>(), Ainv.as<double>());
  }
  RPGO_HIP_CHECK(hipGetLastError());
  *Sc_out = Sc;
  *nseg_out = nseg;
  return RPGO_OK;
}

// One PCG iteration (a = iteration parity), or with first = 1 the set-up
// z0 = M^-1 r0, p0 = z0 written into the slots of iteration 0.
void SolverRun::pcg_step(hipStream_t st, int S, int Sc, int nseg_c, int first, int a) {
  double* s = sc.as<double>();
  const int nseg = (N + S - 1) / S;
  const int nc = 6 * nseg_c;
  const float* P = nseg_c ? Pb.as<float>() : nullptr;
  double* cc0 = ccor.as<double>();
  double* cc1 = cc0 + kCoarseMax;
  // both kernels sum every row in the same order (q is the same to the bit);
  // rows in parallel below kSpmvPoseMin poses, one thread per pose above
  // (5k poses: 46 vs 49 us per PCG iteration; 50k: 78 vs 69 us)
  if (!first && kSpmvChain)
    k_pcg_spmv_chain<<<(unsigned)((N + 63) / 64), 384, 0, st>>>(
        N, a, chain.as<int>(), Ds.as<double>(), Cs.as<double>(), ext_off.as<int>(), ext.as<int>(),
        ext_q.as<int>(), Hij.as<double>(), hdiag.as<double>(), prm.lm_diagonal_damping, s, p.as<double>(),
        q.as<double>());
  else if (!first && N < kSpmvPoseMin)
    k_pcg_spmv_rows<<<grid_for((long long)N * 8, 256, 1 << 22), 256, 0, st>>>(
        N, a, inc_off.as<int>(), inc.as<int>(), inc_q.as<int>(),
        D.as<double>(), Hij.as<double>(), hdiag.as<double>(), prm.lm_diagonal_damping, s, p.as<double>(),
        q.as<double>());
  else if (!first)
    k_pcg_spmv_pose<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(
        N, a, inc_off.as<int>(), inc.as<int>(), fi.as<int>(), fj.as<int>(), kind.as<int>(),
        D.as<double>(), Hij.as<double>(), hdiag.as<double>(), prm.lm_diagonal_damping, s, p.as<double>(),
        q.as<double>());
  // segment lengths with a fixed-S kernel: one workgroup per 8 segments
  const unsigned ugw = (unsigned)((N + 8ll * S - 1) / (8ll * S));
#define RPGO_UPDATE_ARGS                                                                                  \
  N, first, a, s, p.as<double>(), q.as<double>(), dx.as<double>(), r.as<double>(), zv.as<double>(),      \
      Minv.as<float>(), Useg.as<float>(), Sc, P, cres.as<double>(), hist.as<double>()
  switch (S) {
    case 4: k_pcg_update_w<4><<<ugw, 256, 0, st>>>(RPGO_UPDATE_ARGS); break;
    case 8: k_pcg_update_w<8><<<ugw, 256, 0, st>>>(RPGO_UPDATE_ARGS); break;
    case 16: k_pcg_update_w<16><<<ugw, 256, 0, st>>>(RPGO_UPDATE_ARGS); break;
    case 32: k_pcg_update_w<32><<<ugw, 256, 0, st>>>(RPGO_UPDATE_ARGS); break;
    default:
      k_pcg_update<<<grid_for((long long)nseg * 8, 256, 1 << 22), 256, 0, st>>>(
          N, S, first, a, s, p.as<double>(), q.as<double>(), dx.as<double>(), r.as<double>(), zv.as<double>(),
          Minv.as<float>(), Useg.as<float>(), Sc, P, cres.as<double>(), hist.as<double>());
  }
#undef RPGO_UPDATE_ARGS
  if (nseg_c) {
    const dim3 grid((nc + 63) / 64, (nc + 63) / 64);
    k_pcg_coarse<<<grid, 256, 0, st>>>(nc, a, Ainv.as<double>(), cres.as<double>(), a ? cc1 : cc0, s);
  }
  k_pcg_direction<<<grid_for(std::max(6ll * N, (long long)nc), 256, 1 << 22), 256, 0, st>>>(
      N, first, a, s, zv.as<double>(), p.as<double>(), Sc, P, a ? cc1 : cc0, a ? cc0 : cc1,
      cres.as<double>(), nc);
}

// Solve (H + lambda damp) dx = b by PCG (x0 = 0) preconditioned by the
// segment block-tridiagonal solve plus the rigid-segment coarse correction.
int SolverRun::pcg(hipStream_t st, double lambda, int* iters_out) {
  const long long n = (long long)N * 6;
  double* s = sc.as<double>();
  const int S = fine_segment(prm, N);
  const int nseg = (N + S - 1) / S;
  k_seg_factor<<<(nseg + 63) / 64, 64, 0, st>>>(N, S, chain.as<int>(), D.as<double>(), hdiag.as<double>(),
                                                Hij.as<double>(), lambda, prm.lm_diagonal_damping,
                                                Minv.as<float>(), Useg.as<float>());
  int Sc = 0, nseg_c = 0, rc_;
  if ((rc_ = coarse_setup(st, lambda, &Sc, &nseg_c))) return rc_;
  const int max_it = prm.pcg_max_iterations > 0 ? prm.pcg_max_iterations : std::min(6 * N, 4000);
  if ((rc_ = hist.reserve(sizeof(double) * kHistRow * (size_t)(max_it + kPcgChunk)))) return rc_;
  if ((rc_ = ccor.reserve(sizeof(double) * 2 * kCoarseMax))) return rc_;
  if ((rc_ = cres.reserve(sizeof(double) * kCoarseMax))) return rc_;
  RPGO_HIP_CHECK(hipMemsetAsync(hist.ptr, 0, sizeof(double) * kHistRow * (size_t)(max_it + kPcgChunk), st));
  RPGO_HIP_CHECK(hipMemsetAsync(ccor.ptr, 0, sizeof(double) * 2 * kCoarseMax, st));
  RPGO_HIP_CHECK(hipMemsetAsync(cres.ptr, 0, sizeof(double) * kCoarseMax, st));
  RPGO_HIP_CHECK(hipMemsetAsync(dx.ptr, 0, sizeof(double) * n, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s, 0, sizeof(double) * 8, st));
  RPGO_HIP_CHECK(hipMemsetAsync(s + kScCnt, 0, sizeof(double), st));
  RPGO_HIP_CHECK(hipMemsetAsync(s + kScSlots, 0, sizeof(double) * (kScTotal - kScSlots), st));
  RPGO_HIP_CHECK(hipMemcpyAsync(s + kScLam, &lambda, sizeof(double), hipMemcpyHostToDevice, st));
  k_copy_d<<<grid_for(n), 256, 0, st>>>(n, b64.as<double>(), r.as<double>());
  k_dot<<<grid_for(n, 256, 512), 256, 0, st>>>(n, b64.as<double>(), b64.as<double>(), s + 4);
  pcg_step(st, S, Sc, nseg_c, 1, 1);  // z0, p0; r0 . z0 lands in R[0]
  RPGO_HIP_CHECK(hipGetLastError());
  double h[8];
  int rc = read_scalars(st, 8, h);
  if (rc) return rc;
  const double bnorm2 = h[4];
  const double tol2 = prm.pcg_rel_tol * prm.pcg_rel_tol * bnorm2;
  int it = 0;
  if (bnorm2 == 0.0) {
    *iters_out = 0;
    return RPGO_OK;
  }
  auto body = [&]() {
    for (int j = 0; j < kPcgChunk; ++j) pcg_step(st, S, Sc, nseg_c, 0, j & 1);
  };
  // one captured chunk of kPcgChunk iterations, rebuilt when sizes or buffers change
  const GraphKey key{N,
                     S,
                     Sc,
                     nseg_c,
                     (const void*)hist.ptr,
                     (const void*)p.ptr,
                     (const void*)Minv.ptr,
                     (const void*)Hij.ptr,
                     (const void*)inc.ptr,
                     (const void*)inc_q.ptr,
                     nseg_c ? (const void*)Ainv.ptr : nullptr,
                     nseg_c ? (const void*)Pb.ptr : nullptr,
                     (const void*)Ds.ptr,
                     (const void*)ext.ptr};
  const bool graphs = use_graphs && prm.pcg_graphs != 0;
  if (graphs && !(graph_exec && graph_key == key)) {
    if (graph_exec) {
      (void)hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    RPGO_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    body();
    RPGO_HIP_CHECK(hipStreamEndCapture(st, &g));
    hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      graph_exec = nullptr;
      use_graphs = false;  // fall back to direct launches of the same kernels
    } else {
      graph_key = key;
    }
  }
  double best = bnorm2;
  int best_it = 0;
  bool done = false;
  while (it < max_it && !done) {
    if (graphs && graph_exec) RPGO_HIP_CHECK(hipGraphLaunch(graph_exec, st));
    else body();
    RPGO_HIP_CHECK(hipGetLastError());
    double hh[kPcgChunk];
    double* hb = pinned_history();
    if (!hb) {
      set_error("hipHostMalloc of the residual history readback failed");
      return RPGO_ERR_NOMEM;
    }
    RPGO_HIP_CHECK(hipMemcpyAsync(hb, hist.as<double>() + (size_t)it * kHistRow,
                                  sizeof(double) * kHistRow * kPcgChunk, hipMemcpyDeviceToHost, st));
    RPGO_HIP_CHECK(hipStreamSynchronize(st));
    for (int j = 0; j < kPcgChunk; ++j) {
      double v = 0.0;
      for (int k = 0; k < kHistSlots; ++k) v += hb[(size_t)j * kHistRow + k * kRedStride];
      hh[j] = v;
    }
    for (int j = 0; j < kPcgChunk; ++j) {
      const double rr = hh[j];
      if (!(rr == rr)) {  // NaN: keep the iterate of the last chunk boundary
        done = true;
        break;
      }
      if (rr <= tol2) done = true;
      if (rr < 0.98 * best) {
        best = rr;
        best_it = it + j + 1;
      }
    }
    it += kPcgChunk;
    // stagnation: the residual stopped decreasing
    if (it - best_it >= 32 * kPcgChunk) done = true;
  }
  stats.pcg_iterations += it;
  *iters_out = it;
  return RPGO_OK;
}

// model decrease L(0) - L(dx) = b.dx - 0.5 dx^T H dx (undamped H)
int SolverRun::model_decrease(hipStream_t st, double* out) {
  const long long n = (long long)N * 6;
  double* s = sc.as<double>();
  spmv(st, 0.0, dx.as<double>(), hx.as<double>());
  RPGO_HIP_CHECK(hipMemsetAsync(s + 10, 0, sizeof(double) * 2, st));
  k_dot<<<grid_for(n), 256, 0, st>>>(n, b64.as<double>(), dx.as<double>(), s + 10);
  k_dot<<<grid_for(n), 256, 0, st>>>(n, dx.as<double>(), hx.as<double>(), s + 11);
  RPGO_HIP_CHECK(hipGetLastError());
  double h[12];
  int rc = read_scalars(st, 12, h);
  if (rc) return rc;
  *out = h[10] - 0.5 * h[11];
  return RPGO_OK;
}

static bool lm_converged(const rpgo_solve_params& p, double cur, double nw) {
  if (nw <= 0.0) return true;
  const double dec = cur - nw;
  return (dec / cur <= p.lm_relative_error_tol) || (dec <= p.lm_absolute_error_tol);
}

// One LevenbergMarquardtOptimizer / GaussNewtonOptimizer run from x0 with the
// current weights; leaves the result in x and its weighted cost in *cost_out.
int SolverRun::nonlinear_solve(hipStream_t st, bool from_x0, double* cost_out, int* iters_out) {
  int rc;
  if (from_x0) RPGO_HIP_CHECK(hipMemcpyAsync(x.ptr, x0.ptr, sizeof(double) * 12 * N, hipMemcpyDeviceToDevice, st));
  double err;
  if ((rc = total_cost(st, x.as<double>(), &err))) return rc;
  double lambda = prm.lambda_initial;
  int it = 0;
  const bool gn = prm.solver == RPGO_SOLVER_GN;
  while (it < prm.lm_max_iterations) {
    if ((rc = linearize(st))) return rc;
    if (gn) {
      int pit;
      if ((rc = pcg(st, 0.0, &pit))) return rc;
      k_retract<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(N, x.as<double>(), dx.as<double>(), xn.as<double>(),
                                                           prm.pose2_chart);
      double nerr;
      if ((rc = total_cost(st, xn.as<double>(), &nerr))) return rc;
      ++it;
      if (polish && !(nerr < err)) break;  // refinement keeps the better iterate
      std::swap(x, xn);
      const bool conv = lm_converged(prm, err, nerr);
      err = nerr;
      if (conv) break;
      continue;
    }
    bool accepted = false;
    double nerr = err;
    while (true) {
      int pit;
      if ((rc = pcg(st, lambda, &pit))) return rc;
      double lin;
      if ((rc = model_decrease(st, &lin))) return rc;
      k_retract<<<grid_for(N, 128, 1 << 22), 128, 0, st>>>(N, x.as<double>(), dx.as<double>(), xn.as<double>(),
                                                           prm.pose2_chart);
      if ((rc = total_cost(st, xn.as<double>(), &nerr))) return rc;
      const double fid = lin > 0.0 ? (err - nerr) / lin : -1.0;
      if (lin >= 0.0 && fid >= prm.min_model_fidelity) {
        accepted = true;
        lambda = std::max(lambda / prm.lambda_factor, 0.0);
        break;
      }
      lambda = std::max(lambda * prm.lambda_factor, 1e-12);  // a zero lambda must still grow
      ++stats.lm_rejections;
      if (lambda >= prm.lambda_upper_bound) break;
    }
    ++it;
    if (!accepted) break;
    const bool conv = lm_converged(prm, err, nerr);
    std::swap(x, xn);
    err = nerr;
    if (conv) break;
  }
  stats.lm_iterations += it;
  *cost_out = err;
  *iters_out = it;
  return RPGO_OK;
}

int SolverRun::compute_errors(hipStream_t st, const double* xp) {
  k_factor_errors<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, kind.as<int>(), fi.as<int>(), fj.as<int>(),
                                                            z.as<double>(), info.as<double>(), xp,
                                                            err.as<double>(), prm.pose2_chart);
  RPGO_HIP_CHECK(hipGetLastError());
  return RPGO_OK;
}

int SolverRun::run(hipStream_t st, rpgo_solve_report* rep) {
  int rc;
  stats = Stats{};
  if ((rc = build_incidence(st))) return rc;
  int lm_it = 0;
  double cost = 0.0;
  int gnc_it = 0;
  double mu = -1.0;
  if (!prm.use_gnc) {
    if ((rc = nonlinear_solve(st, true, &cost, &lm_it))) return rc;
  } else {
    // first solve with the initial weights (all 1 or bias_odom / caller weights)
    if ((rc = nonlinear_solve(st, true, &cost, &lm_it))) return rc;
    if (prm.init_mu > 0.0) {
      mu = prm.init_mu;
    } else {
      // mu from the residuals at the initial values (TLS, Remark 5)
      if ((rc = compute_errors(st, x0.as<double>()))) return rc;
      k_mu_candidates<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, err.as<double>(), barc.as<double>(),
                                                                cost_buf());
      const unsigned long long inf_bits = 0x7ff0000000000000ull;
      RPGO_HIP_CHECK(hipMemcpyAsync(sc.as<double>() + 20, &inf_bits, 8, hipMemcpyHostToDevice, st));
      k_min<<<grid_for(F), 256, 0, st>>>(F, cost_buf(),
                                         reinterpret_cast<unsigned long long*>(sc.as<double>() + 20));
      RPGO_HIP_CHECK(hipGetLastError());
      double h[21];
      if ((rc = read_scalars(st, 21, h))) return rc;
      mu = h[20];
      if (mu >= 0.0 && mu < 1e-6) mu = 1e-6;
      mu = (mu > 0.0 && !std::isinf(mu)) ? mu : -1.0;
    }
    stats.mu_init = mu;
    double prev_cost = cost;
    const int n_unknown = F - n_known;
    if (!(mu <= 0.0 || n_unknown == 0)) {
      for (int k = 0; k < prm.gnc_max_iterations; ++k) {
        if ((rc = compute_errors(st, x.as<double>()))) return rc;
        k_tls_weights<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, err.as<double>(), barc.as<double>(),
                                                                known.as<unsigned char>(), mu,
                                                                w.as<double>());
        RPGO_HIP_CHECK(hipGetLastError());
        int it2;
        if ((rc = nonlinear_solve(st, !prm.gnc_warm_start, &cost, &it2))) return rc;
        gnc_it = k + 1;
        RPGO_HIP_CHECK(hipMemsetAsync(sc.as<double>() + 24, 0, sizeof(double), st));
        k_weights_binary<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(
            F, w.as<double>(), prm.gnc_weights_tol, reinterpret_cast<unsigned int*>(sc.as<double>() + 24));
        double h[25];
        if ((rc = read_scalars(st, 25, h))) return rc;
        unsigned int bad;
        std::memcpy(&bad, &h[24], sizeof(bad));
        const bool cost_conv = std::fabs(cost - prev_cost) / std::max(prev_cost, 1e-7) < prm.gnc_relative_cost_tol;
        if (cost_conv || bad == 0) break;
        mu *= prm.gnc_mu_step;
        prev_cost = cost;
      }
    }
  }
  if (prm.polish_iterations > 0) {
    // FP64-residual Gauss-Newton refinement of the final weighted problem
    rpgo_solve_params saved = prm;
    prm.solver = RPGO_SOLVER_GN;
    prm.lm_max_iterations = prm.polish_iterations;
    prm.lm_relative_error_tol = 1e-14;
    prm.lm_absolute_error_tol = 0.0;
    prm.pcg_rel_tol = std::min(prm.pcg_rel_tol, 1e-6);
    int it3;
    polish = true;
    rc = nonlinear_solve(st, false, &cost, &it3);
    polish = false;
    prm = saved;
    if (rc) return rc;
  }
  int disagree = 0;
  if (prm.use_gnc) {
    if ((rc = compute_errors(st, x.as<double>()))) return rc;
    RPGO_HIP_CHECK(hipMemsetAsync(sc.as<double>() + 24, 0, sizeof(double), st));
    k_tls_check<<<grid_for(F, 256, 1 << 20), 256, 0, st>>>(F, err.as<double>(), barc.as<double>(),
                                                          known.as<unsigned char>(), w.as<double>(),
                                                          reinterpret_cast<unsigned int*>(sc.as<double>() + 24));
    RPGO_HIP_CHECK(hipGetLastError());
    double h[25];
    if ((rc = read_scalars(st, 25, h))) return rc;
    unsigned int bad;
    std::memcpy(&bad, &h[24], sizeof(bad));
    disagree = (int)bad;
  }
  if (rep) {
    rep->n_tls_disagree = disagree;
    rep->gnc_iterations = gnc_it;
    rep->lm_iterations = stats.lm_iterations;
    rep->pcg_iterations = stats.pcg_iterations;
    rep->linearizations = stats.linearizations;
    rep->mu_init = stats.mu_init;
    rep->mu_final = mu;
    rep->final_cost = cost;
  }
  return RPGO_OK;
}

}  // namespace rpgo
