// FP64 device arithmetic of the PCM path, in the exact operation order of the
// reference's dense expressions (dot products summed k = 0..n-1, no FMA
// contraction: include only from translation units built with
// -ffp-contract=off).
//
//   PoseWithCovariance ops   GeometryUtils.h:91-186
//   Pose3/Pose2 group ops    gtsam::Pose3/Pose2 (operator*, inverse,
//                            AdjointMap, Logmap; LieGroup compose/between
//                            Jacobians H1 = Ad(g^-1), H1 = -Ad(result^-1))
//   LLT success test         Eigen unblocked LLT (pivot <= 0 fails)
//   inverse                  Eigen partial-pivot LU inverse
//
// PoseWithCovariance record (PWC doubles): pose (PL), covariance (D*D,
// row-major), rotation_info flag (1.0 / 0.0).
#pragma once

#include <hip/hip_runtime.h>

namespace rpgo {

constexpr double kPi = 3.14159265358979323846;

template <int N>
__device__ __forceinline__ void mm64(const double* A, const double* B, double* C) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = A[i * N] * B[j];
#pragma unroll
      for (int k = 1; k < N; ++k) s += A[i * N + k] * B[k * N + j];
      C[i * N + j] = s;
    }
}

template <int N>
__device__ __forceinline__ void mmt64(const double* A, const double* B, double* C) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = A[i * N] * B[j * N];
#pragma unroll
      for (int k = 1; k < N; ++k) s += A[i * N + k] * B[j * N + k];
      C[i * N + j] = s;
    }
}

// out = (H S) H^T
template <int N>
__device__ __forceinline__ void congruence64(const double* H, const double* S, double* out) {
  double T[N * N];
  mm64<N>(H, S, T);
  mmt64<N>(T, H, out);
}

// Eigen unblocked LLT: fails when a pivot x <= 0.
template <int N>
__device__ __forceinline__ bool llt_ok64(const double* A) {
  double L[N * N];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = A[k * N + k];
    if (k > 0) {
      double sq = L[k * N] * L[k * N];
#pragma unroll
      for (int j = 1; j < k; ++j) sq += L[k * N + j] * L[k * N + j];
      x -= sq;
    }
    ok = ok && (x > 0.0);
    x = sqrt(x);
    L[k * N + k] = x;
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      double v = A[i * N + k];
      if (k > 0) {
        double dot = L[i * N] * L[k * N];
#pragma unroll
        for (int j = 1; j < k; ++j) dot += L[i * N + j] * L[k * N + j];
        v -= dot;
      }
      L[i * N + k] = v / x;
    }
  }
  return ok;
}

// Partial-pivot LU inverse.  Row swaps are written as register selects so
// every array index stays a compile-time constant.
template <int N>
__device__ __forceinline__ void lu_inverse64(const double* A, double* Ainv) {
  double lu[N * N];
  int perm[N];
#pragma unroll
  for (int e = 0; e < N * N; ++e) lu[e] = A[e];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int p = k;
    double best = fabs(lu[k * N + k]);
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double v = fabs(lu[i * N + k]);
      if (v > best) { best = v; p = i; }
    }
    perm[k] = p;
    if (best != 0.0) {
#pragma unroll
      for (int i = k + 1; i < N; ++i) {
        const bool sw = (p == i);
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double a = lu[k * N + j], b = lu[i * N + j];
          lu[k * N + j] = sw ? b : a;
          lu[i * N + j] = sw ? a : b;
        }
      }
      const double piv = lu[k * N + k];
#pragma unroll
      for (int i = k + 1; i < N; ++i) lu[i * N + k] /= piv;
    }
#pragma unroll
    for (int i = k + 1; i < N; ++i)
#pragma unroll
      for (int j = k + 1; j < N; ++j) lu[i * N + j] -= lu[i * N + k] * lu[k * N + j];
  }
#pragma unroll
  for (int col = 0; col < N; ++col) {
    double b[N];
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = (i == col) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
      for (int i = k + 1; i < N; ++i) {
        const bool sw = (perm[k] == i);
        const double a = b[k], c = b[i];
        b[k] = sw ? c : a;
        b[i] = sw ? a : c;
      }
    }
#pragma unroll
    for (int i = 1; i < N; ++i) {
      double s = lu[i * N] * b[0];
#pragma unroll
      for (int k = 1; k < i; ++k) s += lu[i * N + k] * b[k];
      b[i] -= s;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      double v = b[i];
      if (i < N - 1) {
        double s = lu[i * N + i + 1] * b[i + 1];
#pragma unroll
        for (int k = i + 2; k < N; ++k) s += lu[i * N + k] * b[k];
        v -= s;
      }
      b[i] = v / lu[i * N + i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) Ainv[i * N + col] = b[i];
  }
}

template <int DIM>
struct R64;

// ------------------------------------------------------------------ SE(3)
template <>
struct R64<3> {
  static constexpr int PL = 12, D = 6, RD = 3, PWC = 12 + 36 + 1;

  __device__ static __forceinline__ void rot_vec(const double* R, const double* v, double* o) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = R[i * 3] * v[0];
      s += R[i * 3 + 1] * v[1];
      s += R[i * 3 + 2] * v[2];
      o[i] = s;
    }
  }
  __device__ static __forceinline__ void pose_compose(const double* a, const double* b, double* o) {
    double R[9], Rt[3];
    mm64<3>(a, b, R);
    rot_vec(a, b + 9, Rt);
    const double t0 = a[9] + Rt[0], t1 = a[10] + Rt[1], t2 = a[11] + Rt[2];
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = R[e];
    o[9] = t0; o[10] = t1; o[11] = t2;
  }
  // entry e of a*b (used by the lane-parallel fold)
  __device__ static __forceinline__ double compose_entry(const double* a, const double* b, int e) {
    if (e < 9) {
      const int r = e / 3, c = e % 3;
      double s = a[r * 3] * b[c];
      s += a[r * 3 + 1] * b[3 + c];
      s += a[r * 3 + 2] * b[6 + c];
      return s;
    }
    const int i = e - 9;
    double s = a[i * 3] * b[9];
    s += a[i * 3 + 1] * b[10];
    s += a[i * 3 + 2] * b[11];
    return a[9 + i] + s;
  }
  __device__ static __forceinline__ void pose_inverse(const double* a, double* o) {
    double Rt[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Rt[i * 3 + j] = a[j * 3 + i];
    const double mt[3] = {-a[9], -a[10], -a[11]};
    double t[3];
    rot_vec(Rt, mt, t);
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = Rt[e];
    o[9] = t[0]; o[10] = t[1]; o[11] = t[2];
  }
  __device__ static __forceinline__ void adjoint(const double* a, double* ad) {
    const double* R = a;
    const double tx = a[9], ty = a[10], tz = a[11];
    double A[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      A[0 * 3 + j] = -tz * R[3 + j] + ty * R[6 + j];
      A[1 * 3 + j] = tz * R[j] - tx * R[6 + j];
      A[2 * 3 + j] = -ty * R[j] + tx * R[3 + j];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        ad[i * 6 + j] = R[i * 3 + j];
        ad[i * 6 + j + 3] = 0.0;
        ad[(i + 3) * 6 + j] = A[i * 3 + j];
        ad[(i + 3) * 6 + j + 3] = R[i * 3 + j];
      }
  }
  __device__ static __forceinline__ void rot_logmap(const double* R, double* w) {
    const double R11 = R[0], R12 = R[1], R13 = R[2];
    const double R21 = R[3], R22 = R[4], R23 = R[5];
    const double R31 = R[6], R32 = R[7], R33 = R[8];
    const double tr = (R11 + R22) + R33;
    if (tr + 1.0 < 1e-10) {
      if (fabs(R33 + 1.0) > 1e-5) {
        const double f = kPi / sqrt(2.0 + 2.0 * R33);
        w[0] = f * R13; w[1] = f * R23; w[2] = f * (1.0 + R33);
      } else if (fabs(R22 + 1.0) > 1e-5) {
        const double f = kPi / sqrt(2.0 + 2.0 * R22);
        w[0] = f * R12; w[1] = f * (1.0 + R22); w[2] = f * R32;
      } else {
        const double f = kPi / sqrt(2.0 + 2.0 * R11);
        w[0] = f * (1.0 + R11); w[1] = f * R21; w[2] = f * R31;
      }
    } else {
      double magnitude;
      const double tr_3 = tr - 3.0;
      if (tr_3 < -1e-7) {
        const double theta = acos((tr - 1.0) / 2.0);
        magnitude = theta / (2.0 * sin(theta));
      } else {
        magnitude = 0.5 - tr_3 / 12.0;
      }
      w[0] = magnitude * (R32 - R23);
      w[1] = magnitude * (R13 - R31);
      w[2] = magnitude * (R21 - R12);
    }
  }
  __device__ static __forceinline__ void logmap(const double* a, double* xi) {
    double w[3];
    rot_logmap(a, w);
    const double T0 = a[9], T1 = a[10], T2 = a[11];
    const double t = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
    xi[0] = w[0]; xi[1] = w[1]; xi[2] = w[2];
    if (t < 1e-10) {
      xi[3] = T0; xi[4] = T1; xi[5] = T2;
      return;
    }
    const double n0 = w[0] / t, n1 = w[1] / t, n2 = w[2] / t;
    const double Tan = tan(0.5 * t);
    const double WT0 = -n2 * T1 + n1 * T2;
    const double WT1 = n2 * T0 - n0 * T2;
    const double WT2 = -n1 * T0 + n0 * T1;
    const double WWT0 = -n2 * WT1 + n1 * WT2;
    const double WWT1 = n2 * WT0 - n0 * WT2;
    const double WWT2 = -n1 * WT0 + n0 * WT1;
    const double ht = 0.5 * t;
    const double c = 1.0 - t / (2.0 * Tan);
    xi[3] = (T0 - ht * WT0) + c * WWT0;
    xi[4] = (T1 - ht * WT1) + c * WWT1;
    xi[5] = (T2 - ht * WT2) + c * WWT2;
  }
  __device__ static __forceinline__ double nan_trace(const double* cov) {
    return (cov[0] + cov[7]) + cov[14];
  }
#define RPGO_R64_COMMON
#include "replay64_common.inc"
#undef RPGO_R64_COMMON
};

// ------------------------------------------------------------------ SE(2)
template <>
struct R64<2> {
  static constexpr int PL = 4, D = 3, RD = 1, PWC = 4 + 9 + 1;

  __device__ static __forceinline__ void rot2_fix(double c, double s, double* oc, double* os) {
    double scale = c * c + s * s;
    if (fabs(scale - 1.0) > 1e-10) {
      scale = 1.0 / sqrt(scale);
      c *= scale;
      s *= scale;
    }
    *oc = c; *os = s;
  }
  __device__ static __forceinline__ void pose_compose(const double* a, const double* b, double* o) {
    double c, s;
    rot2_fix(a[0] * b[0] - a[1] * b[1], a[1] * b[0] + a[0] * b[1], &c, &s);
    const double rx = a[0] * b[2] - a[1] * b[3];
    const double ry = a[1] * b[2] + a[0] * b[3];
    o[0] = c; o[1] = s;
    o[2] = a[2] + rx; o[3] = a[3] + ry;
  }
  __device__ static __forceinline__ double compose_entry(const double* a, const double* b, int e) {
    double o[4];
    pose_compose(a, b, o);
    return o[e];
  }
  __device__ static __forceinline__ void pose_inverse(const double* a, double* o) {
    const double c = a[0], s = a[1];
    const double px = -a[2], py = -a[3];
    o[0] = c; o[1] = -s;
    o[2] = c * px + s * py;
    o[3] = -s * px + c * py;
  }
  __device__ static __forceinline__ void adjoint(const double* a, double* ad) {
    ad[0] = a[0]; ad[1] = -a[1]; ad[2] = a[3];
    ad[3] = a[1]; ad[4] = a[0]; ad[5] = -a[2];
    ad[6] = 0.0; ad[7] = 0.0; ad[8] = 1.0;
  }
  __device__ static __forceinline__ void logmap(const double* a, double* xi) {
    const double c = a[0], s = a[1], tx = a[2], ty = a[3];
    const double w = atan2(s, c);
    if (fabs(w) < 1e-10) {
      xi[0] = tx; xi[1] = ty; xi[2] = w;
      return;
    }
    const double c_1 = c - 1.0;
    const double det = c_1 * c_1 + s * s;
    const double ux = (c * tx + s * ty) - tx;
    const double uy = (-s * tx + c * ty) - ty;
    const double px = -uy, py = ux;
    const double f = w / det;
    xi[0] = f * px; xi[1] = f * py; xi[2] = w;
  }
  __device__ static __forceinline__ double nan_trace(const double* cov) { return cov[0]; }
#define RPGO_R64_COMMON
#include "replay64_common.inc"
#undef RPGO_R64_COMMON
};

// Sequential odometry fold of one chain by one wavefront (Pcm.h:516-557):
// for step j < n_steps, T_{j+1} = T_j * delta_j and
// Sigma_{j+1} = (H_j Sigma_j) H_j^T + Sigma_delta_j, in the oracle's order.
// Lanes 0..D*D-1 own the covariance entries, the next lanes the pose entries;
// S, P and ri hold the state on entry (step 0's start) and on exit.  hk, dk:
// H_j and the delta record of step 0 (rows of D*D and PWC doubles); out:
// the trajectory record written by step 0 (then one per step); ob: LDS for
// kFoldAhead records.
// The steps form one dependent chain, so the time per step is the latency of
// its LDS round trips; no global memory access may wait inside it.  Steps go
// in blocks of kFoldAhead: the next block's inputs are loaded into registers
// while this block runs, the block's records collect in LDS and are stored
// after it, and the only wait for global memory is at the block boundary.
// The workgroup is this one wavefront, and LDS operations of a wavefront are
// performed in issue order, so the steps order their LDS traffic with
// wavefront-scope fences (no instruction; they keep the compiler from moving
// memory operations across them).  A __syncthreads() per step is a
// workgroup-scope fence that waits for every global load and store in
// flight.  Callers compiled with -ffp-contract=off, launched with 64 threads.
constexpr int kFoldAhead = 8;

__device__ __forceinline__ void fold_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int DIM>
__device__ __forceinline__ void fold_run(int n_steps, const double* __restrict__ hk,
                                         const double* __restrict__ dk, double* __restrict__ out,
                                         double* S, double* T, double* P, double* Hs, double* Dl, int* ri,
                                         double* ob, int l) {
  constexpr int PL = R64<DIM>::PL, D = R64<DIM>::D, PWC = R64<DIM>::PWC, DD = D * D;
  static_assert(PWC <= 64 && DD <= 64, "one entry per lane");
  constexpr int A = kFoldAhead;
  const int lh = l < DD ? l : DD - 1, ld = l < PWC ? l : PWC - 1;  // in-row addresses for every lane
  double hx[A], dx[A], hy[A], dy[A];
#pragma unroll
  for (int u = 0; u < A; ++u) {
    const int j = u < n_steps ? u : 0;
    hx[u] = hk[(size_t)j * DD + lh];
    dx[u] = dk[(size_t)j * PWC + ld];
  }
  for (int j0 = 0; j0 < n_steps; j0 += A) {
    // the next block's inputs: independent of the chain, in flight meanwhile
#pragma unroll
    for (int u = 0; u < A; ++u) {
      const int j = j0 + A + u < n_steps ? j0 + A + u : 0;
      hy[u] = hk[(size_t)j * DD + lh];
      dy[u] = dk[(size_t)j * PWC + ld];
    }
    const int nb = n_steps - j0 < A ? n_steps - j0 : A;
#pragma unroll
    for (int u = 0; u < A; ++u) {
      if (u >= nb) break;
      if (l < DD) Hs[l] = hx[u];
      if (l < PWC) Dl[l] = dx[u];
      fold_wave_sync();
      // T = H * Sigma
      if (l < DD) {
        const int r = l / D, c = l % D;
        double s = Hs[r * D] * S[c];
#pragma unroll
        for (int m = 1; m < D; ++m) s += Hs[r * D + m] * S[m * D + c];
        T[l] = s;
      }
      double newpose = 0.0;
      if (l >= DD && l < DD + PL) newpose = R64<DIM>::compose_entry(P, Dl, l - DD);
      fold_wave_sync();
      double* o = ob + u * PWC;
      if (l < DD) {
        const int r = l / D, c = l % D;
        double s = T[r * D] * Hs[c * D];
#pragma unroll
        for (int m = 1; m < D; ++m) s += T[r * D + m] * Hs[c * D + m];
        s = s + Dl[PL + l];
        S[l] = s;
        o[PL + l] = s;
      }
      if (l >= DD && l < DD + PL) {
        P[l - DD] = newpose;
        o[l - DD] = newpose;
      }
      if (l == 0) {
        *ri = (*ri && Dl[PL + DD] != 0.0) ? 1 : 0;
        o[PL + DD] = *ri ? 1.0 : 0.0;
      }
      fold_wave_sync();
    }
    // the block's records, stored contiguously
    for (int e = l; e < nb * PWC; e += 64) out[(size_t)j0 * PWC + e] = ob[e];
#pragma unroll
    for (int u = 0; u < A; ++u) {
      hx[u] = hy[u];
      dx[u] = dy[u];
    }
  }
}

}  // namespace rpgo
