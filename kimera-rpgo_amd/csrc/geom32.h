// FP32 SE(3) arithmetic for the PCM fast path.
//
// Poses are (R, t) with R row-major; symmetric 6x6 covariances in the
// (omega, v) tangent order are kept as three 3x3 blocks
//   A = S[omega,omega] (sym, 6), B = S[omega,v] (9), C = S[v,v] (sym, 6).
// The adjoint congruence Ad(T) S Ad(T)^T uses the block structure
// Ad(T) = [[I,0],[skew(t),I]] diag(R,R): a rotation of the three blocks and a
// shear by skew(t), ~200 FMA instead of 342 for a dense symmetric product.
#pragma once

#include <hip/hip_runtime.h>

namespace rpgo {

struct P32 {
  float R[9];
  float t[3];
};

struct S6 {
  float A[6];  // a00 a01 a02 a11 a12 a22
  float B[9];  // row-major
  float C[6];  // c00 c01 c02 c11 c12 c22
};

__device__ __forceinline__ int s3(int i, int j) {
  // packed index of a symmetric 3x3 entry
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

__device__ __forceinline__ void p32_mul(const P32& a, const P32& b, P32& o) {
  P32 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.R[i * 3 + j] = a.R[i * 3] * b.R[j] + a.R[i * 3 + 1] * b.R[3 + j] + a.R[i * 3 + 2] * b.R[6 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    r.t[i] = a.t[i] + (a.R[i * 3] * b.t[0] + a.R[i * 3 + 1] * b.t[1] + a.R[i * 3 + 2] * b.t[2]);
  o = r;
}

__device__ __forceinline__ void p32_inv(const P32& a, P32& o) {
  P32 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.R[i * 3 + j] = a.R[j * 3 + i];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    r.t[i] = -(r.R[i * 3] * a.t[0] + r.R[i * 3 + 1] * a.t[1] + r.R[i * 3 + 2] * a.t[2]);
  o = r;
}

// a^-1 * b
__device__ __forceinline__ void p32_between(const P32& a, const P32& b, P32& o) {
  P32 ai;
  p32_inv(a, ai);
  p32_mul(ai, b, o);
}

__device__ __forceinline__ void s6_zero(S6& s) {
#pragma unroll
  for (int e = 0; e < 6; ++e) s.A[e] = 0.f, s.C[e] = 0.f;
#pragma unroll
  for (int e = 0; e < 9; ++e) s.B[e] = 0.f;
}

__device__ __forceinline__ float s6_trace(const S6& s) {
  return s.A[0] + s.A[3] + s.A[5] + s.C[0] + s.C[3] + s.C[5];
}

// R M R^T for symmetric packed M
__device__ __forceinline__ void rot_sym(const float* R, const float* M, float* o) {
  float RM[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      RM[i * 3 + j] = R[i * 3] * M[s3(0, j)] + R[i * 3 + 1] * M[s3(1, j)] + R[i * 3 + 2] * M[s3(2, j)];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j)
      o[s3(i, j)] = RM[i * 3] * R[j * 3] + RM[i * 3 + 1] * R[j * 3 + 1] + RM[i * 3 + 2] * R[j * 3 + 2];
}

// R M R^T for a general M
__device__ __forceinline__ void rot_gen(const float* R, const float* M, float* o) {
  float RM[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      RM[i * 3 + j] = R[i * 3] * M[j] + R[i * 3 + 1] * M[3 + j] + R[i * 3 + 2] * M[6 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      o[i * 3 + j] = RM[i * 3] * R[j * 3] + RM[i * 3 + 1] * R[j * 3 + 1] + RM[i * 3 + 2] * R[j * 3 + 2];
}

// R M R^T + Y for symmetric packed M and Y (the sums start from Y)
__device__ __forceinline__ void rot_sym_add(const float* R, const float* M, const float* Y, float* o) {
  float RM[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      RM[i * 3 + j] = R[i * 3] * M[s3(0, j)] + R[i * 3 + 1] * M[s3(1, j)] + R[i * 3 + 2] * M[s3(2, j)];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j)
      o[s3(i, j)] = __builtin_fmaf(RM[i * 3], R[j * 3],
                                   __builtin_fmaf(RM[i * 3 + 1], R[j * 3 + 1],
                                                  __builtin_fmaf(RM[i * 3 + 2], R[j * 3 + 2], Y[s3(i, j)])));
}

// o = Ad(T) S Ad(T)^T (+ Y when kAdd).  Ad(T) = [R 0; [t]x R R]: the rotated
// blocks A' = R A R^T, B' = R B R^T, C' = R C R^T, then the shear by [t]x:
//   o.A = A', o.B = A' X^T + B', o.C = X A' X^T + X B' + (X B')^T + C'.
// The X A' X^T terms fold into FMAs on top of the other summands of o.C; with
// kAdd, Y.C enters the sums of C' and Y.A, Y.B are added (one congruence plus
// an accumulation, 12 instructions fewer than the two apart).
template <bool kAdd>
__device__ __forceinline__ void s6_congruence_t(const P32& T, const S6& S, const S6* Y, S6& o) {
  float A[6], B[9], C[6];
  rot_sym(T.R, S.A, A);
  rot_gen(T.R, S.B, B);
  if (kAdd) rot_sym_add(T.R, S.C, Y->C, C);
  else rot_sym(T.R, S.C, C);
  const float tx = T.t[0], ty = T.t[1], tz = T.t[2];
  // XA = skew(t) * A  (column j = t x A[:,j]); P = A X^T = (XA)^T
  float XA[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float a0 = A[s3(0, j)], a1 = A[s3(1, j)], a2 = A[s3(2, j)];
    XA[0 * 3 + j] = ty * a2 - tz * a1;
    XA[1 * 3 + j] = tz * a0 - tx * a2;
    XA[2 * 3 + j] = tx * a1 - ty * a0;
  }
  // V = skew(t) * B
  float V[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float b0 = B[j], b1 = B[3 + j], b2 = B[6 + j];
    V[0 * 3 + j] = ty * b2 - tz * b1;
    V[1 * 3 + j] = tz * b0 - tx * b2;
    V[2 * 3 + j] = tx * b1 - ty * b0;
  }
  // out.B = A X^T + B = (XA)^T + B
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      o.B[i * 3 + j] = kAdd ? XA[j * 3 + i] + B[i * 3 + j] + Y->B[i * 3 + j] : XA[j * 3 + i] + B[i * 3 + j];
  // X A X^T = X * (XA)^T : column j = t x (XA)^T[:,j] = t x XA[j,:]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const float p0 = XA[j * 3], p1 = XA[j * 3 + 1], p2 = XA[j * 3 + 2];
      const float rest = V[i * 3 + j] + V[j * 3 + i] + C[s3(i, j)];
      float v;
      if (i == 0) v = __builtin_fmaf(ty, p2, __builtin_fmaf(-tz, p1, rest));
      else if (i == 1) v = __builtin_fmaf(tz, p0, __builtin_fmaf(-tx, p2, rest));
      else v = __builtin_fmaf(tx, p1, __builtin_fmaf(-ty, p0, rest));
      o.C[s3(i, j)] = v;
    }
#pragma unroll
  for (int e = 0; e < 6; ++e) o.A[e] = kAdd ? A[e] + Y->A[e] : A[e];
}

__device__ __forceinline__ void s6_congruence(const P32& T, const S6& S, S6& o) {
  s6_congruence_t<false>(T, S, nullptr, o);
}

// o = Ad(T) S Ad(T)^T + Y
__device__ __forceinline__ void s6_congruence_add(const P32& T, const S6& S, const S6& Y, S6& o) {
  s6_congruence_t<true>(T, S, &Y, o);
}

// atan2(s, c) for s >= 0: the octant reduction to atan(t), t in [0, 1], and
// the polynomial of Abramowitz & Stegun 4.4.49 (|error| <= 2e-8 on [0, 1]).
__device__ __forceinline__ float atan2_pos(float s, float c) {
  const float ac = fabsf(c);
  const bool steep = s > ac;
  const float num = steep ? ac : s, den = steep ? s : ac;
  const float t = den > 0.f ? num * __builtin_amdgcn_rcpf(den) : 0.f;
  const float t2 = t * t;
  float p = 0.0028662257f;
  p = __builtin_fmaf(p, t2, -0.0161657367f);
  p = __builtin_fmaf(p, t2, 0.0429096138f);
  p = __builtin_fmaf(p, t2, -0.0752896400f);
  p = __builtin_fmaf(p, t2, 0.1065626393f);
  p = __builtin_fmaf(p, t2, -0.1420889944f);
  p = __builtin_fmaf(p, t2, 0.1999355085f);
  p = __builtin_fmaf(p, t2, -0.3333314528f);
  const float a = __builtin_fmaf(t * t2, p, t);  // atan(t)
  const float o = steep ? 1.5707963267948966f - a : a;
  return c < 0.f ? 3.14159265358979f - o : o;
}

// SO(3) Log in FP32: theta from atan2(|axial|, (tr-1)/2); returns false when
// the angle is within ~3e-3 rad of pi (tr + 1 < 1e-5), where the axial vector
// 2 sin(theta) n no longer fixes the axis to FP32 precision.  Farther from pi
// the axis error is ~eps32 / sin(theta) < 2e-5, far inside the recheck band.
// Also returns theta and its sine and cosine.
//
// Square roots, reciprocals and inverse square roots in this file are the
// hardware's (v_sqrt_f32, v_rcp_f32, v_rsq_f32, about 1 ulp) rather than the
// correctly rounded library forms (12-18 instructions each): their error is
// orders of magnitude below the recheck band, and pairs near the threshold
// are decided by the FP64 replay.
__device__ __forceinline__ bool rot_log32(const float* R, float* w, float* th_out, float* s_out,
                                          float* c_out) {
  const float tr = R[0] + R[4] + R[8];
  const float ax = R[7] - R[5], ay = R[2] - R[6], az = R[3] - R[1];
  const float s = 0.5f * __builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az);  // sin(theta)
  const float c = 0.5f * (tr - 1.0f);                                          // cos(theta)
  const float th = atan2_pos(s, c);
  *th_out = th;
  *s_out = s;
  *c_out = c;
  float mag;
  if (th < 1e-3f) mag = 0.5f + th * th * (1.0f / 12.0f);
  else mag = 0.5f * th * __builtin_amdgcn_rcpf(s);
  w[0] = mag * ax;
  w[1] = mag * ay;
  w[2] = mag * az;
  return (tr + 1.0f) > 1e-5f;
}

// Pose3 Logmap in FP32 (same closed form as the reference's Pose3::Logmap)
__device__ __forceinline__ bool se3_log32(const P32& T, float* xi) {
  float w[3], th, s, c;
  const bool ok = rot_log32(T.R, w, &th, &s, &c);
  xi[0] = w[0]; xi[1] = w[1]; xi[2] = w[2];
  const float T0 = T.t[0], T1 = T.t[1], T2 = T.t[2];
  if (th < 1e-10f) {
    xi[3] = T0; xi[4] = T1; xi[5] = T2;
    return ok;
  }
  // u = T - 1/2 w x T + coef * w x (w x T), coef = (1 - (th/2) cot(th/2)) / th^2
  float coef;
  if (th < 1e-2f) coef = (1.0f / 12.0f) + th * th * (1.0f / 720.0f);
  else {
    // (th / 2) cot(th / 2) = (th / 2) (1 + c) / s = (th / 2) s / (1 - c): the
    // first where c >= 0, the second where c < 0, so neither divides by a
    // difference that cancels.
    const float h = c >= 0.f ? (1.0f + c) * __builtin_amdgcn_rcpf(s) : s * __builtin_amdgcn_rcpf(1.0f - c);
    coef = (1.0f - 0.5f * th * h) * __builtin_amdgcn_rcpf(th * th);
  }
  const float c0 = w[1] * T2 - w[2] * T1, c1 = w[2] * T0 - w[0] * T2, c2 = w[0] * T1 - w[1] * T0;
  const float d0 = w[1] * c2 - w[2] * c1, d1 = w[2] * c0 - w[0] * c2, d2 = w[0] * c1 - w[1] * c0;
  xi[3] = T0 - 0.5f * c0 + coef * d0;
  xi[4] = T1 - 0.5f * c1 + coef * d1;
  xi[5] = T2 - 0.5f * c2 + coef * d2;
  return ok;
}

// d^2 = xi^T S^-1 xi through an in-register Cholesky; min_piv gets the
// smallest pivot (L_kk^2).  Returns false when S is not numerically PD.
__device__ __forceinline__ bool mahal2_32(const S6& S, const float* xi, float* d2, float* min_piv) {
  float M[21];
  // full packed upper triangle of the 6x6, row-major
  auto at = [&](int i, int j) -> float {
    if (i > j) { const int t = i; i = j; j = t; }
    if (j < 3) return S.A[s3(i, j)];
    if (i >= 3) return S.C[s3(i - 3, j - 3)];
    return S.B[i * 3 + (j - 3)];
  };
  int e = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) M[e++] = at(i, j);
  float L[21];  // lower factor, row-major packed: L(i,j) j<=i at i(i+1)/2 + j
  float y[6];
  float mp = 3.0e38f;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float x = M[k * 6 - k * (k - 1) / 2];  // M(k,k)
#pragma unroll
    for (int j = 0; j < k; ++j) x -= L[k * (k + 1) / 2 + j] * L[k * (k + 1) / 2 + j];
    ok = ok && (x > 0.f);
    mp = fminf(mp, x);
    const float inv = __builtin_amdgcn_rsqf(fmaxf(x, 1e-30f));  // 1 / L_kk
    L[k * (k + 1) / 2 + k] = inv;
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      float v = M[k * 6 - k * (k - 1) / 2 + (i - k)];  // M(k,i) = M(i,k)
#pragma unroll
      for (int j = 0; j < k; ++j) v -= L[i * (i + 1) / 2 + j] * L[k * (k + 1) / 2 + j];
      L[i * (i + 1) / 2 + k] = v * inv;
    }
  }
  // forward solve L y = xi ; d2 = |y|^2
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float v = xi[i];
#pragma unroll
    for (int j = 0; j < i; ++j) v -= L[i * (i + 1) / 2 + j] * y[j];
    y[i] = v * L[i * (i + 1) / 2 + i];
    acc += y[i] * y[i];
  }
  *d2 = acc;
  *min_piv = mp;
  return ok;
}

}  // namespace rpgo
