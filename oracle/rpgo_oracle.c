/*
 * rpgo_oracle.c -- FP64 CPU restatement of Kimera-RPGO's PCM path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded by
 * tests/, smoke() and bench.py's cpu_baseline leg through oracle/oracle.py.
 *
 * Sources restated (paths relative to the reference tree):
 *   include/KimeraRPGO/utils/GeometryUtils.h:91-186   PoseWithCovariance
 *   include/KimeraRPGO/utils/GraphUtils.h:37-42       Trajectory::getBetween
 *   include/KimeraRPGO/outlier/Pcm.h:516-557          updateOdom (fold)
 *   include/KimeraRPGO/outlier/Pcm.h:604-629          isOdomConsistent
 *   include/KimeraRPGO/outlier/Pcm.h:670-718          areLoopsConsistent
 *   include/KimeraRPGO/outlier/Pcm.h:725-768          incrementAdjMatrix
 *   include/KimeraRPGO/utils/GeometryUtils.h:193-289  PoseWithNode
 *   include/KimeraRPGO/outlier/Pcm.h:581-596,654-662  PcmSimple checks
 *   include/KimeraRPGO/outlier/Pcm.h:460-465,691-698  inter-robot loop closures
 *   include/KimeraRPGO/max_clique_finder/graphIO.cpp:188-257
 *   include/KimeraRPGO/max_clique_finder/findCliqueHeu.cpp:32-209
 *   include/KimeraRPGO/max_clique_finder/findClique.cpp:31-145
 * GTSAM (commit 686e16aa, not vendored) semantics restated from its
 * documented algorithms: Pose3/Pose2 compose/between/inverse/AdjointMap,
 * LieGroup compose/between Jacobians (H1 = Ad(g^-1), H1 = -Ad(result^-1)),
 * SO3/Pose3/Pose2 Logmap.  Eigen semantics restated: unblocked LLT
 * (pivot <= 0 fails), partial-pivot LU inverse.
 *
 * Fixed operation order: every dot product is summed k = 0..n-1 left to
 * right; products never contract into FMA (built with -ffp-contract=off).
 */
#include "rpgo_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Threads used by the row-parallel distance loops (1 = the reference's own
 * single-threaded model; the arithmetic per pair is identical either way). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* small dense helpers (row-major)                                     */
/* ------------------------------------------------------------------ */

/* C = A * B, n x n */
static void mm(int n, const double* A, const double* B, double* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = A[i * n] * B[j];
      for (int k = 1; k < n; ++k) s += A[i * n + k] * B[k * n + j];
      C[i * n + j] = s;
    }
}

/* C = A * B^T, n x n */
static void mmt(int n, const double* A, const double* B, double* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = A[i * n] * B[j * n];
      for (int k = 1; k < n; ++k) s += A[i * n + k] * B[j * n + k];
      C[i * n + j] = s;
    }
}

/* out = (H * S) * H^T  -- the evaluation order of Ha*cov*Ha.transpose() */
static void congruence(int n, const double* H, const double* S, double* out) {
  double T[36];
  mm(n, H, S, T);
  mmt(n, T, H, out);
}

/* ------------------------------------------------------------------ */
/* SO(3) / SE(3)                                                      */
/* ------------------------------------------------------------------ */

static void rot_mul(const double* A, const double* B, double* C) { mm(3, A, B, C); }

static void rot_vec(const double* R, const double* v, double* out) {
  for (int i = 0; i < 3; ++i) {
    double s = R[i * 3] * v[0];
    s += R[i * 3 + 1] * v[1];
    s += R[i * 3 + 2] * v[2];
    out[i] = s;
  }
}

/* Pose3::operator*  : (R_a R_b, t_a + R_a t_b) */
void orc_se3_compose(const double* a, const double* b, double* out) {
  double R[9], Rt[3];
  rot_mul(a, b, R);
  rot_vec(a, b + 9, Rt);
  double t0 = a[9] + Rt[0], t1 = a[10] + Rt[1], t2 = a[11] + Rt[2];
  memcpy(out, R, sizeof(R));
  out[9] = t0; out[10] = t1; out[11] = t2;
}

/* Pose3::inverse : (R^T, R^T * (-t)) */
void orc_se3_inverse(const double* a, double* out) {
  double Rt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i * 3 + j] = a[j * 3 + i];
  double mt[3] = {-a[9], -a[10], -a[11]};
  double t[3];
  rot_vec(Rt, mt, t);
  memcpy(out, Rt, sizeof(Rt));
  out[9] = t[0]; out[10] = t[1]; out[11] = t[2];
}

/* LieGroup::between : inverse(a) * b */
void orc_se3_between(const double* a, const double* b, double* out) {
  double ai[12];
  orc_se3_inverse(a, ai);
  orc_se3_compose(ai, b, out);
}

/* Pose3::AdjointMap : [R 0; skew(t) R, R], tangent (omega, v).
 * skew(t)*R written without its structural zero products. */
void orc_se3_adjoint(const double* a, double* ad) {
  const double* R = a;
  const double tx = a[9], ty = a[10], tz = a[11];
  double A[9];
  for (int j = 0; j < 3; ++j) {
    A[0 * 3 + j] = -tz * R[3 + j] + ty * R[6 + j];
    A[1 * 3 + j] = tz * R[j] - tx * R[6 + j];
    A[2 * 3 + j] = -ty * R[j] + tx * R[3 + j];
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) ad[i * 6 + j] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ad[i * 6 + j] = R[i * 3 + j];
      ad[(i + 3) * 6 + j] = A[i * 3 + j];
      ad[(i + 3) * 6 + j + 3] = R[i * 3 + j];
    }
}

/* SO3::Logmap (GTSAM 4.1 form): special branch at trace ~ -1, Taylor
 * magnitude 0.5 - (tr-3)/12 near the identity. */
void orc_rot3_logmap(const double* R, double* w) {
  const double R11 = R[0], R12 = R[1], R13 = R[2];
  const double R21 = R[3], R22 = R[4], R23 = R[5];
  const double R31 = R[6], R32 = R[7], R33 = R[8];
  const double tr = (R11 + R22) + R33;
  if (tr + 1.0 < 1e-10) {
    if (fabs(R33 + 1.0) > 1e-5) {
      const double f = M_PI / sqrt(2.0 + 2.0 * R33);
      w[0] = f * R13; w[1] = f * R23; w[2] = f * (1.0 + R33);
    } else if (fabs(R22 + 1.0) > 1e-5) {
      const double f = M_PI / sqrt(2.0 + 2.0 * R22);
      w[0] = f * R12; w[1] = f * (1.0 + R22); w[2] = f * R32;
    } else {
      const double f = M_PI / sqrt(2.0 + 2.0 * R11);
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

/* Pose3::Logmap (Agrawal06iros eq. 14 form used by GTSAM). */
void orc_se3_logmap(const double* a, double* xi) {
  double w[3];
  orc_rot3_logmap(a, w);
  const double T0 = a[9], T1 = a[10], T2 = a[11];
  const double t = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  xi[0] = w[0]; xi[1] = w[1]; xi[2] = w[2];
  if (t < 1e-10) {
    xi[3] = T0; xi[4] = T1; xi[5] = T2;
    return;
  }
  const double n0 = w[0] / t, n1 = w[1] / t, n2 = w[2] / t;
  const double Tan = tan(0.5 * t);
  /* WT = skew(n) * T */
  const double WT0 = -n2 * T1 + n1 * T2;
  const double WT1 = n2 * T0 - n0 * T2;
  const double WT2 = -n1 * T0 + n0 * T1;
  /* WWT = skew(n) * WT */
  const double WWT0 = -n2 * WT1 + n1 * WT2;
  const double WWT1 = n2 * WT0 - n0 * WT2;
  const double WWT2 = -n1 * WT0 + n0 * WT1;
  const double ht = 0.5 * t;
  const double c = 1.0 - t / (2.0 * Tan);
  xi[3] = (T0 - ht * WT0) + c * WWT0;
  xi[4] = (T1 - ht * WT1) + c * WWT1;
  xi[5] = (T2 - ht * WT2) + c * WWT2;
}

/* Rodrigues exponential and Pose3::Expmap (used by tests and generators). */
static void rot3_expmap(const double* w, double* R) {
  const double th2 = (w[0] * w[0] + w[1] * w[1]) + w[2] * w[2];
  const double th = sqrt(th2);
  double A, B;
  if (th2 < 1e-16) { A = 1.0 - th2 / 6.0; B = 0.5 - th2 / 24.0; }
  else { A = sin(th) / th; B = (1.0 - cos(th)) / th2; }
  const double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double K2[9];
  mm(3, K, K, K2);
  for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0 ? 1.0 : 0.0) + A * K[i] + B * K2[i];
}

void orc_se3_expmap(const double* xi, double* out) {
  const double* w = xi;
  const double* v = xi + 3;
  double R[9];
  rot3_expmap(w, R);
  const double th2 = (w[0] * w[0] + w[1] * w[1]) + w[2] * w[2];
  memcpy(out, R, sizeof(R));
  if (th2 > 2.220446049250313e-16) {
    const double wv = (w[0] * v[0] + w[1] * v[1]) + w[2] * v[2];
    const double wxv[3] = {w[1] * v[2] - w[2] * v[1], w[2] * v[0] - w[0] * v[2],
                           w[0] * v[1] - w[1] * v[0]};
    double Rwxv[3];
    rot_vec(R, wxv, Rwxv);
    for (int i = 0; i < 3; ++i) out[9 + i] = ((wxv[i] - Rwxv[i]) + w[i] * wv) / th2;
  } else {
    out[9] = v[0]; out[10] = v[1]; out[11] = v[2];
  }
}

/* Unit quaternion -> rotation matrix (standard formula, w first). */
void orc_quat_to_rot(double w, double x, double y, double z, double* R) {
  const double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x;
  const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1.0 - (tyy + tzz); R[1] = txy - twz; R[2] = txz + twy;
  R[3] = txy + twz; R[4] = 1.0 - (txx + tzz); R[5] = tyz - twx;
  R[6] = txz - twy; R[7] = tyz + twx; R[8] = 1.0 - (txx + tyy);
}

/* ------------------------------------------------------------------ */
/* SO(2) / SE(2)  (pose = c, s, x, y; tangent (x, y, theta))          */
/* ------------------------------------------------------------------ */

/* Rot2::fromCosSin with normalize() when |c^2+s^2-1| > 1e-10 */
static void rot2_from_cos_sin(double c, double s, double* oc, double* os) {
  double scale = c * c + s * s;
  if (fabs(scale - 1.0) > 1e-10) {
    scale = 1.0 / sqrt(scale);
    c *= scale;
    s *= scale;
  }
  *oc = c; *os = s;
}

/* Pose2::operator* : (r_a r_b, t_a + r_a.rotate(t_b)) */
void orc_se2_compose(const double* a, const double* b, double* out) {
  double c, s;
  rot2_from_cos_sin(a[0] * b[0] - a[1] * b[1], a[1] * b[0] + a[0] * b[1], &c, &s);
  const double rx = a[0] * b[2] - a[1] * b[3];
  const double ry = a[1] * b[2] + a[0] * b[3];
  out[0] = c; out[1] = s;
  out[2] = a[2] + rx; out[3] = a[3] + ry;
}

/* Pose2::inverse : (r^-1, r.unrotate(-t)) */
void orc_se2_inverse(const double* a, double* out) {
  const double c = a[0], s = a[1];
  const double px = -a[2], py = -a[3];
  out[0] = c; out[1] = -s;
  out[2] = c * px + s * py;
  out[3] = -s * px + c * py;
}

void orc_se2_between(const double* a, const double* b, double* out) {
  double ai[4];
  orc_se2_inverse(a, ai);
  orc_se2_compose(ai, b, out);
}

/* Pose2::AdjointMap : [c -s y; s c -x; 0 0 1] */
void orc_se2_adjoint(const double* a, double* ad) {
  ad[0] = a[0]; ad[1] = -a[1]; ad[2] = a[3];
  ad[3] = a[1]; ad[4] = a[0]; ad[5] = -a[2];
  ad[6] = 0.0; ad[7] = 0.0; ad[8] = 1.0;
}

/* Pose2::Logmap */
void orc_se2_logmap(const double* a, double* xi) {
  const double c = a[0], s = a[1], tx = a[2], ty = a[3];
  const double w = atan2(s, c);
  if (fabs(w) < 1e-10) {
    xi[0] = tx; xi[1] = ty; xi[2] = w;
    return;
  }
  const double c_1 = c - 1.0;
  const double det = c_1 * c_1 + s * s;
  /* R.unrotate(t) - t */
  const double ux = (c * tx + s * ty) - tx;
  const double uy = (-s * tx + c * ty) - ty;
  /* R_PI_2 * u = (-uy, ux) */
  const double px = -uy, py = ux;
  const double f = w / det;
  xi[0] = f * px; xi[1] = f * py; xi[2] = w;
}

/* ------------------------------------------------------------------ */
/* Eigen semantics: LLT success test, partial-pivot LU inverse         */
/* ------------------------------------------------------------------ */

int orc_llt_ok(int n, const double* A) {
  double L[36];
  for (int k = 0; k < n; ++k) {
    double x = A[k * n + k];
    if (k > 0) {
      double sq = L[k * n] * L[k * n];
      for (int j = 1; j < k; ++j) sq += L[k * n + j] * L[k * n + j];
      x -= sq;
    }
    if (x <= 0.0) return 0;
    x = sqrt(x);
    L[k * n + k] = x;
    for (int i = k + 1; i < n; ++i) {
      double v = A[i * n + k];
      if (k > 0) {
        double dot = L[i * n] * L[k * n];
        for (int j = 1; j < k; ++j) dot += L[i * n + j] * L[k * n + j];
        v -= dot;
      }
      L[i * n + k] = v / x;
    }
  }
  return 1;
}

void orc_lu_inverse(int n, const double* A, double* Ainv) {
  double lu[36];
  int perm[6];
  memcpy(lu, A, sizeof(double) * n * n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(lu[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(lu[i * n + k]);
      if (v > best) { best = v; p = i; }
    }
    perm[k] = p;
    if (best != 0.0) {
      if (p != k)
        for (int j = 0; j < n; ++j) {
          const double t = lu[k * n + j];
          lu[k * n + j] = lu[p * n + j];
          lu[p * n + j] = t;
        }
      const double piv = lu[k * n + k];
      for (int i = k + 1; i < n; ++i) lu[i * n + k] /= piv;
    }
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
  }
  for (int col = 0; col < n; ++col) {
    double b[6];
    for (int i = 0; i < n; ++i) b[i] = (i == col) ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k)
      if (perm[k] != k) {
        const double t = b[k];
        b[k] = b[perm[k]];
        b[perm[k]] = t;
      }
    for (int i = 1; i < n; ++i) {
      double s = lu[i * n] * b[0];
      for (int k = 1; k < i; ++k) s += lu[i * n + k] * b[k];
      b[i] -= s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double v = b[i];
      if (i < n - 1) {
        double s = lu[i * n + i + 1] * b[i + 1];
        for (int k = i + 2; k < n; ++k) s += lu[i * n + k] * b[k];
        v -= s;
      }
      b[i] = v / lu[i * n + i];
    }
    for (int i = 0; i < n; ++i) Ainv[i * n + col] = b[i];
  }
}

/* ------------------------------------------------------------------ */
/* PoseWithCovariance (GeometryUtils.h:56-187)                         */
/* ------------------------------------------------------------------ */

static int pose_len(int dim) { return dim == 3 ? 12 : 4; }
static int tan_dim(int dim) { return dim == 3 ? 6 : 3; }
int orc_pwc_size(int dim) {
  const int d = tan_dim(dim);
  return pose_len(dim) + d * d + 1;
}

static void pose_compose(int dim, const double* a, const double* b, double* o) {
  if (dim == 3) orc_se3_compose(a, b, o); else orc_se2_compose(a, b, o);
}
static void pose_inverse(int dim, const double* a, double* o) {
  if (dim == 3) orc_se3_inverse(a, o); else orc_se2_inverse(a, o);
}
static void pose_adjoint(int dim, const double* a, double* o) {
  if (dim == 3) orc_se3_adjoint(a, o); else orc_se2_adjoint(a, o);
}
static void pose_logmap(int dim, const double* a, double* o) {
  if (dim == 3) orc_se3_logmap(a, o); else orc_se2_logmap(a, o);
}

/* PoseWithCovariance(const BetweenFactor&) : GeometryUtils.h:91-115 */
void orc_pwc_from_measurement(int dim, const double* pose, const double* cov, double* pwc) {
  const int pl = pose_len(dim), d = tan_dim(dim);
  memcpy(pwc, pose, sizeof(double) * pl);
  double* C = pwc + pl;
  memcpy(C, cov, sizeof(double) * d * d);
  const int rd = dim == 3 ? 3 : 1;
  double tr = cov[0];
  for (int i = 1; i < rd; ++i) tr += cov[i * d + i];
  double ri = 1.0;
  if (isnan(tr)) {
    ri = 0.0;
    for (int i = 0; i < d * d; ++i) C[i] = 0.0;
    for (int i = rd; i < d; ++i)
      for (int j = rd; j < d; ++j) C[i * d + j] = cov[i * d + j];
  }
  pwc[pl + d * d] = ri;
}

/* compose : GeometryUtils.h:119-129.  H1 = other.pose.inverse().AdjointMap(), H2 = I */
void orc_pwc_compose(int dim, const double* a, const double* b, double* out) {
  const int pl = pose_len(dim), d = tan_dim(dim);
  double pose[12], binv[12], H[36], HSH[36];
  pose_compose(dim, a, b, pose);
  pose_inverse(dim, b, binv);
  pose_adjoint(dim, binv, H);
  congruence(d, H, a + pl, HSH);
  memcpy(out, pose, sizeof(double) * pl);
  for (int i = 0; i < d * d; ++i) out[pl + i] = HSH[i] + b[pl + i];
  out[pl + d * d] = (a[pl + d * d] != 0.0 && b[pl + d * d] != 0.0) ? 1.0 : 0.0;
}

/* inverse : GeometryUtils.h:133-139 (covariance untransported, Q4) */
void orc_pwc_inverse(int dim, const double* a, double* out) {
  const int pl = pose_len(dim), d = tan_dim(dim);
  double pose[12];
  pose_inverse(dim, a, pose);
  memcpy(out, pose, sizeof(double) * pl);
  memcpy(out + pl, a + pl, sizeof(double) * d * d);
  out[pl + d * d] = a[pl + d * d];
}

/* between : GeometryUtils.h:143-170, including the LLT fallback (Q5). */
void orc_pwc_between(int dim, const double* a, const double* b, double* out) {
  const int pl = pose_len(dim), d = tan_dim(dim);
  double pose[12], rinv[12], H[36], HSH[36], C[36];
  double ainv[12];
  pose_inverse(dim, a, ainv);
  pose_compose(dim, ainv, b, pose);          /* result = a^-1 b   */
  pose_inverse(dim, pose, rinv);             /* H1 = -Ad(result^-1) */
  pose_adjoint(dim, rinv, H);
  congruence(d, H, a + pl, HSH);
  for (int i = 0; i < d * d; ++i) C[i] = b[pl + i] - HSH[i];
  if (!orc_llt_ok(d, C)) {
    double r2[12], binv[12], r2inv[12];
    pose_inverse(dim, b, binv);
    pose_compose(dim, binv, a, r2);          /* other.between(this) */
    pose_inverse(dim, r2, r2inv);
    pose_adjoint(dim, r2inv, H);
    congruence(d, H, b + pl, HSH);
    for (int i = 0; i < d * d; ++i) C[i] = a[pl + i] - HSH[i];
  }
  memcpy(out, pose, sizeof(double) * pl);
  memcpy(out + pl, C, sizeof(double) * d * d);
  out[pl + d * d] = (a[pl + d * d] != 0.0 && b[pl + d * d] != 0.0) ? 1.0 : 0.0;
}

/* mahalanobis_norm : GeometryUtils.h:172-186 (Q7 explicit inverse, Q8) */
double orc_pwc_mahalanobis(int dim, const double* a) {
  const int pl = pose_len(dim), d = tan_dim(dim);
  double xi[6];
  pose_logmap(dim, a, xi);
  const double* C = a + pl;
  int off = 0, n = d;
  double M[36], Minv[36];
  if (a[pl + d * d] == 0.0) {
    const int rd = dim == 3 ? 3 : 1;
    off = rd;
    n = d - rd;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) M[i * n + j] = C[(i + off) * d + (j + off)];
  orc_lu_inverse(n, M, Minv);
  const double* v = xi + off;
  /* (v^T * inv) * v */
  double q = 0.0;
  for (int j = 0; j < n; ++j) {
    double r = v[0] * Minv[j];
    for (int k = 1; k < n; ++k) r += v[k] * Minv[k * n + j];
    if (j == 0) q = r * v[0]; else q += r * v[j];
  }
  return sqrt(q);
}

/* ------------------------------------------------------------------ */
/* PCM pipeline                                                        */
/* ------------------------------------------------------------------ */

/* updateOdom fold : Pcm.h:516-557 (first pose from the values, zero cov) */
void orc_fold(int dim, int n_poses, const double* pose0, const double* odom,
              const double* odom_cov, double* traj) {
  const int pl = pose_len(dim), d = tan_dim(dim), ps = orc_pwc_size(dim);
  if (n_poses <= 0) return;
  memcpy(traj, pose0, sizeof(double) * pl);
  for (int i = 0; i < d * d; ++i) traj[pl + i] = 0.0;
  traj[pl + d * d] = 1.0;
  double delta[64];
  for (int k = 0; k + 1 < n_poses; ++k) {
    orc_pwc_from_measurement(dim, odom + (size_t)k * pl, odom_cov + (size_t)k * d * d, delta);
    orc_pwc_compose(dim, traj + (size_t)k * ps, delta, traj + (size_t)(k + 1) * ps);
  }
}

/* isOdomConsistent : Pcm.h:604-629 */
double orc_odom_dist(int dim, const double* traj, int from, int to,
                     const double* lc_pose, const double* lc_cov) {
  const int ps = orc_pwc_size(dim);
  double pij[64], lc[64], pji[64], res[64];
  orc_pwc_between(dim, traj + (size_t)from * ps, traj + (size_t)to * ps, pij);
  orc_pwc_from_measurement(dim, lc_pose, lc_cov, lc);
  orc_pwc_inverse(dim, lc, pji);
  orc_pwc_compose(dim, pij, pji, res);
  return orc_pwc_mahalanobis(dim, res);
}

/* areLoopsConsistent : Pcm.h:670-718 (single prefix, no key swap) */
double orc_pair_dist(int dim, const double* traj,
                     int a, int b, const double* zab, const double* sab,
                     int c, int d_, const double* zcd, const double* scd) {
  const int ps = orc_pwc_size(dim);
  double alcb[64], cld[64], A[64], B[64], X[64], Y[64], W[64], L[64];
  orc_pwc_from_measurement(dim, zab, sab, alcb);
  orc_pwc_from_measurement(dim, zcd, scd, cld);
  orc_pwc_between(dim, traj + (size_t)a * ps, traj + (size_t)c * ps, A);
  orc_pwc_between(dim, traj + (size_t)b * ps, traj + (size_t)d_ * ps, B);
  orc_pwc_compose(dim, A, cld, X);
  orc_pwc_inverse(dim, X, Y);
  orc_pwc_compose(dim, Y, alcb, W);
  orc_pwc_compose(dim, W, B, L);
  return orc_pwc_mahalanobis(dim, L);
}

int orc_pcm_adjacency(int dim, int n_poses, const double* traj,
                      int n_lc, const int32_t* lc_from, const int32_t* lc_to,
                      const double* lc_pose, const double* lc_cov,
                      double odom_thr, double lc_thr,
                      int32_t* accepted, double* odom_dist_out,
                      uint8_t* adj, double* dist) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
  (void)n_poses;
  int m = 0;
  for (int k = 0; k < n_lc; ++k) {
    /* the distance is reported also when the check is off (odom_thr < 0,
     * every loop closure passes: Pcm.h:606) */
    const double od = orc_odom_dist(dim, traj, lc_from[k], lc_to[k], lc_pose + (size_t)k * pl,
                                    lc_cov + (size_t)k * dd);
    const int ok = odom_thr < 0.0 || od < odom_thr;
    if (odom_dist_out) odom_dist_out[k] = od;
    if (ok) accepted[m++] = k;
  }
  if (adj) memset(adj, 0, (size_t)m * m);
  if (dist) memset(dist, 0, sizeof(double) * (size_t)m * m);
  if (lc_thr < 0.0) return m;
  for (int j = 1; j < m; ++j) {
    const int kj = accepted[j];
    for (int i = 0; i < j; ++i) {
      const int ki = accepted[i];
      const double dv = orc_pair_dist(dim, traj, lc_from[ki], lc_to[ki],
                                      lc_pose + (size_t)ki * pl, lc_cov + (size_t)ki * dd,
                                      lc_from[kj], lc_to[kj],
                                      lc_pose + (size_t)kj * pl, lc_cov + (size_t)kj * dd);
      if (dist) {
        dist[(size_t)j * m + i] = dv;
        dist[(size_t)i * m + j] = dv;
      }
      if (adj && dv < lc_thr) {
        adj[(size_t)j * m + i] = 1;
        adj[(size_t)i * m + j] = 1;
      }
    }
  }
  return m;
}

/* ------------------------------------------------------------------ */
/* Several robot trajectories (Pcm.h:516-557 per prefix, 414-502, 670-718) */
/* ------------------------------------------------------------------ */

static int chain_of(int n_chains, const int32_t* chain_off, int k) {
  int lo = 0, hi = n_chains - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (chain_off[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

/* Trajectory::getBetween across two prefixes (GraphUtils.h:43-57): with
 * root poses a0 (of key a's robot) and b0 (of key b's robot), result =
 * between(a0, a)^-1 o between(a0, b0) o between(b0, b).  The records are
 * PoseWithCovariance (pose, covariance, rotation_info). */
void orc_cross_between(int dim, const double* a0, const double* a, const double* b0, const double* b,
                       double* out) {
  double pa[64], pb[64], pab[64], ia[64], t[64];
  orc_pwc_between(dim, a0, a, pa);
  orc_pwc_between(dim, b0, b, pb);
  orc_pwc_between(dim, a0, b0, pab);
  orc_pwc_inverse(dim, pa, ia);
  orc_pwc_compose(dim, ia, pab, t);
  orc_pwc_compose(dim, t, pb, out);
}

/* Fold of every chain: chain c covers poses [chain_off[c], chain_off[c+1]),
 * starts at pose0s[c] with zero covariance (Pcm.h:534-542) and uses odometry
 * rows [chain_off[c] - c, chain_off[c+1] - c - 1). */
void orc_fold_chains(int dim, int n_chains, const int32_t* chain_off, const double* pose0s,
                     const double* odom, const double* odom_cov, double* traj) {
  const int pl = pose_len(dim), d = tan_dim(dim), ps = orc_pwc_size(dim);
  for (int c = 0; c < n_chains; ++c) {
    const int p0 = chain_off[c], n = chain_off[c + 1] - p0, o0 = p0 - c;
    orc_fold(dim, n, pose0s + (size_t)c * pl, odom + (size_t)o0 * pl, odom_cov + (size_t)o0 * d * d,
             traj + (size_t)p0 * ps);
  }
}

/* Adjacency stage of one loop-closure group over several chains:
 *  - a loop closure whose endpoints lie on different chains (an inter-robot
 *    loop closure) passes the odometry check unchecked, distance 0
 *    (Pcm.h:460-465);
 *  - for a pair, when the older loop closure's first endpoint and the newer
 *    one's lie on different chains, the newer one's endpoints are swapped
 *    without inverting its measurement (Pcm.h:691-698).
 * simple != 0 uses the PcmSimple distances (thresholds odom_thr/odom_thr2 and
 * lc_thr/lc_thr2 as translation/rotation); otherwise the Mahalanobis ones
 * (odom_thr, lc_thr). */
int orc_pcm_adjacency_chains(int dim, int n_chains, const int32_t* chain_off, const double* traj,
                             int n_lc, const int32_t* lc_from, const int32_t* lc_to,
                             const double* lc_pose, const double* lc_cov, int simple,
                             double odom_thr, double odom_thr2, double lc_thr, double lc_thr2,
                             int32_t* accepted, double* odom_dist_out, uint8_t* adj, double* dist) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
  const int odom_check = simple ? (odom_thr >= 0.0 && odom_thr2 >= 0.0) : odom_thr >= 0.0;
  const int loop_check = simple ? (lc_thr >= 0.0 && lc_thr2 >= 0.0) : lc_thr >= 0.0;
  int m = 0;
  for (int k = 0; k < n_lc; ++k) {
    const int a = lc_from[k], b = lc_to[k];
    double od = 0.0, rot = 0.0;
    int ok = 1;
    if (chain_of(n_chains, chain_off, a) == chain_of(n_chains, chain_off, b)) {
      if (simple) {
        od = orc_simple_odom(dim, traj, a, b, lc_pose + (size_t)k * pl, lc_cov + (size_t)k * dd, &rot);
        ok = !odom_check || (od < odom_thr && rot < odom_thr2);
      } else {
        od = orc_odom_dist(dim, traj, a, b, lc_pose + (size_t)k * pl, lc_cov + (size_t)k * dd);
        ok = !odom_check || od < odom_thr;
      }
    }
    if (odom_dist_out) odom_dist_out[k] = od;
    if (ok) accepted[m++] = k;
  }
  if (adj) memset(adj, 0, (size_t)m * m);
  if (dist) memset(dist, 0, sizeof(double) * (size_t)m * m);
  if (!loop_check) return m;
  for (int j = 1; j < m; ++j) {
    const int kj = accepted[j];
    for (int i = 0; i < j; ++i) {
      const int ki = accepted[i];
      const int a = lc_from[ki], b = lc_to[ki];
      int c = lc_from[kj], d_ = lc_to[kj];
      if (chain_of(n_chains, chain_off, a) != chain_of(n_chains, chain_off, c)) {
        const int t = c;
        c = d_;
        d_ = t;
      }
      double dv, rot = 0.0;
      int cons;
      if (simple) {
        dv = orc_simple_pair(dim, traj, a, b, lc_pose + (size_t)ki * pl, lc_cov + (size_t)ki * dd, c, d_,
                             lc_pose + (size_t)kj * pl, lc_cov + (size_t)kj * dd, &rot);
        cons = dv < lc_thr && rot < lc_thr2;
      } else {
        dv = orc_pair_dist(dim, traj, a, b, lc_pose + (size_t)ki * pl, lc_cov + (size_t)ki * dd, c, d_,
                           lc_pose + (size_t)kj * pl, lc_cov + (size_t)kj * dd);
        cons = dv < lc_thr;
      }
      if (dist) {
        dist[(size_t)j * m + i] = dv;
        dist[(size_t)i * m + j] = dv;
      }
      if (adj && cons) {
        adj[(size_t)j * m + i] = 1;
        adj[(size_t)i * m + j] = 1;
      }
    }
  }
  return m;
}

void orc_pcm_dist_rows(int dim, const double* traj, int m,
                       const int32_t* lc_from, const int32_t* lc_to,
                       const double* lc_pose, const double* lc_cov,
                       int j0, int j1, double* dist) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = j0; j < j1; ++j)
    for (int i = 0; i < j; ++i) {
      const double dv = orc_pair_dist(dim, traj, lc_from[i], lc_to[i],
                                      lc_pose + (size_t)i * pl, lc_cov + (size_t)i * dd,
                                      lc_from[j], lc_to[j],
                                      lc_pose + (size_t)j * pl, lc_cov + (size_t)j * dd);
      dist[(size_t)j * m + i] = dv;
      dist[(size_t)i * m + j] = dv;
    }
}

/* incrementAdjMatrix's storage pattern (Pcm.h:737-767), one loop closure at
 * a time, for the CPU baseline's timing: zero-filled k x k adjacency and
 * distance matrices (column-major doubles, Eigen::MatrixXd::Zero), the
 * previous (k-1) x (k-1) matrices copied into their top-left corner, row and
 * column k-1 filled from areLoopsConsistent against every earlier loop closure
 * (:752-763), then assigned to the stored matrices (:765-766, a second copy of
 * k x k doubles each).  Single-threaded, like the reference.  do_pairs = 0
 * moves the same bytes without the pair checks (every pair consistent, as
 * with lc_threshold < 0, :673).  Returns the consistent pairs (i < j). */
long long orc_adj_growth(int dim, const double* traj, int m, const int32_t* lc_from,
                         const int32_t* lc_to, const double* lc_pose, const double* lc_cov,
                         double lc_thr, int do_pairs) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
  double* adj = NULL;  /* stored matrices, (k-1) x (k-1) before step k */
  double* dst = NULL;
  long long n_cons = 0;
  for (int k = 1; k <= m; ++k) {
    const size_t kk = (size_t)k * k;
    double* nadj = (double*)calloc(kk, sizeof(double));
    double* ndst = (double*)calloc(kk, sizeof(double));
    if (!nadj || !ndst) {
      free(nadj); free(ndst); free(adj); free(dst);
      return -1;
    }
    if (k > 1) {
      for (int c = 0; c < k - 1; ++c) {  /* topLeftCorner(k-1, k-1) = old, column by column */
        memcpy(nadj + (size_t)c * k, adj + (size_t)c * (k - 1), sizeof(double) * (k - 1));
        memcpy(ndst + (size_t)c * k, dst + (size_t)c * (k - 1), sizeof(double) * (k - 1));
      }
      const int j = k - 1;
      for (int i = 0; i < j; ++i) {
        double dv = 0.0;
        int ok = 1;
        if (do_pairs && lc_thr >= 0.0) {
          dv = orc_pair_dist(dim, traj, lc_from[i], lc_to[i], lc_pose + (size_t)i * pl,
                             lc_cov + (size_t)i * dd, lc_from[j], lc_to[j],
                             lc_pose + (size_t)j * pl, lc_cov + (size_t)j * dd);
          ok = dv < lc_thr;
        }
        ndst[(size_t)i * k + j] = dv;  /* (row j, col i) and (row i, col j) */
        ndst[(size_t)j * k + i] = dv;
        if (ok) {
          nadj[(size_t)i * k + j] = 1.0;
          nadj[(size_t)j * k + i] = 1.0;
          ++n_cons;
        }
      }
    }
    /* adj_matrix = new_adj_matrix: the stored matrix is reallocated and copied */
    free(adj);
    free(dst);
    adj = (double*)malloc(sizeof(double) * kk);
    dst = (double*)malloc(sizeof(double) * kk);
    if (!adj || !dst) {
      free(nadj); free(ndst); free(adj); free(dst);
      return -1;
    }
    memcpy(adj, nadj, sizeof(double) * kk);
    memcpy(dst, ndst, sizeof(double) * kk);
    free(nadj);
    free(ndst);
  }
  free(adj);
  free(dst);
  return n_cons;
}

/* ------------------------------------------------------------------ */
/* PcmSimple (PoseWithNode, GeometryUtils.h:193-289)                    */
/* ------------------------------------------------------------------ */

/* avg_trans_norm / avg_rot_norm : GeometryUtils.h:275-289.  The norms take
 * tail(translation size) and head(rotation dimension) of T::Logmap: for
 * Pose3 (omega, v) that is |v| and |omega|; for Pose2 (v_x, v_y, theta) it is
 * |(v_y, theta)| and |v_x| (translation size 2, Rot2 dimension 1). */
void orc_pwn_norms(int dim, const double* pose, int node, int rot_info, double* trans,
                   double* rot) {
  const int d = tan_dim(dim), td = dim == 3 ? 3 : 2, rd = dim == 3 ? 3 : 1;
  double xi[6];
  pose_logmap(dim, pose, xi);
  double st = xi[d - td] * xi[d - td];
  for (int k = d - td + 1; k < d; ++k) st += xi[k] * xi[k];
  *trans = sqrt(st) / node;
  if (!rot_info) {
    *rot = 0.0;
    return;
  }
  double sr = xi[0] * xi[0];
  for (int k = 1; k < rd; ++k) sr += xi[k] * xi[k];
  *rot = sqrt(sr) / node;
}

/* Trajectory records come from orc_fold: the pose chain and rotation_info
 * flags of PoseWithNode equal those of PoseWithCovariance (compose / between
 * / inverse move the pose the same way and AND the flags the same way), and
 * the node count of chain pose k is k (Pcm.h:534-556: node 0 at the chain
 * start, 1 per odometry factor). */
static int pwn_flag(int dim, const double* rec) {
  const int d = tan_dim(dim);
  return rec[pose_len(dim) + d * d] != 0.0;
}

/* isOdomConsistent with checkOdomConsistent(PoseWithNode) : Pcm.h:581-596,
 * 604-629.  result = getBetween(i, j).compose(LC.inverse()), node |j-i| + 1. */
double orc_simple_odom(int dim, const double* traj, int from, int to, const double* lc_pose,
                       const double* lc_cov, double* rot_out) {
  const int ps = orc_pwc_size(dim);
  double lc[64], ainv[12], pij[12], lcinv[12], res[12];
  orc_pwc_from_measurement(dim, lc_pose, lc_cov, lc);
  const double* Pi = traj + (size_t)from * ps;
  const double* Pj = traj + (size_t)to * ps;
  pose_inverse(dim, Pi, ainv);
  pose_compose(dim, ainv, Pj, pij);
  pose_inverse(dim, lc, lcinv);
  pose_compose(dim, pij, lcinv, res);
  const int node = abs(to - from) + 1;
  const int ri = pwn_flag(dim, Pi) && pwn_flag(dim, Pj) && pwn_flag(dim, lc);
  double t, r;
  orc_pwn_norms(dim, res, node, ri, &t, &r);
  if (rot_out) *rot_out = r;
  return t;
}

/* areLoopsConsistent with checkLoopConsistent(PoseWithNode) : Pcm.h:654-662,
 * 670-718; node |c-a| + 1 + 1 + |d-b|. */
double orc_simple_pair(int dim, const double* traj, int a, int b, const double* zab,
                       const double* sab, int c, int d_, const double* zcd, const double* scd,
                       double* rot_out) {
  const int ps = orc_pwc_size(dim);
  double alcb[64], cld[64], inv[12], A[12], B[12], X[12], Y[12], W[12], L[12];
  orc_pwc_from_measurement(dim, zab, sab, alcb);
  orc_pwc_from_measurement(dim, zcd, scd, cld);
  const double* Pa = traj + (size_t)a * ps;
  const double* Pb = traj + (size_t)b * ps;
  const double* Pc = traj + (size_t)c * ps;
  const double* Pd = traj + (size_t)d_ * ps;
  pose_inverse(dim, Pa, inv);
  pose_compose(dim, inv, Pc, A);
  pose_compose(dim, A, cld, X);
  pose_inverse(dim, X, Y);
  pose_compose(dim, Y, alcb, W);
  pose_inverse(dim, Pb, inv);
  pose_compose(dim, inv, Pd, B);
  pose_compose(dim, W, B, L);
  const int node = abs(c - a) + 2 + abs(d_ - b);
  const int ri = pwn_flag(dim, Pa) && pwn_flag(dim, Pc) && pwn_flag(dim, cld) &&
                 pwn_flag(dim, alcb) && pwn_flag(dim, Pb) && pwn_flag(dim, Pd);
  double t, r;
  orc_pwn_norms(dim, L, node, ri, &t, &r);
  if (rot_out) *rot_out = r;
  return t;
}

/* ------------------------------------------------------------------ */
/* Landmark observations (Pcm.h:775-844): loop closures (i, l) and (j, l) */
/* from poses i, j of one trajectory to one landmark l.                   */
/* ------------------------------------------------------------------ */

/* incrementLandmarkAdjMatrix, PoseWithCovariance: loop = (between(i, j) o
 * Z_jl)^-1 o Z_il, Mahalanobis norm (Pcm.h:818-833) */
double orc_landmark_pair_dist(int dim, const double* traj, int i, const double* zil, const double* sil,
                              int j, const double* zjl, const double* sjl) {
  const int ps = orc_pwc_size(dim);
  double il[64], jl[64], A[64], X[64], Y[64], L[64];
  orc_pwc_from_measurement(dim, zil, sil, il);
  orc_pwc_from_measurement(dim, zjl, sjl, jl);
  orc_pwc_between(dim, traj + (size_t)i * ps, traj + (size_t)j * ps, A);
  orc_pwc_compose(dim, A, jl, X);
  orc_pwc_inverse(dim, X, Y);
  orc_pwc_compose(dim, Y, il, L);
  return orc_pwc_mahalanobis(dim, L);
}

/* the same with PoseWithNode: node |j - i| + 1 + 1 */
double orc_simple_landmark_pair(int dim, const double* traj, int i, const double* zil,
                                const double* sil, int j, const double* zjl, const double* sjl,
                                double* rot_out) {
  const int ps = orc_pwc_size(dim);
  double il[64], jl[64], inv[12], A[12], X[12], Y[12], L[12];
  orc_pwc_from_measurement(dim, zil, sil, il);
  orc_pwc_from_measurement(dim, zjl, sjl, jl);
  const double* Pi = traj + (size_t)i * ps;
  const double* Pj = traj + (size_t)j * ps;
  pose_inverse(dim, Pi, inv);
  pose_compose(dim, inv, Pj, A);
  pose_compose(dim, A, jl, X);
  pose_inverse(dim, X, Y);
  pose_compose(dim, Y, il, L);
  const int node = abs(j - i) + 2;
  const int ri = pwn_flag(dim, Pi) && pwn_flag(dim, Pj) && pwn_flag(dim, jl) && pwn_flag(dim, il);
  double t, r;
  orc_pwn_norms(dim, L, node, ri, &t, &r);
  if (rot_out) *rot_out = r;
  return t;
}

/* Adjacency of one landmark's observations from poses lc_from[k]: every
 * observation is kept (no odometry check), pair (i < j) consistent when its
 * distance is below lc_thr (and the rotation norm below lc_thr2 for simple);
 * a negative threshold leaves the adjacency empty. */
void orc_pcm_adjacency_landmark(int dim, const double* traj, int n_lc, const int32_t* lc_from,
                                const double* lc_pose, const double* lc_cov, int simple, double lc_thr,
                                double lc_thr2, uint8_t* adj, double* dist) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
  const int loop_check = simple ? (lc_thr >= 0.0 && lc_thr2 >= 0.0) : lc_thr >= 0.0;
  if (adj) memset(adj, 0, (size_t)n_lc * n_lc);
  if (dist) memset(dist, 0, sizeof(double) * (size_t)n_lc * n_lc);
  if (!loop_check) return;
  for (int j = 1; j < n_lc; ++j)
    for (int i = 0; i < j; ++i) {
      double dv, rot = 0.0;
      int cons;
      if (simple) {
        dv = orc_simple_landmark_pair(dim, traj, lc_from[i], lc_pose + (size_t)i * pl, lc_cov + (size_t)i * dd,
                                      lc_from[j], lc_pose + (size_t)j * pl, lc_cov + (size_t)j * dd, &rot);
        cons = dv < lc_thr && rot < lc_thr2;
      } else {
        dv = orc_landmark_pair_dist(dim, traj, lc_from[i], lc_pose + (size_t)i * pl, lc_cov + (size_t)i * dd,
                                    lc_from[j], lc_pose + (size_t)j * pl, lc_cov + (size_t)j * dd);
        cons = dv < lc_thr;
      }
      if (dist) {
        dist[(size_t)j * n_lc + i] = dv;
        dist[(size_t)i * n_lc + j] = dv;
      }
      if (adj && cons) {
        adj[(size_t)j * n_lc + i] = 1;
        adj[(size_t)i * n_lc + j] = 1;
      }
    }
}

int orc_pcm_simple_adjacency(int dim, int n_poses, const double* traj, int n_lc,
                             const int32_t* lc_from, const int32_t* lc_to, const double* lc_pose,
                             const double* lc_cov, double odom_trans, double odom_rot,
                             double dist_trans, double dist_rot, int32_t* accepted,
                             double* odom_dist_out, uint8_t* adj, double* dist) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
  (void)n_poses;
  const int odom_check = odom_trans >= 0.0 && odom_rot >= 0.0;
  const int loop_check = dist_trans >= 0.0 && dist_rot >= 0.0;
  int m = 0;
  for (int k = 0; k < n_lc; ++k) {
    double rot = 0.0;
    const double od = orc_simple_odom(dim, traj, lc_from[k], lc_to[k], lc_pose + (size_t)k * pl,
                                      lc_cov + (size_t)k * dd, &rot);
    const int ok = !odom_check || (od < odom_trans && rot < odom_rot);
    if (odom_dist_out) odom_dist_out[k] = od;
    if (ok) accepted[m++] = k;
  }
  if (adj) memset(adj, 0, (size_t)m * m);
  if (dist) memset(dist, 0, sizeof(double) * (size_t)m * m);
  if (!loop_check) return m;
  for (int j = 1; j < m; ++j) {
    const int kj = accepted[j];
    for (int i = 0; i < j; ++i) {
      const int ki = accepted[i];
      double rot = 0.0;
      const double t = orc_simple_pair(dim, traj, lc_from[ki], lc_to[ki],
                                       lc_pose + (size_t)ki * pl, lc_cov + (size_t)ki * dd,
                                       lc_from[kj], lc_to[kj], lc_pose + (size_t)kj * pl,
                                       lc_cov + (size_t)kj * dd, &rot);
      if (dist) {
        dist[(size_t)j * m + i] = t;
        dist[(size_t)i * m + j] = t;
      }
      if (adj && t < dist_trans && rot < dist_rot) {
        adj[(size_t)j * m + i] = 1;
        adj[(size_t)i * m + j] = 1;
      }
    }
  }
  return m;
}

/* ------------------------------------------------------------------ */
/* max clique                                                          */
/* ------------------------------------------------------------------ */

/* ReadEigenAdjacencyMatrix : graphIO.cpp:188-230.  Column-major scan of the
 * lower triangle gives every vertex its neighbours in ascending order.  The
 * duplicate scan (:203-209) can never fire on a lower-triangle scan, so it
 * is not repeated here. */
int orc_csr_from_dense(int n, const uint8_t* adj, int32_t* offsets, int32_t* edges,
                       int capacity) {
  int e = 0;
  offsets[0] = 0;
  for (int v = 0; v < n; ++v) {
    for (int u = 0; u < n; ++u) {
      if (u == v) continue;
      /* lower triangle entry (max(u,v), min(u,v)) decides the edge */
      const int r = u > v ? u : v, c = u > v ? v : u;
      if (adj[(size_t)r * n + c] != 0) {
        if (e < capacity) edges[e] = u;
        ++e;
      }
    }
    offsets[v + 1] = e;
  }
  return e;
}

static int deg(const int32_t* off, int v) { return off[v + 1] - off[v]; }

/* Shared body of maxCliqueHeu / maxCliqueHeuIncremental (findCliqueHeu.cpp). */
static int heu_core(int n, const int32_t* off, const int32_t* edg, int first,
                    int init_best, int32_t* out) {
  int maxDegree = -1;
  for (int v = 0; v < n; ++v)
    if (deg(off, v) > maxDegree) maxDegree = deg(off, v);
  const int bufn = maxDegree + 1 > 0 ? maxDegree + 1 : 0;
  int* S = (int*)calloc((size_t)bufn + 1, sizeof(int));
  int* S1 = (int*)calloc((size_t)bufn + 1, sizeof(int));
  int maxClq = init_best;
  for (int c = first; c < n; ++c) {
    if (maxClq > deg(off, c)) continue; /* pruning 1 */
    int iPos = 0;
    S[iPos++] = c;
    for (int j = off[c]; j < off[c + 1]; ++j) /* pruning 3 */
      if (maxClq <= deg(off, edg[j])) S[iPos++] = edg[j];
    int icc = 0;
    while (iPos > 0) {
      ++icc;
      const int m = S[iPos - 1];
      int iPos1 = 0;
      for (int j = 0; j < iPos; ++j)
        for (int k = off[m]; k < off[m + 1]; ++k)
          if (S[j] == edg[k] && maxClq <= deg(off, edg[k])) { /* pruning 5 */
            S1[iPos1++] = S[j];
            break;
          }
      for (int j = 0; j < iPos1; ++j) S[j] = S1[j];
      iPos = iPos1;
    }
    if (maxClq < icc) {
      for (int j = 0; j < bufn; ++j) out[j] = S[j];
      maxClq = icc;
    }
  }
  free(S);
  free(S1);
  return maxClq;
}

int orc_max_clique_heu(int n, const int32_t* off, const int32_t* edg, int32_t* out) {
  return heu_core(n, off, edg, 0, -1, out);
}

int orc_max_clique_heu_incremental(int n, const int32_t* off, const int32_t* edg,
                                   int num_new, int prev_size, int32_t* out) {
  const int got = heu_core(n, off, edg, n - num_new, prev_size, out);
  return got > prev_size ? got : 0;
}

/* maxClique + maxCliqueHelper : findClique.cpp:31-145 */
typedef struct {
  int* data;
  int len;
} ivec;

static void helper(const int32_t* off, const int32_t* edg, int* U, int ulen,
                   int size, int* maxClq, ivec* inter, int n) {
  if (ulen == 0) {
    if (size > *maxClq) {
      *maxClq = size;
      inter->len = 0;
    }
    return;
  }
  int* Unew = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  while (ulen > 0) {
    if (size + ulen <= *maxClq) break;
    const int index = U[--ulen];
    int nn = 0;
    for (int j = off[index]; j < off[index + 1]; ++j)
      if (deg(off, edg[j]) >= *maxClq)
        for (int i = 0; i < ulen; ++i)
          if (edg[j] == U[i]) Unew[nn++] = edg[j];
    const int prev = *maxClq;
    helper(off, edg, Unew, nn, size + 1, maxClq, inter, n);
    if (*maxClq > prev) inter->data[inter->len++] = index;
  }
  free(Unew);
}

int orc_max_clique_exact(int n, const int32_t* off, const int32_t* edg, int32_t* out) {
  int maxClq = 0;
  int* bitVec = (int*)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  int* U = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  ivec inter = {(int*)malloc(sizeof(int) * (size_t)(n + 1)), 0};
  int outlen = 0;
  for (int i = n - 1; i >= 0; --i) {
    bitVec[i] = 1;
    const int prev = maxClq;
    int ulen = 0;
    if (deg(off, i) < maxClq) continue; /* pruning 1 */
    for (int j = off[i]; j < off[i + 1]; ++j)
      if (bitVec[edg[j]] != 1 && deg(off, edg[j]) >= maxClq) U[ulen++] = edg[j];
    helper(off, edg, U, ulen, 1, &maxClq, &inter, n);
    if (maxClq > prev) {
      inter.data[inter.len++] = i;
      for (int k = 0; k < inter.len; ++k) out[k] = inter.data[k];
      outlen = inter.len;
    }
    inter.len = 0;
  }
  (void)outlen;
  free(bitVec);
  free(U);
  free(inter.data);
  return maxClq;
}

/* Same semantics as orc_max_clique_heu (findCliqueHeu.cpp:32-117) on a dense
 * byte adjacency, with neighbourhood tests through a bitset instead of the
 * CSR scan.  Candidate greedy sizes are computed on bitsets; the winning
 * candidate is then replayed with the list-based loop to recover the stale
 * buffer it leaves behind. */
static int heu_winner_replay(int n, const uint64_t* bits, int W, const int* degv, int c,
                             int best, int32_t* out) {
  int* S = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int* S1 = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  int iPos = 0;
  S[iPos++] = c;
  for (int u = 0; u < n; ++u)
    if (u != c && ((bits[(size_t)c * W + (u >> 6)] >> (u & 63)) & 1ull) && best <= degv[u])
      S[iPos++] = u;
  int icc = 0;
  int written = iPos;
  for (int j = 0; j < iPos; ++j) out[j] = S[j];
  while (iPos > 0) {
    ++icc;
    const int m = S[iPos - 1];
    int iPos1 = 0;
    for (int j = 0; j < iPos; ++j) {
      const int v = S[j];
      if (((bits[(size_t)m * W + (v >> 6)] >> (v & 63)) & 1ull) && best <= degv[v]) S1[iPos1++] = v;
    }
    for (int j = 0; j < iPos1; ++j) { S[j] = S1[j]; out[j] = S1[j]; }
    iPos = iPos1;
  }
  (void)written;
  free(S);
  free(S1);
  return icc;
}

int orc_max_clique_heu_dense(int n, const uint8_t* adj, int32_t* out) {
  if (n <= 0) return -1;
  const int W = (n + 63) / 64;
  uint64_t* bits = (uint64_t*)calloc((size_t)n * W, sizeof(uint64_t));
  int* degv = (int*)calloc((size_t)n, sizeof(int));
  for (int v = 0; v < n; ++v)
    for (int u = 0; u < n; ++u) {
      if (u == v) continue;
      const int r = u > v ? u : v, cc = u > v ? v : u;
      if (adj[(size_t)r * n + cc]) {
        bits[(size_t)v * W + (u >> 6)] |= 1ull << (u & 63);
        ++degv[v];
      }
    }
  uint64_t* S = (uint64_t*)malloc(sizeof(uint64_t) * W);
  uint64_t* D = (uint64_t*)malloc(sizeof(uint64_t) * W);
  int best = -1, win = -1, win_best = -1;
  for (int c = 0; c < n; ++c) {
    if (best > degv[c]) continue;
    /* D = { u : deg(u) >= best } */
    for (int w = 0; w < W; ++w) D[w] = 0;
    for (int u = 0; u < n; ++u)
      if (degv[u] >= best) D[u >> 6] |= 1ull << (u & 63);
    for (int w = 0; w < W; ++w) S[w] = bits[(size_t)c * W + w] & D[w];
    int icc = 0;
    for (;;) {
      ++icc;
      int top = -1;
      for (int w = W - 1; w >= 0; --w)
        if (S[w]) { top = w * 64 + 63 - __builtin_clzll(S[w]); break; }
      if (top < 0) break; /* only the seed c was left: it is removed, loop ends */
      const uint64_t* Nm = bits + (size_t)top * W;
      for (int w = 0; w < W; ++w) S[w] &= Nm[w];
    }
    if (best < icc) {
      win = c;
      win_best = best;
      best = icc;
    }
  }
  if (win >= 0) heu_winner_replay(n, bits, W, degv, win, win_best, out);
  free(S); free(D); free(bits); free(degv);
  return best;
}

/* Pairwise distances for an explicit list of (i, j) pairs (i older, j newer)
 * of an accepted loop-closure list; used to spot-check large graphs. */
void orc_pair_dist_list(int dim, const double* traj, const int32_t* lc_from,
                        const int32_t* lc_to, const double* lc_pose, const double* lc_cov,
                        int npairs, const int32_t* pi, const int32_t* pj, double* out) {
  const int pl = pose_len(dim), dd = tan_dim(dim) * tan_dim(dim);
#pragma omp parallel for schedule(dynamic, 256)
  for (int t = 0; t < npairs; ++t) {
    const int i = pi[t], j = pj[t];
    out[t] = orc_pair_dist(dim, traj, lc_from[i], lc_to[i], lc_pose + (size_t)i * pl,
                           lc_cov + (size_t)i * dd, lc_from[j], lc_to[j],
                           lc_pose + (size_t)j * pl, lc_cov + (size_t)j * dd);
  }
}
