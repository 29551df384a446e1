// Host-side pose algebra, noise models and containers (FP64).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "kimera_rpgo_amd/Types.h"

namespace kimera_rpgo_amd {

Pose3 Pose3::FromQuaternion(double qw, double qx, double qy, double qz, double x, double y, double z) {
  // unit-quaternion rotation matrix (as Eigen::Quaternion::toRotationMatrix)
  const double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
  const double twx = tx * qw, twy = ty * qw, twz = tz * qw;
  const double txx = tx * qx, txy = ty * qx, txz = tz * qx;
  const double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  Pose3 p;
  p.R = {1 - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1 - (txx + tzz), tyz - twx,
         txz - twy, tyz + twx, 1 - (txx + tyy)};
  p.t = {x, y, z};
  return p;
}

Pose3 Pose3::compose(const Pose3& o) const {
  Pose3 r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      r.R[i * 3 + j] = R[i * 3] * o.R[j] + R[i * 3 + 1] * o.R[3 + j] + R[i * 3 + 2] * o.R[6 + j];
    r.t[i] = R[i * 3] * o.t[0] + R[i * 3 + 1] * o.t[1] + R[i * 3 + 2] * o.t[2] + t[i];
  }
  return r;
}

Pose3 Pose3::inverse() const {
  Pose3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.R[i * 3 + j] = R[j * 3 + i];
  for (int i = 0; i < 3; ++i) r.t[i] = -(r.R[i * 3] * t[0] + r.R[i * 3 + 1] * t[1] + r.R[i * 3 + 2] * t[2]);
  return r;
}

Pose3 Pose3::Expmap(const std::array<double, 6>& xi) {
  const double wx = xi[0], wy = xi[1], wz = xi[2];
  const double th2 = wx * wx + wy * wy + wz * wz, th = std::sqrt(th2);
  double A, B, C;
  if (th2 < 1e-20) {
    A = 1 - th2 / 6;
    B = 0.5 - th2 / 24;
    C = 1.0 / 6 - th2 / 120;
  } else {
    A = std::sin(th) / th;
    B = (1 - std::cos(th)) / th2;
    C = (th - std::sin(th)) / (th2 * th);
  }
  const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
  double W2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
  Pose3 p;
  double V[9];
  for (int e = 0; e < 9; ++e) {
    const double I = (e % 4 == 0) ? 1.0 : 0.0;
    p.R[e] = I + A * W[e] + B * W2[e];
    V[e] = I + B * W[e] + C * W2[e];
  }
  for (int i = 0; i < 3; ++i) p.t[i] = V[i * 3] * xi[3] + V[i * 3 + 1] * xi[4] + V[i * 3 + 2] * xi[5];
  return p;
}

bool Pose3::equals(const Pose3& o, double tol) const {
  for (int e = 0; e < 9; ++e)
    if (std::fabs(R[e] - o.R[e]) > tol) return false;
  for (int e = 0; e < 3; ++e)
    if (std::fabs(t[e] - o.t[e]) > tol) return false;
  return true;
}

static double wrap_angle(double a) {
  // Rot2 normalisation: the angle of (cos a, sin a)
  return std::atan2(std::sin(a), std::cos(a));
}

Pose2 Pose2::compose(const Pose2& o) const {
  const double c = std::cos(theta), s = std::sin(theta);
  return Pose2(x + c * o.x - s * o.y, y + s * o.x + c * o.y, wrap_angle(theta + o.theta));
}

Pose2 Pose2::inverse() const {
  const double c = std::cos(theta), s = std::sin(theta);
  return Pose2(-(c * x + s * y), -(-s * x + c * y), wrap_angle(-theta));
}

bool Pose2::equals(const Pose2& o, double tol) const {
  return std::fabs(x - o.x) <= tol && std::fabs(y - o.y) <= tol &&
         std::fabs(wrap_angle(theta - o.theta)) <= tol;
}

// inverse of a symmetric positive-definite d x d matrix (Gauss-Jordan with
// partial pivoting; d <= 6)
static std::vector<double> inverse_spd(int d, const std::vector<double>& A) {
  std::vector<double> M(A), I(d * d, 0.0);
  for (int i = 0; i < d; ++i) I[i * d + i] = 1.0;
  for (int c = 0; c < d; ++c) {
    int piv = c;
    for (int r = c + 1; r < d; ++r)
      if (std::fabs(M[r * d + c]) > std::fabs(M[piv * d + c])) piv = r;
    if (M[piv * d + c] == 0.0) throw std::invalid_argument("noise model: singular matrix");
    if (piv != c)
      for (int k = 0; k < d; ++k) {
        std::swap(M[c * d + k], M[piv * d + k]);
        std::swap(I[c * d + k], I[piv * d + k]);
      }
    const double inv = 1.0 / M[c * d + c];
    for (int k = 0; k < d; ++k) {
      M[c * d + k] *= inv;
      I[c * d + k] *= inv;
    }
    for (int r = 0; r < d; ++r) {
      if (r == c) continue;
      const double f = M[r * d + c];
      if (f == 0.0) continue;
      for (int k = 0; k < d; ++k) {
        M[r * d + k] -= f * M[c * d + k];
        I[r * d + k] -= f * I[c * d + k];
      }
    }
  }
  return I;
}

std::shared_ptr<NoiseModel> NoiseModel::Information(int dim, const std::vector<double>& info) {
  if ((int)info.size() != dim * dim) throw std::invalid_argument("noise model: size mismatch");
  auto m = std::make_shared<NoiseModel>();
  m->dim_ = dim;
  m->info_ = info;
  m->cov_ = inverse_spd(dim, info);
  return m;
}

std::shared_ptr<NoiseModel> NoiseModel::Covariance(int dim, const std::vector<double>& cov) {
  if ((int)cov.size() != dim * dim) throw std::invalid_argument("noise model: size mismatch");
  auto m = std::make_shared<NoiseModel>();
  m->dim_ = dim;
  m->cov_ = cov;
  m->info_ = inverse_spd(dim, cov);
  return m;
}

std::shared_ptr<NoiseModel> NoiseModel::Variances(const std::vector<double>& v) {
  const int d = (int)v.size();
  auto m = std::make_shared<NoiseModel>();
  m->dim_ = d;
  m->info_.assign(d * d, 0.0);
  m->cov_.assign(d * d, 0.0);
  for (int i = 0; i < d; ++i) {
    m->cov_[i * d + i] = v[i];
    m->info_[i * d + i] = 1.0 / v[i];
  }
  return m;
}

std::shared_ptr<NoiseModel> NoiseModel::Precisions(const std::vector<double>& p) {
  const int d = (int)p.size();
  auto m = std::make_shared<NoiseModel>();
  m->dim_ = d;
  m->info_.assign(d * d, 0.0);
  for (int i = 0; i < d; ++i) m->info_[i * d + i] = p[i];
  // Rinv = R^-1 column by column (R = diag(sqrt(p)), upper-triangular
  // back-substitution in IEEE arithmetic: 0 * inf and 0 / 0 give NaN where a
  // precision is zero), covariance = Rinv Rinv^T
  std::vector<double> R(d * d, 0.0), Ri(d * d, 0.0);
  for (int i = 0; i < d; ++i) R[i * d + i] = std::sqrt(p[i]);
  for (int c = 0; c < d; ++c)
    for (int r = d - 1; r >= 0; --r) {
      double s = (r == c) ? 1.0 : 0.0;
      for (int k = r + 1; k < d; ++k) s -= R[r * d + k] * Ri[k * d + c];
      Ri[r * d + c] = s / R[r * d + r];
    }
  m->cov_.assign(d * d, 0.0);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) s += Ri[i * d + k] * Ri[j * d + k];
      m->cov_[i * d + j] = s;
    }
  return m;
}

std::shared_ptr<NoiseModel> NoiseModel::Sigmas(const std::vector<double>& s) {
  std::vector<double> v(s.size());
  for (size_t i = 0; i < s.size(); ++i) v[i] = s[i] * s[i];
  return Variances(v);
}

std::shared_ptr<NoiseModel> NoiseModel::Isotropic(int dim, double sigma) {
  return Sigmas(std::vector<double>(dim, sigma));
}

std::shared_ptr<NoiseModel> NoiseModel::IsotropicVariance(int dim, double variance) {
  return Variances(std::vector<double>(dim, variance));
}

static FactorPtr make_factor(FactorType t, Key i, Key j, const Pose3* p3, const Pose2* p2,
                             const SharedNoiseModel& noise, int dim) {
  if (!noise || noise->dim() != dim) throw std::invalid_argument("factor: noise model of the wrong dimension");
  auto f = std::make_shared<Factor>();
  f->type = t;
  f->key1 = i;
  f->key2 = j;
  if (p3) f->pose3 = *p3;
  if (p2) f->pose2 = *p2;
  f->noise = noise;
  return f;
}

FactorPtr BetweenFactorPose3(Key i, Key j, const Pose3& z, const SharedNoiseModel& noise) {
  return make_factor(FactorType::BETWEEN_POSE3, i, j, &z, nullptr, noise, 6);
}
FactorPtr PriorFactorPose3(Key k, const Pose3& prior, const SharedNoiseModel& noise) {
  return make_factor(FactorType::PRIOR_POSE3, k, k, &prior, nullptr, noise, 6);
}
FactorPtr BetweenFactorPose2(Key i, Key j, const Pose2& z, const SharedNoiseModel& noise) {
  return make_factor(FactorType::BETWEEN_POSE2, i, j, nullptr, &z, noise, 3);
}
FactorPtr PriorFactorPose2(Key k, const Pose2& prior, const SharedNoiseModel& noise) {
  return make_factor(FactorType::PRIOR_POSE2, k, k, nullptr, &prior, noise, 3);
}

void Values::insert(Key k, const Pose3& p) {
  if (exists(k)) throw std::invalid_argument("Values::insert: key already exists");
  p3_[k] = p;
}
void Values::insert(Key k, const Pose2& p) {
  if (exists(k)) throw std::invalid_argument("Values::insert: key already exists");
  p2_[k] = p;
}
void Values::insert(const Values& o) {
  for (const auto& kv : o.p3_) insert(kv.first, kv.second);
  for (const auto& kv : o.p2_) insert(kv.first, kv.second);
}
void Values::update(Key k, const Pose3& p) {
  auto it = p3_.find(k);
  if (it == p3_.end()) throw std::invalid_argument("Values::update: no Pose3 at key");
  it->second = p;
}
void Values::update(Key k, const Pose2& p) {
  auto it = p2_.find(k);
  if (it == p2_.end()) throw std::invalid_argument("Values::update: no Pose2 at key");
  it->second = p;
}
const Pose3& Values::atPose3(Key k) const {
  auto it = p3_.find(k);
  if (it == p3_.end()) throw std::out_of_range("Values::at: no Pose3 at key");
  return it->second;
}
const Pose2& Values::atPose2(Key k) const {
  auto it = p2_.find(k);
  if (it == p2_.end()) throw std::out_of_range("Values::at: no Pose2 at key");
  return it->second;
}
std::vector<Key> Values::keys() const {
  std::vector<Key> k;
  k.reserve(size());
  for (const auto& kv : p3_) k.push_back(kv.first);
  for (const auto& kv : p2_) k.push_back(kv.first);
  std::sort(k.begin(), k.end());
  return k;
}

}  // namespace kimera_rpgo_amd
