// Host-side value types of the robust back end: keys, SE(3)/SE(2) poses,
// Gaussian noise models, the factors the robust solver handles and the
// containers that hold them.  They stand in for the GTSAM types the
// reference's RobustSolver API takes (gtsam::Symbol, Pose3, Pose2,
// noiseModel::Gaussian, BetweenFactor, PriorFactor, NonlinearFactorGraph,
// Values; include/KimeraRPGO/RobustSolver.h:30-139), with the same meaning:
//   * keys pack a character and an index as gtsam::Symbol does
//     (chr << 56 | index);
//   * Pose3 is R (row-major) and t, tangent order (omega, v);
//   * Pose2 is (x, y, theta), tangent order (x, y, theta);
//   * a noise model is kept as its information matrix (R^T R) and its
//     covariance, both in the tangent order of the pose type.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace kimera_rpgo_amd {

using Key = std::uint64_t;

// gtsam::Symbol packing: 8-bit character, 56-bit index
struct Symbol {
  char c = 0;
  std::uint64_t j = 0;
  Symbol() = default;
  Symbol(char chr, std::uint64_t index) : c(chr), j(index) {}
  explicit Symbol(Key key) : c((char)((key >> 56) & 0xFF)), j(key & ((std::uint64_t(1) << 56) - 1)) {}
  Key key() const { return (Key((unsigned char)c) << 56) | j; }
  operator Key() const { return key(); }
  char chr() const { return c; }
  std::uint64_t index() const { return j; }
};

struct Pose3 {
  std::array<double, 9> R{1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::array<double, 3> t{0, 0, 0};
  Pose3() = default;
  Pose3(const std::array<double, 9>& r, const std::array<double, 3>& tt) : R(r), t(tt) {}
  // rotation from a unit quaternion (w, x, y, z)
  static Pose3 FromQuaternion(double qw, double qx, double qy, double qz, double x, double y, double z);
  Pose3 compose(const Pose3& o) const;
  Pose3 inverse() const;
  Pose3 between(const Pose3& o) const { return inverse().compose(o); }
  // SO(3) exponential / SE(3) exponential in (omega, v) order
  static Pose3 Expmap(const std::array<double, 6>& xi);
  double x() const { return t[0]; }
  double y() const { return t[1]; }
  double z() const { return t[2]; }
  bool equals(const Pose3& o, double tol = 1e-9) const;
};

struct Pose2 {
  double x = 0, y = 0, theta = 0;
  Pose2() = default;
  Pose2(double xx, double yy, double th) : x(xx), y(yy), theta(th) {}
  Pose2 compose(const Pose2& o) const;
  Pose2 inverse() const;
  Pose2 between(const Pose2& o) const { return inverse().compose(o); }
  bool equals(const Pose2& o, double tol = 1e-9) const;
};

// Gaussian noise model of dimension d (6 for Pose3, 3 for Pose2), kept as
// information and covariance matrices (row-major d x d).
class NoiseModel {
 public:
  static std::shared_ptr<NoiseModel> Information(int dim, const std::vector<double>& info);
  static std::shared_ptr<NoiseModel> Covariance(int dim, const std::vector<double>& cov);
  static std::shared_ptr<NoiseModel> Sigmas(const std::vector<double>& sigmas);
  static std::shared_ptr<NoiseModel> Variances(const std::vector<double>& variances);
  // diagonal information; a zero precision leaves that direction unobserved
  // and, as gtsam's Gaussian::covariance() (R^-1 R^-T by back-substitution)
  // does, fills the covariance rows and columns that meet it with NaN
  static std::shared_ptr<NoiseModel> Precisions(const std::vector<double>& precisions);
  static std::shared_ptr<NoiseModel> Isotropic(int dim, double sigma);
  static std::shared_ptr<NoiseModel> IsotropicVariance(int dim, double variance);
  int dim() const { return dim_; }
  const std::vector<double>& information() const { return info_; }
  const std::vector<double>& covariance() const { return cov_; }

 private:
  int dim_ = 0;
  std::vector<double> info_, cov_;
};
using SharedNoiseModel = std::shared_ptr<NoiseModel>;

// noiseModel namespace spelled as GTSAM's so call sites read the same
namespace noiseModel {
struct Isotropic {
  static SharedNoiseModel Sigma(int dim, double s) { return NoiseModel::Isotropic(dim, s); }
  static SharedNoiseModel Variance(int dim, double v) { return NoiseModel::IsotropicVariance(dim, v); }
};
struct Diagonal {
  static SharedNoiseModel Sigmas(const std::vector<double>& s) { return NoiseModel::Sigmas(s); }
  static SharedNoiseModel Variances(const std::vector<double>& v) { return NoiseModel::Variances(v); }
  static SharedNoiseModel Precisions(const std::vector<double>& p) { return NoiseModel::Precisions(p); }
};
struct Gaussian {
  static SharedNoiseModel Information(int dim, const std::vector<double>& m) {
    return NoiseModel::Information(dim, m);
  }
  static SharedNoiseModel Covariance(int dim, const std::vector<double>& m) {
    return NoiseModel::Covariance(dim, m);
  }
};
}  // namespace noiseModel

enum class FactorType { BETWEEN_POSE3, PRIOR_POSE3, BETWEEN_POSE2, PRIOR_POSE2 };

// BetweenFactor<Pose3|Pose2>(key1, key2, measured, noise) or
// PriorFactor<Pose3|Pose2>(key1, prior, noise)
struct Factor {
  FactorType type;
  Key key1 = 0, key2 = 0;
  Pose3 pose3;
  Pose2 pose2;
  SharedNoiseModel noise;
  bool isBetween() const { return type == FactorType::BETWEEN_POSE3 || type == FactorType::BETWEEN_POSE2; }
  bool isPrior() const { return !isBetween(); }
  bool is3D() const { return type == FactorType::BETWEEN_POSE3 || type == FactorType::PRIOR_POSE3; }
  Key front() const { return key1; }
  Key back() const { return isBetween() ? key2 : key1; }
};
using FactorPtr = std::shared_ptr<const Factor>;

FactorPtr BetweenFactorPose3(Key i, Key j, const Pose3& z, const SharedNoiseModel& noise);
FactorPtr PriorFactorPose3(Key k, const Pose3& prior, const SharedNoiseModel& noise);
FactorPtr BetweenFactorPose2(Key i, Key j, const Pose2& z, const SharedNoiseModel& noise);
FactorPtr PriorFactorPose2(Key k, const Pose2& prior, const SharedNoiseModel& noise);

class NonlinearFactorGraph {
 public:
  void add(const FactorPtr& f) { f_.push_back(f); }
  void add(const NonlinearFactorGraph& g) { f_.insert(f_.end(), g.f_.begin(), g.f_.end()); }
  void push_back(const FactorPtr& f) { f_.push_back(f); }
  size_t size() const { return f_.size(); }
  bool empty() const { return f_.empty(); }
  const FactorPtr& operator[](size_t i) const { return f_[i]; }
  const FactorPtr& at(size_t i) const { return f_.at(i); }
  void erase_last() { f_.pop_back(); }
  std::vector<FactorPtr>::const_iterator begin() const { return f_.begin(); }
  std::vector<FactorPtr>::const_iterator end() const { return f_.end(); }

 private:
  std::vector<FactorPtr> f_;
};

// Values: key -> Pose3 or Pose2 (one pose type per key)
class Values {
 public:
  void insert(Key k, const Pose3& p);
  void insert(Key k, const Pose2& p);
  // inserts every entry of o; throws if a key already exists (gtsam::Values::insert)
  void insert(const Values& o);
  void update(Key k, const Pose3& p);
  void update(Key k, const Pose2& p);
  bool exists(Key k) const { return p3_.count(k) || p2_.count(k); }
  size_t size() const { return p3_.size() + p2_.size(); }
  bool empty() const { return size() == 0; }
  const Pose3& atPose3(Key k) const;
  const Pose2& atPose2(Key k) const;
  bool isPose3(Key k) const { return p3_.count(k) != 0; }
  std::vector<Key> keys() const;
  const std::map<Key, Pose3>& pose3s() const { return p3_; }
  const std::map<Key, Pose2>& pose2s() const { return p2_; }

 private:
  std::map<Key, Pose3> p3_;
  std::map<Key, Pose2> p2_;
};

// (front, back) keys of a removed loop closure (TypeUtils.h Edge)
struct Edge {
  Key from_key = 0, to_key = 0;
  Edge() = default;
  Edge(Key a, Key b) : from_key(a), to_key(b) {}
};
using EdgePtr = std::unique_ptr<Edge>;

}  // namespace kimera_rpgo_amd
