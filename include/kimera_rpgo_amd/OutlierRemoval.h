// Outlier-rejection plugin interface and its PCM implementation, with the
// interface of include/KimeraRPGO/outlier/OutlierRemoval.h:19-102 and the
// behaviour of Pcm<poseT, PoseWithCovariance> (outlier/Pcm.h): factors are
// classified into odometry (i -> i+1 with a new value for i+1), loop closures
// and special (non-between) factors; every loop closure is checked against
// the odometry chain (Mahalanobis distance < odom_threshold) and pairwise
// against the accepted ones (< lc_threshold); the maximum clique of the
// consistency graph (FMC::maxCliqueHeu) is the inlier set.  The pairwise
// matrix and the clique run on the GPU (rpgo_pcm_adjacency, rpgo_max_clique).
//
// PcmSimple2D/3D (Pcm<poseT, PoseWithNode>, Pcm.h:1169-1170) replace both
// distances with the per-node translation and rotation norms of
// PoseWithNode (GeometryUtils.h:193-289) against odom_trans/rot_threshold
// and dist_trans/rot_threshold (rpgo_pcm_simple_adjacency).  As in the
// reference constructor (Pcm.h:74-82) a negative odom_threshold,
// odom_trans_threshold or odom_rot_threshold turns the odometry check off
// and a negative lc_threshold, dist_trans_threshold or dist_rot_threshold
// the pairwise check (every accepted loop closure is then an inlier).
//
// Scope: one robot prefix (the intra-robot group of the reference).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "kimera_rpgo_amd/SolverParams.h"
#include "kimera_rpgo_amd/Types.h"

struct rpgo_ctx;

namespace kimera_rpgo_amd {

class OutlierRemoval {
 public:
  virtual ~OutlierRemoval() = default;
  // inserts new_values into *values, replaces *nfg with the graph to optimize;
  // returns whether the graph should be optimized
  virtual bool removeOutliers(const NonlinearFactorGraph& new_factors, const Values& new_values,
                              NonlinearFactorGraph* nfg, Values* values) = 0;
  virtual size_t getNumLC() const = 0;
  virtual size_t getNumLCInliers() const = 0;
  virtual size_t getNumOdomFactors() const = 0;
  virtual size_t getNumSpecialFactors() const = 0;
  virtual EdgePtr removeLastLoopClosure(char prefix_1, char prefix_2, NonlinearFactorGraph* nfg) = 0;
  virtual EdgePtr removeLastLoopClosure(NonlinearFactorGraph* nfg) = 0;
  virtual void ignoreLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) = 0;
  virtual void reviveLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) = 0;
  virtual std::vector<char> getIgnoredPrefixes() const = 0;
  virtual void removePriorFactorsWithPrefix(const char& prefix, NonlinearFactorGraph* nfg) = 0;
  virtual void saveData(const std::string& folder_path) {}
  void setQuiet() { debug_ = false; }
  void logOutput(const std::string& folder) {
    log_output_ = true;
    log_folder_ = folder;
  }

 protected:
  bool debug_ = true;
  bool log_output_ = false;
  std::string log_folder_;
};

class Pcm : public OutlierRemoval {
 public:
  // dim 3: Pose3 (Pcm3D), dim 2: Pose2 (Pcm2D); simple: PoseWithNode
  // distances (PcmSimple3D / PcmSimple2D)
  Pcm(int dim, const PcmParams& params, const std::vector<char>& special_symbols = {},
      const DeviceParams& device = DeviceParams(), bool simple = false);
  ~Pcm() override;
  Pcm(const Pcm&) = delete;
  Pcm& operator=(const Pcm&) = delete;

  bool removeOutliers(const NonlinearFactorGraph& new_factors, const Values& new_values,
                      NonlinearFactorGraph* nfg, Values* values) override;
  size_t getNumLC() const override { return lc_.size(); }
  size_t getNumLCInliers() const override { return inliers_.size(); }
  size_t getNumOdomFactors() const override { return odom_.size(); }
  size_t getNumSpecialFactors() const override { return nfg_special_.size(); }
  EdgePtr removeLastLoopClosure(char prefix_1, char prefix_2, NonlinearFactorGraph* nfg) override;
  EdgePtr removeLastLoopClosure(NonlinearFactorGraph* nfg) override;
  void ignoreLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) override;
  void reviveLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) override;
  std::vector<char> getIgnoredPrefixes() const override { return ignored_prefixes_; }
  void removePriorFactorsWithPrefix(const char& prefix, NonlinearFactorGraph* nfg) override;

  // device time and pair count of the last consistency pass
  double lastPcmMs() const { return pcm_ms_; }
  double lastCliqueMs() const { return clique_ms_; }
  long long lastPairs() const { return pairs_; }

 private:
  bool isSpecialSymbol(char c) const;
  void updateOdom(const FactorPtr& f, const Values& values);
  void runPcm(const std::vector<FactorPtr>& pending);
  NonlinearFactorGraph buildGraphToOptimize() const;
  void check(int rc, const char* what) const;

  int dim_;
  bool simple_;
  bool odom_check_, loop_check_;  // Pcm.h:74-82
  PcmParams params_;
  std::vector<char> special_symbols_;
  DeviceParams device_;
  rpgo_ctx* ctx_ = nullptr;
  char prefix_ = 0;
  bool have_traj_ = false;
  Key first_key_ = 0;
  Pose3 pose0_3_;
  Pose2 pose0_2_;
  std::vector<FactorPtr> odom_;       // odometry factors in chain order
  NonlinearFactorGraph nfg_special_;  // non-between factors, in arrival order
  std::vector<FactorPtr> lc_;         // loop closures that passed the odometry check
  std::vector<int> inliers_;          // clique output (indices into lc_)
  std::vector<char> ignored_prefixes_;
  double pcm_ms_ = 0.0, clique_ms_ = 0.0;
  long long pairs_ = 0;
};

class Pcm3D : public Pcm {
 public:
  explicit Pcm3D(const PcmParams& p, const std::vector<char>& special = {}, const DeviceParams& d = DeviceParams())
      : Pcm(3, p, special, d) {}
};

class Pcm2D : public Pcm {
 public:
  explicit Pcm2D(const PcmParams& p, const std::vector<char>& special = {}, const DeviceParams& d = DeviceParams())
      : Pcm(2, p, special, d) {}
};

class PcmSimple3D : public Pcm {
 public:
  explicit PcmSimple3D(const PcmParams& p, const std::vector<char>& special = {},
                       const DeviceParams& d = DeviceParams())
      : Pcm(3, p, special, d, true) {}
};

class PcmSimple2D : public Pcm {
 public:
  explicit PcmSimple2D(const PcmParams& p, const std::vector<char>& special = {},
                       const DeviceParams& d = DeviceParams())
      : Pcm(2, p, special, d, true) {}
};

}  // namespace kimera_rpgo_amd
