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
// Several robots (key prefixes): one odometry chain per prefix (robot order
// = first-seen order), loop closures grouped per unordered prefix pair
// (ObservationId); inter-robot loop closures skip the odometry check and each
// group is one device problem over its one or two chains.  Landmarks (keys
// whose character is a special symbol): the observation that brings the
// landmark's value starts its group, re-observations are checked pairwise
// against each other (Pcm.h:775-844) and the group's max clique is kept.
// With a MultiRobotAlignMethod other than NONE the values of every robot
// after the first (smallest prefix) are re-initialised from its odometry
// chain and the GNC pose average of the frame transforms its inter-robot
// inliers imply (Pcm.h:1011-1102), solved on the device (rpgo_solve).
//
// Differences from the reference, where its behaviour is undefined: an
// odometry factor must extend its robot's chain at the end; a landmark
// factor must run pose -> landmark and all observations of one landmark come
// from one robot; loop closures must join poses of known chains.  Each
// raises std::invalid_argument.  Group and landmark order in the graph to
// optimize is first-seen order (the reference iterates hash maps).
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
  Pcm(int dim, const PcmParams& params, MultiRobotAlignMethod align, double align_gnc_probability = 0.99,
      const std::vector<char>& special_symbols = {}, const DeviceParams& device = DeviceParams(),
      bool simple = false);
  ~Pcm() override;
  Pcm(const Pcm&) = delete;
  Pcm& operator=(const Pcm&) = delete;

  bool removeOutliers(const NonlinearFactorGraph& new_factors, const Values& new_values,
                      NonlinearFactorGraph* nfg, Values* values) override;
  // Pcm.h:135-138: every loop closure and landmark observation taken in
  // (never decremented), the inliers of the last max-clique pass
  size_t getNumLC() const override { return total_lc_; }
  size_t getNumLCInliers() const override { return total_good_lc_; }
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
  struct Chain {  // one robot's odometry (Pcm.h:516-557 trajectory)
    char prefix = 0;
    Key first_key = 0;
    Pose3 pose0_3;
    Pose2 pose0_2;
    std::vector<FactorPtr> odom;
  };
  struct Group {  // loop closures of one ObservationId (unordered prefix pair)
    char id1 = 0, id2 = 0;
    std::vector<FactorPtr> lc;  // factors that passed the odometry check
    std::vector<int> inliers;   // consistent factors: clique indices into lc
  };
  struct Landmark {  // observations of one landmark (Pcm.h:207-220, 431-448)
    Key key = 0;
    std::vector<FactorPtr> obs;
    std::vector<int> inliers;
  };

  bool isSpecialSymbol(char c) const;
  int chainOf(char prefix) const;
  int groupOf(char a, char b) const;
  void updateOdom(const FactorPtr& f, const Values& values);
  // device stages of one group: odometry check of the pending loop closures,
  // pairwise check of all, max clique; returns how many pending ones passed
  int runGroup(Group& g, const std::vector<FactorPtr>& pending, bool incremental);
  void runLandmark(Landmark& l);
  void countInliers();
  NonlinearFactorGraph buildGraphToOptimize() const;
  Values multirobotValueInitialization(const Values& values) const;
  void robotValues(const Chain& c, const Pose3* T3, const Pose2* T2, Values* out) const;
  void check(int rc, const char* what) const;

  int dim_;
  bool simple_;
  bool odom_check_, loop_check_;  // Pcm.h:74-82
  PcmParams params_;
  MultiRobotAlignMethod align_ = MultiRobotAlignMethod::NONE;
  double align_prob_ = 0.99;
  std::vector<char> special_symbols_;
  DeviceParams device_;
  rpgo_ctx* ctx_ = nullptr;
  std::vector<Chain> chains_;         // robots in first-seen order (robot_order_)
  std::vector<FactorPtr> odom_;       // every odometry factor in arrival order
  NonlinearFactorGraph nfg_special_;  // non-between factors, in arrival order
  std::vector<Group> groups_;
  std::vector<int> lc_order_;         // group of every loop closure taken in, in order
  std::vector<Landmark> landmarks_;
  size_t total_lc_ = 0, total_good_lc_ = 0;
  std::vector<char> ignored_prefixes_;
  double pcm_ms_ = 0.0, clique_ms_ = 0.0;
  long long pairs_ = 0;
};

// Pcm3D(params, align_method, align_probability, special_symbols) as the
// reference constructs them (Pcm.h:52-95), or without the alignment method.
#define RPGO_PCM_VARIANT(NAME, DIM, SIMPLE)                                                              \
  class NAME : public Pcm {                                                                              \
   public:                                                                                               \
    explicit NAME(const PcmParams& p, const std::vector<char>& special = {},                             \
                  const DeviceParams& d = DeviceParams())                                                \
        : Pcm(DIM, p, special, d, SIMPLE) {}                                                             \
    NAME(const PcmParams& p, MultiRobotAlignMethod align, double align_gnc_probability = 0.99,           \
         const std::vector<char>& special = {}, const DeviceParams& d = DeviceParams())                  \
        : Pcm(DIM, p, align, align_gnc_probability, special, d, SIMPLE) {}                              \
  };
RPGO_PCM_VARIANT(Pcm3D, 3, false)
RPGO_PCM_VARIANT(Pcm2D, 2, false)
RPGO_PCM_VARIANT(PcmSimple3D, 3, true)
RPGO_PCM_VARIANT(PcmSimple2D, 2, true)
#undef RPGO_PCM_VARIANT

}  // namespace kimera_rpgo_amd
