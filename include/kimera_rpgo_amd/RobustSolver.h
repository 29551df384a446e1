// RobustSolver: the robust pose-graph back end behind the reference's API
// (include/KimeraRPGO/RobustSolver.h:30-139, GenericSolver.h:24-90): PCM
// outlier rejection (pairwise consistency + max clique) and GNC-TLS over
// LM/GN, every heavy stage run on the GPU through the C ABI of rpgo_hip.h.
//
// Scope: Pose3 or Pose2 graphs of one or several robots (one odometry chain
// per key prefix), with inter-robot loop closures grouped per robot pair and
// landmark (special-symbol) observations grouped per landmark, as the
// reference's Pcm does (Pcm.h:414-502, 775-844).  std::invalid_argument is
// thrown where the reference's behaviour is undefined: odometry that does not
// extend its robot's chain at the end, a landmark factor from landmark to pose,
// a landmark observed from two robots, a loop closure on a robot without
// odometry.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "kimera_rpgo_amd/OutlierRemoval.h"
#include "kimera_rpgo_amd/SolverParams.h"
#include "kimera_rpgo_amd/Types.h"

struct rpgo_ctx;

namespace kimera_rpgo_amd {

class RobustSolver {
 public:
  explicit RobustSolver(const RobustSolverParams& params);
  ~RobustSolver();
  RobustSolver(const RobustSolver&) = delete;
  RobustSolver& operator=(const RobustSolver&) = delete;

  // RobustSolver.h:65-67: classify new factors (odometry, loop closure,
  // special), run PCM when loop closures arrive, then optimize
  void update(const NonlinearFactorGraph& factors, const Values& values = Values(),
              bool optimize_graph = true);
  // RobustSolver.h:54-56: same, and always optimize
  void forceUpdate(const NonlinearFactorGraph& factors = NonlinearFactorGraph(),
                   const Values& values = Values());
  // README.md:114-118: the graph with a prior on its first pose / a between
  // factor that connects it to what is loaded already
  void loadGraph(const NonlinearFactorGraph& factors, const Values& values, const FactorPtr& prior);
  void addGraph(const NonlinearFactorGraph& factors, const Values& values, const FactorPtr& between);

  size_t getNumLC() const;
  size_t getNumLCInliers() const;
  const std::vector<double>& getGncWeights() const { return gnc_weights_; }
  const std::vector<double>& getGncTempWeights() const { return gnc_temp_weights_; }

  EdgePtr removeLastLoopClosure(char prefix_1, char prefix_2);
  EdgePtr removeLastLoopClosure();
  void ignorePrefix(char prefix);
  void revivePrefix(char prefix);
  std::vector<char> getIgnoredPrefixes() const;
  void removePriorFactorsWithPrefix(const char& prefix, bool optimize_graph = false);
  void saveData(const std::string& folder_path) const;

  // GenericSolver.h:39-69
  size_t size() const { return nfg_.size(); }
  Values calculateEstimate() const { return values_; }
  Values calculateBestEstimate() const { return values_; }
  Values getLinearizationPoint() const { return values_; }
  NonlinearFactorGraph getFactorsUnsafe() const { return nfg_; }
  Values getTempValues() const { return temp_values_; }
  NonlinearFactorGraph getTempFactorsUnsafe() const { return temp_nfg_; }
  void updateTempFactorsValues(const NonlinearFactorGraph& f, const Values& v);
  void replaceTempFactorsValues(const NonlinearFactorGraph& f, const Values& v);
  void clearTempFactorsValues();
  void setQuiet() { debug_ = false; }

  // statistics of the last PCM and solve (device time, iterations)
  struct Stats {
    double pcm_ms = 0.0, clique_ms = 0.0, solve_ms = 0.0;
    int gnc_iterations = 0, lm_iterations = 0, pcg_iterations = 0;
    long long pcm_pairs = 0;
  };
  const Stats& lastStats() const { return stats_; }

 private:
  bool addAndCheckIfOptimize(const NonlinearFactorGraph& nfg, const Values& values);
  void optimize();
  void check(int rc, const char* what) const;
  void logStatus(long long spin_us);

  RobustSolverParams params_;
  std::unique_ptr<OutlierRemoval> outlier_removal_;
  rpgo_ctx* ctx_ = nullptr;
  bool debug_ = true;
  Values values_, temp_values_;
  NonlinearFactorGraph nfg_, temp_nfg_;
  std::vector<double> gnc_weights_, gnc_temp_weights_;
  size_t gnc_num_inliers_ = 0;
  size_t latest_num_lc_ = 0;
  Stats stats_;
};

}  // namespace kimera_rpgo_amd
