// RobustSolverParams: the parameter set of the robust solver, with the field
// names, enums, defaults and setters of include/KimeraRPGO/SolverParams.h:11-298
// so that configuration code written for the reference compiles unchanged
// against this namespace.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace kimera_rpgo_amd {

enum class Solver { LM, GN };

enum class OutlierRemovalMethod { NONE, PCM2D, PCM3D, PCM_Simple2D, PCM_Simple3D };

enum class Verbosity { UPDATE, QUIET, VERBOSE };

enum class MultiRobotAlignMethod { NONE, L2, GNC };

struct PcmParams {
  double odom_threshold = 10.0;  // < 0 disables the check
  double lc_threshold = 5.0;
  double odom_trans_threshold = 0.05;  // PcmSimple
  double odom_rot_threshold = 0.005;
  double dist_trans_threshold = 0.01;
  double dist_rot_threshold = 0.001;
  bool incremental = false;
};

struct GncParams {
  enum class GncThresholdMode { COST = 0u, PROBABILITY = 1u };
  GncThresholdMode gnc_threshold_mode_ = GncThresholdMode::PROBABILITY;
  double gnc_inlier_threshold_ = 0.9;
  size_t max_iterations_ = 100;
  double mu_step_ = 1.4;
  double relative_cost_tol_ = 1e-5;
  double weights_tol_ = 1e-4;
  bool fix_prev_inliers_ = false;
  bool bias_odom_ = false;
};

// Settings of the GPU linear solver (no counterpart in the reference, whose
// linear step is GTSAM's sparse Cholesky): PCG relative tolerance, fine
// segment length of the block-tridiagonal preconditioner, coarse segments.
struct DeviceParams {
  int device = 0;
  double pcg_rel_tol = 1e-6;
  int pcg_segment = 0;  // <= 0: automatic (16 poses, 32 from 10,000 poses)
  int pcg_coarse = 0;   // 0: automatic (128 segments, 256 from 10,000 poses); < 0: off
  int pcm_precision = 1;  // 1: FP32 pairs + FP64 recheck, 0: FP64 replay
};

struct RobustSolverParams {
  Solver solver = Solver::LM;
  OutlierRemovalMethod outlierRemovalMethod = OutlierRemovalMethod::PCM3D;
  std::vector<char> specialSymbols;
  Verbosity verbosity = Verbosity::UPDATE;
  bool log_output = false;
  std::string log_folder;
  PcmParams pcm_params;
  GncParams gnc_params;
  bool lm_diagonal_damping = true;
  MultiRobotAlignMethod multirobot_align_method = MultiRobotAlignMethod::NONE;
  double multirobot_align_gnc_prob = 0.99;
  bool use_gnc_ = false;
  DeviceParams device_params;

  void setNoRejection(Verbosity v = Verbosity::UPDATE) {
    outlierRemovalMethod = OutlierRemovalMethod::NONE;
    verbosity = v;
  }
  void setIncremental() { pcm_params.incremental = true; }
  void setLmDiagonalDamping(bool d) { lm_diagonal_damping = d; }
  void setPcm2DParams(double odom, double lc, Verbosity v = Verbosity::UPDATE) {
    setPcm(OutlierRemovalMethod::PCM2D, odom, lc, v);
  }
  void setPcm3DParams(double odom, double lc, Verbosity v = Verbosity::UPDATE) {
    setPcm(OutlierRemovalMethod::PCM3D, odom, lc, v);
  }
  void setPcmSimple2DParams(double trans, double rot, Verbosity v = Verbosity::UPDATE) {
    setPcmSimple(OutlierRemovalMethod::PCM_Simple2D, trans, rot, trans, rot, v);
  }
  void setPcmSimple3DParams(double trans, double rot, Verbosity v = Verbosity::UPDATE) {
    setPcmSimple(OutlierRemovalMethod::PCM_Simple3D, trans, rot, trans, rot, v);
  }
  void setPcmSimple2DParams(double to, double ro, double tp, double rp, Verbosity v = Verbosity::UPDATE) {
    setPcmSimple(OutlierRemovalMethod::PCM_Simple2D, to, ro, tp, rp, v);
  }
  void setPcmSimple3DParams(double to, double ro, double tp, double rp, Verbosity v = Verbosity::UPDATE) {
    setPcmSimple(OutlierRemovalMethod::PCM_Simple3D, to, ro, tp, rp, v);
  }
  void setGncInlierCostThresholdsAtProbability(const double& alpha, const size_t& max_iterations = 100,
                                               const double& mu_step = 1.4,
                                               const double& rel_cost_tol = 1e-5,
                                               const double& weight_tol = 1e-4,
                                               const bool& fix_prev_inliers = false,
                                               const bool& bias_odom = false) {
    setGnc(GncParams::GncThresholdMode::PROBABILITY, alpha, max_iterations, mu_step, rel_cost_tol,
           weight_tol, fix_prev_inliers, bias_odom);
  }
  void setGncInlierCostThresholds(const double& cost, const size_t& max_iterations = 100,
                                  const double& mu_step = 1.4, const double& rel_cost_tol = 1e-5,
                                  const double& weight_tol = 1e-4, const bool& fix_prev_inliers = false,
                                  const bool& bias_odom = false) {
    setGnc(GncParams::GncThresholdMode::COST, cost, max_iterations, mu_step, rel_cost_tol, weight_tol,
           fix_prev_inliers, bias_odom);
  }
  void gncFixInliers() { gnc_params.fix_prev_inliers_ = true; }
  void gncBiasOdom() { gnc_params.bias_odom_ = true; }
  void setMultiRobotAlignMethod(MultiRobotAlignMethod m) { multirobot_align_method = m; }
  void setMultiRobotGncAlignProbability(double p) { multirobot_align_gnc_prob = p; }
  void logOutput(const std::string& folder) {
    log_output = true;
    log_folder = folder;
  }

 private:
  void setPcm(OutlierRemovalMethod m, double odom, double lc, Verbosity v) {
    outlierRemovalMethod = m;
    pcm_params.odom_threshold = odom;
    pcm_params.lc_threshold = lc;
    verbosity = v;
  }
  void setPcmSimple(OutlierRemovalMethod m, double to, double ro, double tp, double rp, Verbosity v) {
    outlierRemovalMethod = m;
    pcm_params.odom_trans_threshold = to;
    pcm_params.odom_rot_threshold = ro;
    pcm_params.dist_trans_threshold = tp;
    pcm_params.dist_rot_threshold = rp;
    verbosity = v;
  }
  void setGnc(GncParams::GncThresholdMode mode, double thr, size_t it, double step, double rel, double wt,
              bool fix, bool bias) {
    use_gnc_ = true;
    gnc_params.gnc_threshold_mode_ = mode;
    gnc_params.gnc_inlier_threshold_ = thr;
    gnc_params.max_iterations_ = it;
    gnc_params.mu_step_ = step;
    gnc_params.relative_cost_tol_ = rel;
    gnc_params.weights_tol_ = wt;
    gnc_params.fix_prev_inliers_ = fix;
    gnc_params.bias_odom_ = bias;
  }
};

}  // namespace kimera_rpgo_amd
