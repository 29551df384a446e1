// RobustSolver facade: the reference's update / optimize flow
// (src/RobustSolver.cpp:28-466, GenericSolver.cpp:37-140) over the GPU
// stages of rpgo_hip.h.  Host code keeps the bookkeeping (graph assembly,
// weights, values); outlier rejection is the Pcm plugin, and every nonlinear
// solve runs on the device through rpgo_solve.
#include "kimera_rpgo_amd/RobustSolver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>

#include "kimera_rpgo_amd/Dataset.h"
#include "rpgo_hip.h"

namespace kimera_rpgo_amd {

namespace {

void put_pose3(const Pose3& p, double* o) {
  for (int e = 0; e < 9; ++e) o[e] = p.R[e];
  for (int e = 0; e < 3; ++e) o[9 + e] = p.t[e];
}

// Pose2 (x, y, theta) as the planar Pose3 (Rz(theta), (x, y, 0))
Pose3 embed(const Pose2& p) {
  const double c = std::cos(p.theta), s = std::sin(p.theta);
  Pose3 q;
  q.R = {c, -s, 0, s, c, 0, 0, 0, 1};
  q.t = {p.x, p.y, 0.0};
  return q;
}

// Pose2 information (x, y, theta) inside the SE(3) tangent (omega, v): theta
// -> omega_z, x -> v_x, y -> v_y.  The out-of-plane directions decouple from
// the planar ones for planar poses (Ad, Jr^-1 and the Hessian are
// block-diagonal across the two sets), so they carry unit information that
// changes neither the planar solution nor its cost (their residuals vanish).
void embed_info(const std::vector<double>& I3, double* I6) {
  static const int map3[3] = {3, 4, 2};
  for (int e = 0; e < 36; ++e) I6[e] = 0.0;
  I6[0 * 7] = I6[1 * 7] = I6[5 * 7] = 1.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) I6[map3[a] * 6 + map3[b]] = I3[a * 3 + b];
}

}  // namespace

RobustSolver::RobustSolver(const RobustSolverParams& params) : params_(params) {
  switch (params.outlierRemovalMethod) {
    case OutlierRemovalMethod::NONE:
      break;
    case OutlierRemovalMethod::PCM2D:
      outlier_removal_.reset(new Pcm2D(params.pcm_params, params.multirobot_align_method,
                                          params.multirobot_align_gnc_prob, params.specialSymbols,
                                          params.device_params));
      break;
    case OutlierRemovalMethod::PCM3D:
      outlier_removal_.reset(new Pcm3D(params.pcm_params, params.multirobot_align_method,
                                          params.multirobot_align_gnc_prob, params.specialSymbols,
                                          params.device_params));
      break;
    case OutlierRemovalMethod::PCM_Simple2D:
      outlier_removal_.reset(new PcmSimple2D(params.pcm_params, params.multirobot_align_method,
                                          params.multirobot_align_gnc_prob, params.specialSymbols,
                                          params.device_params));
      break;
    case OutlierRemovalMethod::PCM_Simple3D:
      outlier_removal_.reset(new PcmSimple3D(params.pcm_params, params.multirobot_align_method,
                                          params.multirobot_align_gnc_prob, params.specialSymbols,
                                          params.device_params));
      break;
  }
  if (params.verbosity != Verbosity::VERBOSE && outlier_removal_) outlier_removal_->setQuiet();
  if (params.verbosity == Verbosity::QUIET) debug_ = false;
  check(rpgo_ctx_create(params.device_params.device, 0, &ctx_), "rpgo_ctx_create");
  if (params.log_output) {
    if (outlier_removal_) outlier_removal_->logOutput(params.log_folder);
    std::ofstream f(params.log_folder + "/rpgo_status.csv");
    f << "graph-size,spin-time(mu-s),num-lc,num-inliers\n";
  }
}

RobustSolver::~RobustSolver() {
  outlier_removal_.reset();
  if (ctx_) rpgo_ctx_destroy(ctx_);
}

void RobustSolver::check(int rc, const char* what) const {
  if (rc != RPGO_OK) {
    const char* m = rpgo_last_error();
    throw std::runtime_error(std::string(what) + " failed: " + (m ? m : "?"));
  }
}

// GenericSolver.cpp:47-62
bool RobustSolver::addAndCheckIfOptimize(const NonlinearFactorGraph& nfg, const Values& values) {
  nfg_.add(nfg);
  values_.insert(values);
  if (nfg.size() == 1 && nfg[0]->isBetween() && values.size() == 1) return false;
  if (nfg.size() == 0 && values.size() == 0) return false;
  return true;
}

// RobustSolver.cpp:115-306
void RobustSolver::optimize() {
  Values full_values = values_;
  full_values.insert(temp_values_);
  NonlinearFactorGraph full_nfg = nfg_;
  full_nfg.add(temp_nfg_);
  const bool pcm = outlier_removal_ != nullptr;
  const size_t n_known = pcm ? outlier_removal_->getNumOdomFactors() + outlier_removal_->getNumSpecialFactors() : 0;
  const bool use_gnc = params_.use_gnc_ && pcm &&
                       !(params_.gnc_params.fix_prev_inliers_ && getNumLC() == latest_num_lc_);
  if (!use_gnc && params_.gnc_params.fix_prev_inliers_ && pcm && gnc_weights_.size() >= latest_num_lc_) {
    // keep only the loop closures GNC accepted before (RobustSolver.cpp:183-196)
    const size_t prev_k = gnc_weights_.size() - latest_num_lc_;
    full_nfg = NonlinearFactorGraph();
    for (size_t i = 0; i < n_known && i < nfg_.size(); ++i) full_nfg.add(nfg_[i]);
    for (size_t i = 0; i < latest_num_lc_ && n_known + i < nfg_.size(); ++i)
      if (gnc_weights_[prev_k + i] > 0.5) full_nfg.add(nfg_[n_known + i]);
    full_nfg.add(temp_nfg_);
  }
  // dense pose indices in key order
  const std::vector<Key> keys = full_values.keys();
  if (keys.empty()) return;
  const bool planar = !full_values.pose2s().empty();
  if (planar && !full_values.pose3s().empty())
    throw std::invalid_argument("RobustSolver: mixed Pose2 / Pose3 values");
  std::map<Key, int> index;
  for (size_t k = 0; k < keys.size(); ++k) index[keys[k]] = (int)k;
  const int N = (int)keys.size(), F = (int)full_nfg.size();
  std::vector<double> x0((size_t)N * 12);
  for (int k = 0; k < N; ++k)
    put_pose3(planar ? embed(full_values.atPose2(keys[k])) : full_values.atPose3(keys[k]), &x0[(size_t)k * 12]);
  std::vector<int32_t> kind(F), ki(F), kj(F);
  std::vector<double> meas((size_t)F * 12), info((size_t)F * 36);
  for (int f = 0; f < F; ++f) {
    const Factor& fa = *full_nfg[f];
    if (fa.is3D() == planar) throw std::invalid_argument("RobustSolver: factor and value types differ");
    auto it1 = index.find(fa.front()), it2 = index.find(fa.back());
    if (it1 == index.end() || it2 == index.end())
      throw std::invalid_argument("RobustSolver: factor on a key without a value");
    kind[f] = fa.isBetween() ? RPGO_FACTOR_BETWEEN : RPGO_FACTOR_PRIOR;
    ki[f] = it1->second;
    kj[f] = it2->second;
    put_pose3(planar ? embed(fa.pose2) : fa.pose3, &meas[(size_t)f * 12]);
    if (planar) embed_info(fa.noise->information(), &info[(size_t)f * 36]);
    else std::copy(fa.noise->information().begin(), fa.noise->information().end(), &info[(size_t)f * 36]);
  }
  rpgo_solve_problem prob{};
  prob.n_poses = N;
  prob.poses = x0.data();
  prob.n_factors = F;
  prob.kind = kind.data();
  prob.key_i = ki.data();
  prob.key_j = kj.data();
  prob.measured = meas.data();
  prob.information = info.data();
  rpgo_solve_params sp;
  rpgo_solve_params_default(&sp);
  sp.solver = params_.solver == Solver::GN ? RPGO_SOLVER_GN : RPGO_SOLVER_LM;
  sp.lm_diagonal_damping = params_.lm_diagonal_damping ? 1 : 0;
  sp.pcg_rel_tol = params_.device_params.pcg_rel_tol;
  sp.pcg_segment = params_.device_params.pcg_segment;
  sp.pcg_coarse = params_.device_params.pcg_coarse;
  sp.gnc_factor_dim = planar ? 3 : 6;
  sp.pose2_chart = planar ? 1 : 0;  // Pose2 factors keep GTSAM's Pose2 chart
  std::vector<double> w0;
  if (use_gnc) {
    const auto& g = params_.gnc_params;
    sp.use_gnc = 1;
    sp.known_inliers = (int32_t)std::min<size_t>(n_known, (size_t)F);
    sp.gnc_threshold_mode =
        g.gnc_threshold_mode_ == GncParams::GncThresholdMode::COST ? RPGO_GNC_COST : RPGO_GNC_PROBABILITY;
    sp.gnc_inlier_threshold = g.gnc_inlier_threshold_;
    sp.gnc_max_iterations = (int32_t)g.max_iterations_;
    sp.gnc_mu_step = g.mu_step_;
    sp.gnc_relative_cost_tol = g.relative_cost_tol_;
    sp.gnc_weights_tol = g.weights_tol_;
    sp.gnc_bias_odom = g.bias_odom_ ? 1 : 0;
  }
  std::vector<double> out((size_t)N * 12), w(F > 0 ? F : 1);
  rpgo_solve_report rep{};
  check(rpgo_solve(ctx_, &prob, &sp, out.data(), w.data(), &rep), "rpgo_solve");
  stats_.solve_ms = rep.ms_total;
  stats_.gnc_iterations = rep.gnc_iterations;
  stats_.lm_iterations = rep.lm_iterations;
  stats_.pcg_iterations = rep.pcg_iterations;
  // GenericSolver::updateValues (GenericSolver.cpp:37-45)
  for (int k = 0; k < N; ++k) {
    const double* o = &out[(size_t)k * 12];
    const Key key = keys[k];
    if (planar) {
      const Pose2 p(o[9], o[10], std::atan2(o[3], o[0]));
      if (values_.exists(key)) values_.update(key, p);
      else if (temp_values_.exists(key)) temp_values_.update(key, p);
    } else {
      Pose3 p;
      for (int e = 0; e < 9; ++e) p.R[e] = o[e];
      for (int e = 0; e < 3; ++e) p.t[e] = o[9 + e];
      if (values_.exists(key)) values_.update(key, p);
      else if (temp_values_.exists(key)) temp_values_.update(key, p);
    }
  }
  if (use_gnc) {
    gnc_weights_.assign(w.begin(), w.begin() + std::min<size_t>(nfg_.size(), (size_t)F));
    gnc_temp_weights_.assign(w.begin() + std::min<size_t>(nfg_.size(), (size_t)F), w.begin() + F);
    double sum = 0.0;
    for (double v : gnc_weights_) sum += v;
    gnc_num_inliers_ = (size_t)sum - n_known;
  } else {
    // loop-closure weights of the last GNC run carried over, everything else 1
    std::vector<double> lc_w;
    if (gnc_weights_.size() >= latest_num_lc_)
      lc_w.assign(gnc_weights_.end() - latest_num_lc_, gnc_weights_.end());
    gnc_weights_.assign(nfg_.size(), 1.0);
    if (lc_w.size() <= gnc_weights_.size())
      std::copy(lc_w.begin(), lc_w.end(), gnc_weights_.end() - lc_w.size());
    gnc_temp_weights_.assign(temp_nfg_.size(), 1.0);
  }
  if (pcm) latest_num_lc_ = outlier_removal_->getNumLC();
}

void RobustSolver::logStatus(long long spin_us) {
  std::ofstream f(params_.log_folder + "/rpgo_status.csv", std::ofstream::app);
  f << nfg_.size() << ',' << spin_us << ',' << getNumLC() << ',' << getNumLCInliers() << '\n';
  f.close();
  saveData(params_.log_folder);
}

// RobustSolver.cpp:337-369
void RobustSolver::update(const NonlinearFactorGraph& factors, const Values& values, bool optimize_graph) {
  const auto t0 = std::chrono::steady_clock::now();
  const bool do_optimize = outlier_removal_ ? outlier_removal_->removeOutliers(factors, values, &nfg_, &values_)
                                            : addAndCheckIfOptimize(factors, values);
  if (do_optimize && optimize_graph) optimize();
  if (outlier_removal_) {
    stats_.pcm_ms = static_cast<Pcm*>(outlier_removal_.get())->lastPcmMs();
    stats_.clique_ms = static_cast<Pcm*>(outlier_removal_.get())->lastCliqueMs();
    stats_.pcm_pairs = static_cast<Pcm*>(outlier_removal_.get())->lastPairs();
  }
  if (params_.log_output && optimize_graph)
    logStatus(std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count());
}

// RobustSolver.cpp:308-335
void RobustSolver::forceUpdate(const NonlinearFactorGraph& factors, const Values& values) {
  const auto t0 = std::chrono::steady_clock::now();
  if (outlier_removal_) outlier_removal_->removeOutliers(factors, values, &nfg_, &values_);
  else addAndCheckIfOptimize(factors, values);
  optimize();
  if (params_.log_output)
    logStatus(std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count());
}

void RobustSolver::loadGraph(const NonlinearFactorGraph& factors, const Values& values, const FactorPtr& prior) {
  NonlinearFactorGraph g = factors;
  if (prior) g.add(prior);
  update(g, values);
}

void RobustSolver::addGraph(const NonlinearFactorGraph& factors, const Values& values, const FactorPtr& between) {
  NonlinearFactorGraph g = factors;
  if (between) g.add(between);
  update(g, values);
}

size_t RobustSolver::getNumLC() const { return outlier_removal_ ? outlier_removal_->getNumLC() : 0; }

size_t RobustSolver::getNumLCInliers() const {
  if (params_.use_gnc_) return gnc_num_inliers_;
  return outlier_removal_ ? outlier_removal_->getNumLCInliers() : 0;
}

// RobustSolver.cpp:385-412
EdgePtr RobustSolver::removeLastLoopClosure(char prefix_1, char prefix_2) {
  EdgePtr removed;
  if (outlier_removal_) {
    removed = outlier_removal_->removeLastLoopClosure(prefix_1, prefix_2, &nfg_);
  } else if (!nfg_.empty()) {
    removed.reset(new Edge(nfg_[nfg_.size() - 1]->front(), nfg_[nfg_.size() - 1]->back()));
    nfg_.erase_last();
  }
  optimize();
  return removed;
}

EdgePtr RobustSolver::removeLastLoopClosure() {
  EdgePtr removed;
  if (outlier_removal_) {
    removed = outlier_removal_->removeLastLoopClosure(&nfg_);
  } else if (!nfg_.empty()) {
    removed.reset(new Edge(nfg_[nfg_.size() - 1]->front(), nfg_[nfg_.size() - 1]->back()));
    nfg_.erase_last();
  }
  optimize();
  return removed;
}

void RobustSolver::ignorePrefix(char prefix) {
  if (outlier_removal_) outlier_removal_->ignoreLoopClosureWithPrefix(prefix, &nfg_);
  else if (debug_) std::fprintf(stderr, "[rpgo] 'ignorePrefix' is not implemented without outlier rejection\n");
  optimize();
}

void RobustSolver::revivePrefix(char prefix) {
  if (outlier_removal_) outlier_removal_->reviveLoopClosureWithPrefix(prefix, &nfg_);
  else if (debug_) std::fprintf(stderr, "[rpgo] 'revivePrefix' is not implemented without outlier rejection\n");
  optimize();
}

std::vector<char> RobustSolver::getIgnoredPrefixes() const {
  return outlier_removal_ ? outlier_removal_->getIgnoredPrefixes() : std::vector<char>();
}

// RobustSolver.cpp:371-383 / GenericSolver.cpp removePriorsWithPrefix
void RobustSolver::removePriorFactorsWithPrefix(const char& prefix, bool optimize_graph) {
  if (outlier_removal_) {
    outlier_removal_->removePriorFactorsWithPrefix(prefix, &nfg_);
  } else {
    NonlinearFactorGraph s;
    for (const auto& f : nfg_)
      if (!(f->isPrior() && Symbol(f->front()).chr() == prefix)) s.add(f);
    nfg_ = s;
  }
  if (optimize_graph) optimize();
}

void RobustSolver::updateTempFactorsValues(const NonlinearFactorGraph& f, const Values& v) {
  temp_nfg_.add(f);
  temp_values_.insert(v);
}

void RobustSolver::replaceTempFactorsValues(const NonlinearFactorGraph& f, const Values& v) {
  temp_nfg_ = f;
  temp_values_ = v;
}

void RobustSolver::clearTempFactorsValues() {
  temp_nfg_ = NonlinearFactorGraph();
  temp_values_ = Values();
}

// RobustSolver.cpp:448-466
void RobustSolver::saveData(const std::string& folder_path) const {
  writeG2o(nfg_, values_, folder_path + "/result.g2o");
  if (outlier_removal_) outlier_removal_->saveData(folder_path);
  if (params_.use_gnc_) {
    std::ofstream f(folder_path + "/gnc_weights.csv");
    f.precision(17);
    for (size_t i = 0; i < gnc_weights_.size(); ++i) f << gnc_weights_[i] << '\n';
  }
}

}  // namespace kimera_rpgo_amd
