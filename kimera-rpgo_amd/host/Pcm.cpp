// PCM outlier-rejection plugin (see OutlierRemoval.h): host bookkeeping of
// outlier/Pcm.h:148-1005 for one robot prefix over the device stages
// rpgo_pcm_adjacency / rpgo_pcm_simple_adjacency (odometry check + pairwise
// consistency) and rpgo_max_clique (FMC::maxCliqueHeu).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "kimera_rpgo_amd/OutlierRemoval.h"
#include "rpgo_hip.h"

namespace kimera_rpgo_amd {

namespace {

void put_pose(const Factor& f, int dim, double* o) {
  if (dim == 3) {
    for (int e = 0; e < 9; ++e) o[e] = f.pose3.R[e];
    for (int e = 0; e < 3; ++e) o[9 + e] = f.pose3.t[e];
  } else {
    o[0] = std::cos(f.pose2.theta);
    o[1] = std::sin(f.pose2.theta);
    o[2] = f.pose2.x;
    o[3] = f.pose2.y;
  }
}

}  // namespace

Pcm::Pcm(int dim, const PcmParams& params, const std::vector<char>& special_symbols, const DeviceParams& device,
         bool simple)
    : dim_(dim),
      simple_(simple),
      // Pcm.h:74-82: any negative threshold of a check turns that check off
      odom_check_(!(params.odom_threshold < 0 || params.odom_rot_threshold < 0 ||
                    params.odom_trans_threshold < 0)),
      loop_check_(!(params.lc_threshold < 0 || params.dist_rot_threshold < 0 ||
                    params.dist_trans_threshold < 0)),
      params_(params),
      special_symbols_(special_symbols),
      device_(device) {
  if (dim != 2 && dim != 3) throw std::invalid_argument("Pcm: dim must be 2 or 3");
  check(rpgo_ctx_create(device.device, 0, &ctx_), "rpgo_ctx_create");
}

Pcm::~Pcm() {
  if (ctx_) rpgo_ctx_destroy(ctx_);
}

void Pcm::check(int rc, const char* what) const {
  if (rc != RPGO_OK) {
    const char* m = rpgo_last_error();
    throw std::runtime_error(std::string(what) + " failed: " + (m ? m : "?"));
  }
}

bool Pcm::isSpecialSymbol(char c) const {
  return std::find(special_symbols_.begin(), special_symbols_.end(), c) != special_symbols_.end();
}

// Pcm.h:516-557
void Pcm::updateOdom(const FactorPtr& f, const Values& values) {
  if ((dim_ == 3) != f->is3D()) throw std::invalid_argument("Pcm: odometry of the wrong pose type");
  const Symbol to(f->back());
  if (!have_traj_) {
    prefix_ = to.chr();
    first_key_ = f->front();
    have_traj_ = true;
    // the first pose comes from the values when the chain starts (Pcm.h:534-542)
    if (dim_ == 3) pose0_3_ = values.atPose3(f->front());
    else pose0_2_ = values.atPose2(f->front());
  } else if (to.chr() != prefix_) {
    throw std::invalid_argument("Pcm: odometry of a second robot prefix (multi-robot graphs are not "
                                "available in this build)");
  }
  if (f->front() != first_key_ + odom_.size())
    throw std::invalid_argument("Pcm: odometry does not extend the end of the chain");
  odom_.push_back(f);
}

// Pcm.h:148-281
bool Pcm::removeOutliers(const NonlinearFactorGraph& new_factors, const Values& new_values,
                         NonlinearFactorGraph* nfg, Values* values) {
  values->insert(new_values);
  if (new_factors.size() == 0) return false;
  bool do_optimize = false;
  std::vector<FactorPtr> lcs;
  for (const auto& f : new_factors) {
    if (!f) continue;
    if (f->isBetween()) {
      const Symbol a(f->front()), b(f->back());
      if (isSpecialSymbol(a.chr()) || isSpecialSymbol(b.chr()))
        throw std::invalid_argument("Pcm: landmark observations are not available in this build");
      if (f->front() + 1 == f->back() && new_values.exists(f->back())) {
        updateOdom(f, *values);
      } else if (f->front() != f->back()) {
        lcs.push_back(f);
      } else if (debug_) {
        std::fprintf(stderr, "[rpgo] attempting to close loop against self\n");
      }
    } else {
      nfg_special_.add(f);
      do_optimize = true;
    }
  }
  if (!lcs.empty()) {
    // Pcm.h:414-502
    std::vector<FactorPtr> pending;
    for (const auto& f : lcs) {
      if (!values->exists(f->front()) || !values->exists(f->back())) {
        if (debug_) std::fprintf(stderr, "[rpgo] cannot add loop closure with non-existing keys\n");
        continue;
      }
      if ((dim_ == 3) != f->is3D()) throw std::invalid_argument("Pcm: loop closure of the wrong pose type");
      if (Symbol(f->front()).chr() != Symbol(f->back()).chr() || Symbol(f->front()).chr() != prefix_)
        throw std::invalid_argument("Pcm: inter-robot loop closures are not available in this build");
      pending.push_back(f);
    }
    runPcm(pending);
    do_optimize = true;
  }
  *nfg = buildGraphToOptimize();
  return do_optimize;
}

// Odometry check of the new loop closures, pairwise consistency of every
// accepted pair and the max clique over the accepted set, all on the device.
// Loop closures accepted earlier pass the odometry check again unchanged (the
// chain of existing poses never changes), so recomputing the set gives the
// adjacency the reference grows row by row (Pcm.h:725-768).
void Pcm::runPcm(const std::vector<FactorPtr>& pending) {
  std::vector<FactorPtr> all(lc_);
  all.insert(all.end(), pending.begin(), pending.end());
  if (all.empty()) return;
  if (!have_traj_) throw std::invalid_argument("Pcm: loop closures before any odometry");
  const int plen = dim_ == 3 ? 12 : 4, d = dim_ == 3 ? 6 : 3;
  const int n = (int)odom_.size() + 1, m = (int)all.size();
  std::vector<double> pose0(plen), odom((size_t)(n - 1) * plen), ocov((size_t)(n - 1) * d * d);
  if (dim_ == 3) {
    for (int e = 0; e < 9; ++e) pose0[e] = pose0_3_.R[e];
    for (int e = 0; e < 3; ++e) pose0[9 + e] = pose0_3_.t[e];
  } else {
    pose0 = {std::cos(pose0_2_.theta), std::sin(pose0_2_.theta), pose0_2_.x, pose0_2_.y};
  }
  for (int k = 0; k + 1 < n; ++k) {
    put_pose(*odom_[k], dim_, &odom[(size_t)k * plen]);
    const auto& c = odom_[k]->noise->covariance();
    std::copy(c.begin(), c.end(), &ocov[(size_t)k * d * d]);
  }
  std::vector<int32_t> from(m), to(m);
  std::vector<double> lpose((size_t)m * plen), lcov((size_t)m * d * d);
  const std::uint64_t base = Symbol(first_key_).index();
  for (int t = 0; t < m; ++t) {
    const Factor& f = *all[t];
    const long long a = (long long)Symbol(f.front()).index() - (long long)base;
    const long long b = (long long)Symbol(f.back()).index() - (long long)base;
    if (a < 0 || b < 0 || a >= n || b >= n)
      throw std::invalid_argument("Pcm: loop closure outside the odometry chain");
    from[t] = (int32_t)a;
    to[t] = (int32_t)b;
    put_pose(f, dim_, &lpose[(size_t)t * plen]);
    const auto& c = f.noise->covariance();
    std::copy(c.begin(), c.end(), &lcov[(size_t)t * d * d]);
  }
  rpgo_pcm_problem prob{};
  prob.dim = dim_;
  prob.n_poses = n;
  prob.pose0 = pose0.data();
  prob.odom = odom.data();
  prob.odom_cov = ocov.data();
  prob.n_lc = m;
  prob.lc_from = from.data();
  prob.lc_to = to.data();
  prob.lc_pose = lpose.data();
  prob.lc_cov = lcov.data();
  rpgo_pcm_stats st{};
  std::vector<int32_t> accepted(m);
  if (simple_) {
    rpgo_pcm_simple_params sp{};
    sp.odom_trans_threshold = odom_check_ ? params_.odom_trans_threshold : -1.0;
    sp.odom_rot_threshold = odom_check_ ? params_.odom_rot_threshold : -1.0;
    sp.dist_trans_threshold = loop_check_ ? params_.dist_trans_threshold : -1.0;
    sp.dist_rot_threshold = loop_check_ ? params_.dist_rot_threshold : -1.0;
    check(rpgo_pcm_simple_adjacency(ctx_, &prob, &sp, accepted.data(), nullptr, nullptr, nullptr, &st),
          "rpgo_pcm_simple_adjacency");
  } else {
    rpgo_pcm_params pp{};
    pp.odom_threshold = odom_check_ ? params_.odom_threshold : -1.0;
    pp.lc_threshold = loop_check_ ? params_.lc_threshold : -1.0;
    pp.precision = device_.pcm_precision;
    check(rpgo_pcm_adjacency(ctx_, &prob, &pp, accepted.data(), nullptr, nullptr, nullptr, &st),
          "rpgo_pcm_adjacency");
  }
  lc_.clear();
  for (int k = 0; k < st.n_accepted; ++k) lc_.push_back(all[accepted[k]]);
  pcm_ms_ = st.ms_total;
  pairs_ = st.n_pairs;
  inliers_.clear();
  const int ma = (int)lc_.size();
  if (!loop_check_) {
    for (int k = 0; k < ma; ++k) inliers_.push_back(k);  // Pcm.h:870-873
  } else if (ma > 0) {
    std::vector<int32_t> idx(ma);
    int32_t k = 0;
    double ms = 0.0;
    check(rpgo_max_clique(ctx_, RPGO_CLIQUE_HEU_REF, idx.data(), &k, &ms), "rpgo_max_clique");
    clique_ms_ = ms;
    for (int q = 0; q < k; ++q) inliers_.push_back(idx[q]);
  }
}

// Pcm.h:977-1005: odometry, then special factors, then the inlier loop closures
NonlinearFactorGraph Pcm::buildGraphToOptimize() const {
  NonlinearFactorGraph g;
  for (const auto& f : odom_) g.add(f);
  g.add(nfg_special_);
  const bool ignored =
      std::find(ignored_prefixes_.begin(), ignored_prefixes_.end(), prefix_) != ignored_prefixes_.end();
  if (!ignored)
    for (int i : inliers_) g.add(lc_[i]);
  return g;
}

// Pcm.h:296-340
EdgePtr Pcm::removeLastLoopClosure(char prefix_1, char prefix_2, NonlinearFactorGraph* nfg) {
  if (lc_.empty() || prefix_1 != prefix_ || prefix_2 != prefix_) return nullptr;
  EdgePtr removed(new Edge(lc_.back()->front(), lc_.back()->back()));
  lc_.pop_back();
  if (lc_.size() < 2) {
    inliers_.clear();
    for (size_t k = 0; k < lc_.size(); ++k) inliers_.push_back((int)k);
  } else {
    std::vector<FactorPtr> keep;
    keep.swap(lc_);
    runPcm(keep);
  }
  *nfg = buildGraphToOptimize();
  return removed;
}

EdgePtr Pcm::removeLastLoopClosure(NonlinearFactorGraph* nfg) {
  if (lc_.empty()) return nullptr;
  return removeLastLoopClosure(prefix_, prefix_, nfg);
}

void Pcm::ignoreLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) {
  if (std::find(ignored_prefixes_.begin(), ignored_prefixes_.end(), prefix) == ignored_prefixes_.end())
    ignored_prefixes_.push_back(prefix);
  *nfg = buildGraphToOptimize();
}

void Pcm::reviveLoopClosureWithPrefix(char prefix, NonlinearFactorGraph* nfg) {
  ignored_prefixes_.erase(std::remove(ignored_prefixes_.begin(), ignored_prefixes_.end(), prefix),
                          ignored_prefixes_.end());
  *nfg = buildGraphToOptimize();
}

void Pcm::removePriorFactorsWithPrefix(const char& prefix, NonlinearFactorGraph* nfg) {
  NonlinearFactorGraph s;
  for (const auto& f : nfg_special_)
    if (!(f->isPrior() && Symbol(f->front()).chr() == prefix)) s.add(f);
  nfg_special_ = s;
  *nfg = buildGraphToOptimize();
}

}  // namespace kimera_rpgo_amd
