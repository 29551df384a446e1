// PCM outlier-rejection plugin (see OutlierRemoval.h): host bookkeeping of
// outlier/Pcm.h:148-1131 (robots, loop-closure groups, landmarks, frame
// alignment) over the device stages rpgo_pcm_adjacency /
// rpgo_pcm_simple_adjacency (odometry check + pairwise consistency),
// rpgo_max_clique[_incremental] (FMC::maxCliqueHeu) and rpgo_solve (GNC pose
// averaging).
#include <algorithm>
#include <cfloat>
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

// Pose2 (x, y, theta) as the planar Pose3 (Rz(theta), (x, y, 0))
Pose3 embed(const Pose2& p) {
  const double c = std::cos(p.theta), s = std::sin(p.theta);
  Pose3 q;
  q.R = {c, -s, 0, s, c, 0, 0, 0, 1};
  q.t = {p.x, p.y, 0.0};
  return q;
}

Pose2 unembed(const Pose3& q) { return Pose2(q.t[0], q.t[1], std::atan2(q.R[3], q.R[0])); }

}  // namespace

Pcm::Pcm(int dim, const PcmParams& params, const std::vector<char>& special_symbols, const DeviceParams& device,
         bool simple)
    : Pcm(dim, params, MultiRobotAlignMethod::NONE, 0.99, special_symbols, device, simple) {}

Pcm::Pcm(int dim, const PcmParams& params, MultiRobotAlignMethod align, double align_gnc_probability,
         const std::vector<char>& special_symbols, const DeviceParams& device, bool simple)
    : dim_(dim),
      simple_(simple),
      // Pcm.h:74-82: any negative threshold of a check turns that check off
      odom_check_(!(params.odom_threshold < 0 || params.odom_rot_threshold < 0 ||
                    params.odom_trans_threshold < 0)),
      loop_check_(!(params.lc_threshold < 0 || params.dist_rot_threshold < 0 ||
                    params.dist_trans_threshold < 0)),
      params_(params),
      align_(align),
      align_prob_(align_gnc_probability),
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

int Pcm::chainOf(char prefix) const {
  for (size_t c = 0; c < chains_.size(); ++c)
    if (chains_[c].prefix == prefix) return (int)c;
  return -1;
}

// ObservationId equality (TypeUtils.h:44-58): the unordered prefix pair
int Pcm::groupOf(char a, char b) const {
  for (size_t g = 0; g < groups_.size(); ++g)
    if ((groups_[g].id1 == a && groups_[g].id2 == b) || (groups_[g].id1 == b && groups_[g].id2 == a))
      return (int)g;
  return -1;
}

// Pcm.h:516-557
void Pcm::updateOdom(const FactorPtr& f, const Values& values) {
  if ((dim_ == 3) != f->is3D()) throw std::invalid_argument("Pcm: odometry of the wrong pose type");
  const char prefix = Symbol(f->back()).chr();
  int c = chainOf(prefix);
  if (c < 0) {
    // a robot seen for the first time starts at the value of its first key
    Chain ch;
    ch.prefix = prefix;
    ch.first_key = f->front();
    if (dim_ == 3) ch.pose0_3 = values.atPose3(f->front());
    else ch.pose0_2 = values.atPose2(f->front());
    chains_.push_back(ch);
    c = (int)chains_.size() - 1;
  }
  Chain& ch = chains_[c];
  if (f->front() != ch.first_key + ch.odom.size())
    throw std::invalid_argument("Pcm: odometry does not extend the end of its robot's chain");
  ch.odom.push_back(f);
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
      if ((dim_ == 3) != f->is3D()) throw std::invalid_argument("Pcm: between factor of the wrong pose type");
      const Symbol a(f->front()), b(f->back());
      if (isSpecialSymbol(a.chr()) || isSpecialSymbol(b.chr())) {
        if (new_values.exists(f->front()) || new_values.exists(f->back())) {
          // first observation: the landmark's group starts (anew) with it
          const Key lkey = isSpecialSymbol(a.chr()) ? f->front() : f->back();
          auto it = std::find_if(landmarks_.begin(), landmarks_.end(),
                                 [&](const Landmark& l) { return l.key == lkey; });
          if (it == landmarks_.end()) it = landmarks_.insert(landmarks_.end(), Landmark{lkey, {}, {}});
          it->obs.assign(1, f);
          it->inliers.assign(1, 0);
          ++total_lc_;
        } else {
          lcs.push_back(f);  // re-observation: checked as a loop closure
        }
      } else if (f->front() + 1 == f->back() && new_values.exists(f->back())) {
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
    // Pcm.h:414-502: sort the loop closures into their groups
    std::vector<std::vector<FactorPtr>> pending(groups_.size());
    std::vector<int> lc_group;  // group of each loop closure kept, arrival order (-1: landmark)
    std::vector<int> dirty_landmarks;
    for (const auto& f : lcs) {
      if (!values->exists(f->front()) || !values->exists(f->back())) {
        if (debug_) std::fprintf(stderr, "[rpgo] cannot add loop closure with non-existing keys\n");
        continue;
      }
      const Symbol a(f->front()), b(f->back());
      if (isSpecialSymbol(a.chr()) || isSpecialSymbol(b.chr())) {
        if (isSpecialSymbol(a.chr()))
          throw std::invalid_argument("Pcm: landmark observations must run pose -> landmark");
        auto it = std::find_if(landmarks_.begin(), landmarks_.end(),
                               [&](const Landmark& l) { return l.key == f->back(); });
        if (it == landmarks_.end())
          throw std::invalid_argument("Pcm: re-observation of a landmark that was never initialised");
        if (!it->obs.empty() && Symbol(it->obs[0]->front()).chr() != a.chr())
          throw std::invalid_argument("Pcm: a landmark must be observed from one robot's trajectory");
        it->obs.push_back(f);
        ++total_lc_;
        const int li = (int)(it - landmarks_.begin());
        if (std::find(dirty_landmarks.begin(), dirty_landmarks.end(), li) == dirty_landmarks.end())
          dirty_landmarks.push_back(li);
        lc_group.push_back(-1);
        continue;
      }
      if (chainOf(a.chr()) < 0 || chainOf(b.chr()) < 0)
        throw std::invalid_argument("Pcm: loop closure on a robot without odometry");
      int g = groupOf(a.chr(), b.chr());
      if (g < 0) {
        Group ng;
        ng.id1 = a.chr();
        ng.id2 = b.chr();
        groups_.push_back(ng);
        pending.emplace_back();
        g = (int)groups_.size() - 1;
      }
      pending[g].push_back(f);
      lc_group.push_back(g);
    }
    // device stages per group; kept[g]: pending loop closures of g that passed
    std::vector<std::vector<bool>> kept(groups_.size());
    for (size_t g = 0; g < groups_.size(); ++g) {
      if (pending[g].empty()) continue;
      const size_t before = groups_[g].lc.size();
      runGroup(groups_[g], pending[g], params_.incremental);
      // accepted pending ones keep their order after the earlier factors
      kept[g].assign(pending[g].size(), false);
      size_t q = before;
      for (size_t t = 0; t < pending[g].size() && q < groups_[g].lc.size(); ++t)
        if (groups_[g].lc[q] == pending[g][t]) kept[g][t] = true, ++q;
    }
    // Pcm.h:476-486: chronological order of the loop closures taken in
    std::vector<size_t> next(groups_.size(), 0);
    for (int g : lc_group) {
      if (g < 0) continue;
      const size_t t = next[g]++;
      if (kept[g][t]) {
        lc_order_.push_back(g);
        ++total_lc_;
      }
    }
    for (int li : dirty_landmarks) runLandmark(landmarks_[li]);
    countInliers();
    do_optimize = true;
  }
  *nfg = buildGraphToOptimize();
  if (align_ != MultiRobotAlignMethod::NONE && chains_.size() > 1) *values = multirobotValueInitialization(*values);
  return do_optimize;
}

// Odometry check of the pending loop closures, pairwise consistency of every
// accepted pair of the group and the max clique over the accepted set, all
// on the device.  Loop closures accepted earlier pass the odometry check again
// unchanged (the chain poses of existing keys never change), so recomputing
// the set gives the adjacency the reference grows row by row (Pcm.h:725-768).
int Pcm::runGroup(Group& g, const std::vector<FactorPtr>& pending, bool incremental) {
  std::vector<FactorPtr> all(g.lc);
  all.insert(all.end(), pending.begin(), pending.end());
  const int m_before = (int)g.lc.size();
  if (all.empty()) {
    g.inliers.clear();
    return 0;
  }
  const int plen = dim_ == 3 ? 12 : 4, d = dim_ == 3 ? 6 : 3;
  // the group's chains: one robot, or two (id1's first)
  std::vector<int> cs{chainOf(g.id1)};
  if (g.id2 != g.id1) cs.push_back(chainOf(g.id2));
  for (int c : cs)
    if (c < 0) throw std::invalid_argument("Pcm: loop closure on a robot without odometry");
  std::vector<int32_t> chain_off{0};
  std::vector<double> pose0s, odom, ocov;
  for (int c : cs) {
    const Chain& ch = chains_[c];
    chain_off.push_back(chain_off.back() + (int32_t)ch.odom.size() + 1);
    if (dim_ == 3) {
      for (int e = 0; e < 9; ++e) pose0s.push_back(ch.pose0_3.R[e]);
      for (int e = 0; e < 3; ++e) pose0s.push_back(ch.pose0_3.t[e]);
    } else {
      const Pose2& p = ch.pose0_2;
      pose0s.insert(pose0s.end(), {std::cos(p.theta), std::sin(p.theta), p.x, p.y});
    }
    for (const auto& f : ch.odom) {
      const size_t o = odom.size();
      odom.resize(o + plen);
      put_pose(*f, dim_, &odom[o]);
      const auto& cv = f->noise->covariance();
      ocov.insert(ocov.end(), cv.begin(), cv.end());
    }
  }
  auto index_of = [&](Key k) -> int32_t {
    const Symbol s(k);
    for (size_t q = 0; q < cs.size(); ++q) {
      const Chain& ch = chains_[cs[q]];
      if (ch.prefix != s.chr()) continue;
      const long long j = (long long)s.index() - (long long)Symbol(ch.first_key).index();
      if (j < 0 || j > (long long)ch.odom.size())
        throw std::invalid_argument("Pcm: loop closure outside its robot's odometry chain");
      return chain_off[q] + (int32_t)j;
    }
    throw std::invalid_argument("Pcm: loop closure outside its group's robots");
  };
  const int m = (int)all.size();
  std::vector<int32_t> from(m), to(m);
  std::vector<double> lpose((size_t)m * plen), lcov;
  for (int t = 0; t < m; ++t) {
    const Factor& f = *all[t];
    from[t] = index_of(f.front());
    to[t] = index_of(f.back());
    put_pose(f, dim_, &lpose[(size_t)t * plen]);
    const auto& cv = f.noise->covariance();
    lcov.insert(lcov.end(), cv.begin(), cv.end());
  }
  rpgo_pcm_problem prob{};
  prob.dim = dim_;
  prob.n_poses = chain_off.back();
  prob.pose0 = pose0s.data();
  prob.odom = odom.data();
  prob.odom_cov = ocov.data();
  prob.n_lc = m;
  prob.lc_from = from.data();
  prob.lc_to = to.data();
  prob.lc_pose = lpose.data();
  prob.lc_cov = lcov.data();
  prob.n_chains = (int32_t)cs.size();
  prob.chain_off = cs.size() > 1 ? chain_off.data() : nullptr;
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
  g.lc.clear();
  for (int k = 0; k < st.n_accepted; ++k) g.lc.push_back(all[accepted[k]]);
  pcm_ms_ = st.ms_total;
  pairs_ = st.n_pairs;
  const int ma = (int)g.lc.size();
  const int num_new = ma - m_before;
  if (!loop_check_) {
    g.inliers.clear();
    for (int k = 0; k < ma; ++k) g.inliers.push_back(k);  // Pcm.h:870-873
  } else if (ma > 0 && incremental) {
    // Pcm.h:906-951: a strictly larger clique among the new loop closures
    // replaces the previous inliers, else they stay
    if (num_new > 0) {
      std::vector<int32_t> idx(ma);
      int32_t k = 0;
      check(rpgo_max_clique_incremental(ctx_, num_new, (int32_t)g.inliers.size(), idx.data(), &k),
            "rpgo_max_clique_incremental");
      if (k > 0) g.inliers.assign(idx.begin(), idx.begin() + k);
    }
  } else if (ma > 0) {
    std::vector<int32_t> idx(ma);
    int32_t k = 0;
    double ms = 0.0;
    check(rpgo_max_clique(ctx_, RPGO_CLIQUE_HEU_REF, idx.data(), &k, &ms), "rpgo_max_clique");
    clique_ms_ = ms;
    g.inliers.assign(idx.begin(), idx.begin() + std::max(k, 0));
  } else {
    g.inliers.clear();
  }
  return num_new;
}

// incrementLandmarkAdjMatrix (Pcm.h:775-844) over all observations, then the
// landmark's max clique (Pcm.h:877-896).  The landmark check compares with
// lc_threshold / dist_*_threshold directly: a negative one leaves the
// adjacency empty and the clique is the first observation.
void Pcm::runLandmark(Landmark& l) {
  const int m = (int)l.obs.size();
  if (m <= 1) {
    l.inliers.assign(m, 0);
    return;
  }
  const char prefix = Symbol(l.obs[0]->front()).chr();
  const int c = chainOf(prefix);
  if (c < 0) throw std::invalid_argument("Pcm: landmark observed from a robot without odometry");
  const Chain& ch = chains_[c];
  const int plen = dim_ == 3 ? 12 : 4;
  std::vector<double> pose0(plen), odom(ch.odom.size() * plen), ocov;
  if (dim_ == 3) {
    for (int e = 0; e < 9; ++e) pose0[e] = ch.pose0_3.R[e];
    for (int e = 0; e < 3; ++e) pose0[9 + e] = ch.pose0_3.t[e];
  } else {
    pose0 = {std::cos(ch.pose0_2.theta), std::sin(ch.pose0_2.theta), ch.pose0_2.x, ch.pose0_2.y};
  }
  for (size_t k = 0; k < ch.odom.size(); ++k) {
    put_pose(*ch.odom[k], dim_, &odom[k * plen]);
    const auto& cv = ch.odom[k]->noise->covariance();
    ocov.insert(ocov.end(), cv.begin(), cv.end());
  }
  std::vector<int32_t> from(m);
  std::vector<double> lpose((size_t)m * plen), lcov;
  for (int t = 0; t < m; ++t) {
    const long long j = (long long)Symbol(l.obs[t]->front()).index() - (long long)Symbol(ch.first_key).index();
    if (j < 0 || j > (long long)ch.odom.size())
      throw std::invalid_argument("Pcm: landmark observed from a pose outside the odometry chain");
    from[t] = (int32_t)j;
    put_pose(*l.obs[t], dim_, &lpose[(size_t)t * plen]);
    const auto& cv = l.obs[t]->noise->covariance();
    lcov.insert(lcov.end(), cv.begin(), cv.end());
  }
  rpgo_pcm_problem prob{};
  prob.dim = dim_;
  prob.n_poses = (int32_t)ch.odom.size() + 1;
  prob.pose0 = pose0.data();
  prob.odom = odom.data();
  prob.odom_cov = ocov.data();
  prob.n_lc = m;
  prob.lc_from = from.data();
  prob.lc_to = nullptr;
  prob.lc_pose = lpose.data();
  prob.lc_cov = lcov.data();
  prob.landmark = 1;
  rpgo_pcm_stats st{};
  std::vector<int32_t> accepted(m);
  if (simple_) {
    rpgo_pcm_simple_params sp{};
    sp.odom_trans_threshold = -1.0;
    sp.odom_rot_threshold = -1.0;
    sp.dist_trans_threshold = params_.dist_trans_threshold;
    sp.dist_rot_threshold = params_.dist_rot_threshold;
    check(rpgo_pcm_simple_adjacency(ctx_, &prob, &sp, accepted.data(), nullptr, nullptr, nullptr, &st),
          "rpgo_pcm_simple_adjacency");
  } else {
    rpgo_pcm_params pp{};
    pp.odom_threshold = -1.0;
    pp.lc_threshold = params_.lc_threshold;
    pp.precision = device_.pcm_precision;
    check(rpgo_pcm_adjacency(ctx_, &prob, &pp, accepted.data(), nullptr, nullptr, nullptr, &st),
          "rpgo_pcm_adjacency");
  }
  std::vector<int32_t> idx(m);
  int32_t k = 0;
  double ms = 0.0;
  check(rpgo_max_clique(ctx_, RPGO_CLIQUE_HEU_REF, idx.data(), &k, &ms), "rpgo_max_clique");
  l.inliers.assign(idx.begin(), idx.begin() + std::max(k, 0));
}

// Pcm.h:851-970: inliers of every group and landmark
void Pcm::countInliers() {
  total_good_lc_ = 0;
  for (const auto& g : groups_) total_good_lc_ += g.inliers.size();
  for (const auto& l : landmarks_) total_good_lc_ += l.inliers.size();
}

// Pcm.h:977-1005: odometry, then special factors, then the inliers of every
// group whose robots are not ignored, then the landmark inliers
NonlinearFactorGraph Pcm::buildGraphToOptimize() const {
  NonlinearFactorGraph out;
  for (const auto& f : odom_) out.add(f);
  out.add(nfg_special_);
  auto ignored = [&](char p) {
    return std::find(ignored_prefixes_.begin(), ignored_prefixes_.end(), p) != ignored_prefixes_.end();
  };
  for (const auto& g : groups_)
    if (!ignored(g.id1) && !ignored(g.id2))
      for (int i : g.inliers) out.add(g.lc[i]);
  for (const auto& l : landmarks_)
    for (int i : l.inliers) out.add(l.obs[i]);
  return out;
}

// Pcm.h:1080-1089: a robot's chain poses, composed from its start, after T
void Pcm::robotValues(const Chain& c, const Pose3* T3, const Pose2* T2, Values* out) const {
  Pose3 p3 = c.pose0_3;
  Pose2 p2 = c.pose0_2;
  for (size_t k = 0; k <= c.odom.size(); ++k) {
    const Key key = c.first_key + k;
    if (k > 0) {
      if (dim_ == 3) p3 = p3.compose(c.odom[k - 1]->pose3);
      else p2 = p2.compose(c.odom[k - 1]->pose2);
    }
    if (dim_ == 3) {
      const Pose3 v = T3 ? T3->compose(p3) : p3;
      if (out->exists(key)) out->update(key, v);
      else out->insert(key, v);
    } else {
      const Pose2 v = T2 ? T2->compose(p2) : p2;
      if (out->exists(key)) out->update(key, v);
      else out->insert(key, v);
    }
  }
}

// Pcm.h:1011-1131: the first robot (smallest prefix) keeps its frame; every
// other robot's chain is moved by the GNC (or L2) pose average of the frame
// transforms T_w0_front Z T_wi_back^-1 of its inter-robot inliers with the
// first robot.  The average is a one-pose problem with one prior factor per
// transform (sigmas 0.1 rotation, 0.5 translation), GncOptimizer's TLS at the
// given probability (or an infinite threshold for L2), from the identity.
Values Pcm::multirobotValueInitialization(const Values& values) const {
  Values out = values;
  std::vector<int> order(chains_.size());
  for (size_t c = 0; c < order.size(); ++c) order[c] = (int)c;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return chains_[a].prefix < chains_[b].prefix; });
  const Chain& c0 = chains_[order[0]];
  robotValues(c0, nullptr, nullptr, &out);
  // chain pose of a key (composed odometry from the robot's start)
  auto chain_pose = [&](const Chain& c, Key key) {
    const long long j = (long long)Symbol(key).index() - (long long)Symbol(c.first_key).index();
    if (j < 0 || j > (long long)c.odom.size()) throw std::invalid_argument("Pcm: key outside its chain");
    Pose3 p = dim_ == 3 ? c.pose0_3 : embed(c.pose0_2);
    for (long long k = 0; k < j; ++k) p = p.compose(dim_ == 3 ? c.odom[k]->pose3 : embed(c.odom[k]->pose2));
    return p;
  };
  for (size_t q = 1; q < order.size(); ++q) {
    const Chain& ci = chains_[order[q]];
    const int gi = groupOf(c0.prefix, ci.prefix);
    if (gi < 0) {
      if (debug_)
        std::fprintf(stderr, "[rpgo] no inter-robot loop closures between %c and %c for frame alignment\n",
                     c0.prefix, ci.prefix);
      continue;
    }
    const Group& g = groups_[gi];
    std::vector<Pose3> T;
    for (int i : g.inliers) {
      const Factor& f = *g.lc[i];
      Key front = f.front(), back = f.back();
      Pose3 z = dim_ == 3 ? f.pose3 : embed(f.pose2);
      if (Symbol(front).chr() != c0.prefix) {
        std::swap(front, back);
        z = z.inverse();
      }
      T.push_back(chain_pose(c0, front).compose(z).compose(chain_pose(ci, back).inverse()));
    }
    // one pose, one prior per transform (planar problems embedded in SE(3))
    const int F = (int)T.size();
    Pose3 est;
    if (F > 0) {
      std::vector<double> pose(12), meas((size_t)F * 12), info((size_t)F * 36, 0.0);
      std::vector<int32_t> kind(F, RPGO_FACTOR_PRIOR), ki(F, 0), kj(F, 0);
      for (int e = 0; e < 9; ++e) pose[e] = (e % 4 == 0) ? 1.0 : 0.0;
      for (int t = 0; t < F; ++t) {
        for (int e = 0; e < 9; ++e) meas[(size_t)t * 12 + e] = T[t].R[e];
        for (int e = 0; e < 3; ++e) meas[(size_t)t * 12 + 9 + e] = T[t].t[e];
        double* I = &info[(size_t)t * 36];
        for (int e = 0; e < 3; ++e) I[e * 7] = 1.0 / (0.1 * 0.1);
        for (int e = 3; e < 6; ++e) I[e * 7] = 1.0 / (0.5 * 0.5);
        if (dim_ == 2) I[0] = I[7] = I[35] = 1.0;  // out-of-plane directions
      }
      rpgo_solve_problem prob{};
      prob.n_poses = 1;
      prob.poses = pose.data();
      prob.n_factors = F;
      prob.kind = kind.data();
      prob.key_i = ki.data();
      prob.key_j = kj.data();
      prob.measured = meas.data();
      prob.information = info.data();
      rpgo_solve_params sp;
      rpgo_solve_params_default(&sp);
      sp.use_gnc = 1;
      sp.known_inliers = 0;
      if (align_ == MultiRobotAlignMethod::L2) {
        sp.gnc_threshold_mode = RPGO_GNC_COST;
        sp.gnc_inlier_threshold = DBL_MAX;
      } else {
        sp.gnc_threshold_mode = RPGO_GNC_PROBABILITY;
        sp.gnc_inlier_threshold = align_prob_;
        sp.gnc_factor_dim = dim_ == 3 ? 6 : 3;
        sp.pose2_chart = dim_ == 3 ? 0 : 1;
      }
      std::vector<double> res(12);
      rpgo_solve_report rep{};
      check(rpgo_solve(ctx_, &prob, &sp, res.data(), nullptr, &rep), "rpgo_solve (frame alignment)");
      for (int e = 0; e < 9; ++e) est.R[e] = res[e];
      for (int e = 0; e < 3; ++e) est.t[e] = res[9 + e];
    }
    if (dim_ == 3) {
      robotValues(ci, &est, nullptr, &out);
    } else {
      const Pose2 e2 = unembed(est);
      robotValues(ci, nullptr, &e2, &out);
    }
  }
  return out;
}

// Pcm.h:296-340.  The reference sizes the removal by the group's adjacency,
// which is only grown while the pairwise check is on: with it off nothing is
// removed.
EdgePtr Pcm::removeLastLoopClosure(char prefix_1, char prefix_2, NonlinearFactorGraph* nfg) {
  const int gi = groupOf(prefix_1, prefix_2);
  if (gi < 0 || !loop_check_) return nullptr;
  Group& g = groups_[gi];
  if (g.lc.empty()) return nullptr;
  EdgePtr removed(new Edge(g.lc.back()->front(), g.lc.back()->back()));
  g.lc.pop_back();
  if (g.lc.size() < 2) {
    g.inliers.clear();
    for (size_t k = 0; k < g.lc.size(); ++k) g.inliers.push_back((int)k);
  } else {
    std::vector<FactorPtr> keep;
    keep.swap(g.lc);
    runGroup(g, keep, false);
  }
  *nfg = buildGraphToOptimize();
  return removed;
}

// Pcm.h:345-352: the group of the last loop closure taken in
EdgePtr Pcm::removeLastLoopClosure(NonlinearFactorGraph* nfg) {
  if (lc_order_.empty()) return nullptr;
  const int gi = lc_order_.back();
  lc_order_.pop_back();
  return removeLastLoopClosure(groups_[gi].id1, groups_[gi].id2, nfg);
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
