// PcmSimple plugin tests, following the reference's tests/testPcmSimple.cpp:18-400
// (same graphs, thresholds and expected factor / value counts), plus a Pose2
// chain through RobustSolver with PCM_Simple2D.
#include <memory>

#include "harness.h"
#include "kimera_rpgo_amd/OutlierRemoval.h"
#include "kimera_rpgo_amd/RobustSolver.h"

using namespace kimera_rpgo_amd;

namespace {

Pose3 rz(double a, double x, double y, double z) {
  const double c = std::cos(a), s = std::sin(a);
  Pose3 p;
  p.R = {c, -s, 0, s, c, 0, 0, 0, 1};
  p.t = {x, y, z};
  return p;
}

// R rows (0,-1,0), (1,0,0), (0,0,1), written out exactly
Pose3 r90(double x) {
  Pose3 p;
  p.R = {0, -1, 0, 1, 0, 0, 0, 0, 1};
  p.t = {x, 0, 0};
  return p;
}

Pose3 r180(double x, double y) {
  Pose3 p;
  p.R = {-1, 0, 0, 0, -1, 0, 0, 0, 1};
  p.t = {x, y, 0};
  return p;
}

void start(OutlierRemoval* pcm, NonlinearFactorGraph* nfg, Values* est) {
  Values init_vals;
  NonlinearFactorGraph init_factors;
  init_vals.insert(0, Pose3());
  init_factors.add(PriorFactorPose3(0, Pose3(), noiseModel::Isotropic::Variance(6, 0.01)));
  pcm->removeOutliers(init_factors, init_vals, nfg, est);
}

void odom(OutlierRemoval* pcm, size_t i, const Pose3& p, NonlinearFactorGraph* nfg, Values* est) {
  Values v;
  NonlinearFactorGraph f;
  v.insert(i + 1, p);
  f.add(BetweenFactorPose3(i, i + 1, p, noiseModel::Isotropic::Variance(6, 0.1)));
  pcm->removeOutliers(f, v, nfg, est);
}

bool lc(OutlierRemoval* pcm, size_t a, size_t b, const Pose3& p, double var, NonlinearFactorGraph* nfg,
        Values* est) {
  NonlinearFactorGraph f;
  f.add(BetweenFactorPose3(a, b, p, noiseModel::Isotropic::Variance(6, var)));
  return pcm->removeOutliers(f, Values(), nfg, est);
}

}  // namespace

TEST(PcmSimple, OdomTransCheck) {
  PcmParams params;
  params.dist_trans_threshold = -1;
  params.dist_rot_threshold = -1;
  params.odom_trans_threshold = 0.051;
  params.odom_rot_threshold = 100.0;
  std::unique_ptr<OutlierRemoval> pcm(new PcmSimple3D(params));
  pcm->setQuiet();
  NonlinearFactorGraph nfg;
  Values est;
  start(pcm.get(), &nfg, &est);
  for (size_t i = 0; i < 3; i++) odom(pcm.get(), i, r90(1), &nfg, &est);
  EXPECT(size_t(4) == nfg.size());
  EXPECT(size_t(4) == est.size());
  // translation per node 0.05
  bool do_optimize = lc(pcm.get(), 3, 0, r90(0.8), 0.1, &nfg, &est);
  EXPECT(size_t(5) == nfg.size());
  EXPECT(size_t(4) == est.size());
  EXPECT(do_optimize == true);
  // translation per node 0.052
  do_optimize = lc(pcm.get(), 3, 0, r90(0.792), 0.05, &nfg, &est);
  EXPECT(size_t(5) == nfg.size());
  EXPECT(size_t(4) == est.size());
  EXPECT(do_optimize == true);
}

TEST(PcmSimple, OdomRotCheck) {
  PcmParams params;
  params.dist_trans_threshold = -1;
  params.dist_rot_threshold = -1;
  params.odom_trans_threshold = 100.0;
  params.odom_rot_threshold = 0.005;
  std::unique_ptr<OutlierRemoval> pcm(new PcmSimple3D(params));
  pcm->setQuiet();
  NonlinearFactorGraph nfg;
  Values est;
  start(pcm.get(), &nfg, &est);
  for (size_t i = 0; i < 3; i++) odom(pcm.get(), i, r90(1), &nfg, &est);
  EXPECT(size_t(4) == nfg.size());
  // rotation per node 0.0049
  bool do_optimize = lc(pcm.get(), 3, 0, rz(1.551, 1, 0, 0), 0.1, &nfg, &est);
  EXPECT(size_t(5) == nfg.size());
  EXPECT(do_optimize == true);
  // rotation per node 0.0052
  do_optimize = lc(pcm.get(), 3, 0, rz(1.55, 1, 0, 0), 0.1, &nfg, &est);
  EXPECT(size_t(5) == nfg.size());
  EXPECT(size_t(4) == est.size());
  EXPECT(do_optimize == true);
}

namespace {

void consistency_case(const PcmParams& params, const Pose3 lcs[4], int& rpgo_failures_) {
  std::unique_ptr<OutlierRemoval> pcm(new PcmSimple3D(params));
  pcm->setQuiet();
  NonlinearFactorGraph nfg;
  Values est;
  start(pcm.get(), &nfg, &est);
  for (size_t i = 0; i < 2; i++) odom(pcm.get(), i, r90(1), &nfg, &est);
  for (size_t i = 2; i < 6; i++) odom(pcm.get(), i, Pose3({1, 0, 0, 0, 1, 0, 0, 0, 1}, {1, 0, 0}), &nfg, &est);
  EXPECT(size_t(7) == nfg.size());
  EXPECT(size_t(7) == est.size());
  lc(pcm.get(), 3, 0, lcs[0], 0.1, &nfg, &est);
  bool do_optimize = lc(pcm.get(), 4, 0, lcs[1], 0.1, &nfg, &est);
  EXPECT(size_t(9) == nfg.size());
  EXPECT(size_t(7) == est.size());
  EXPECT(do_optimize == true);
  do_optimize = lc(pcm.get(), 5, 0, lcs[2], 0.1, &nfg, &est);
  EXPECT(size_t(10) == nfg.size());
  EXPECT(do_optimize == true);
  // inconsistent: passes the (disabled) odometry check, not the pairwise one
  do_optimize = lc(pcm.get(), 6, 0, lcs[3], 0.1, &nfg, &est);
  EXPECT(size_t(10) == nfg.size());
  EXPECT(size_t(7) == est.size());
  EXPECT(do_optimize == true);
  EXPECT(pcm->getNumLC() == 4);
  EXPECT(pcm->getNumLCInliers() == 3);
}

}  // namespace

TEST(PcmSimple, ConsistencyTransCheck) {
  PcmParams params;
  params.odom_trans_threshold = -1;
  params.odom_rot_threshold = -1;
  params.dist_trans_threshold = 0.05;
  params.dist_rot_threshold = 100.0;
  const Pose3 lcs[4] = {r180(0, 0.9), rz(3.1416, -0.9, 0.9, 0), rz(3.1416, -1.9, 0.8, 0),
                        rz(3.1416, -2.8, 0.75, 0)};
  consistency_case(params, lcs, rpgo_failures_);
}

TEST(PcmSimple, ConsistencyRotCheck) {
  PcmParams params;
  params.odom_trans_threshold = -1;
  params.odom_rot_threshold = -1;
  params.dist_trans_threshold = 100.0;
  params.dist_rot_threshold = 0.005;
  const Pose3 lcs[4] = {r180(0, 1.0), rz(3.141, -1.0, 1.0, 0), rz(3.13, -2.0, 1.0, 0),
                        rz(3.12, -3.0, 1.0, 0)};
  consistency_case(params, lcs, rpgo_failures_);
}

// A straight Pose2 chain with one exact and one far-off loop closure through
// the solver facade (setPcmSimple2DParams, the batch examples' outlier method).
TEST(PcmSimple, Pose2SolverRejectsBadLoop) {
  RobustSolverParams params;
  params.setPcmSimple2DParams(0.05, 0.05, Verbosity::QUIET);
  RobustSolver solver(params);
  NonlinearFactorGraph f;
  Values v;
  const auto noise = noiseModel::Isotropic::Variance(3, 0.01);
  v.insert(0, Pose2(0, 0, 0));
  f.add(PriorFactorPose2(0, Pose2(0, 0, 0), noise));
  for (int i = 0; i < 8; ++i) {
    v.insert(i + 1, Pose2(i + 1, 0, 0));
    f.add(BetweenFactorPose2(i, i + 1, Pose2(1, 0, 0), noise));
  }
  solver.update(f, v);
  NonlinearFactorGraph lcs;
  lcs.add(BetweenFactorPose2(8, 0, Pose2(-8, 0, 0), noise));   // exact
  lcs.add(BetweenFactorPose2(6, 1, Pose2(5, 3, 1.0), noise));  // far off
  solver.update(lcs, Values());
  EXPECT(solver.getFactorsUnsafe().size() == size_t(1 + 8 + 1));
  EXPECT(std::abs(solver.calculateEstimate().atPose2(8).x - 8.0) < 1e-6);
}

RPGO_TEST_MAIN
