// RobustSolver facade tests, following the single-robot cases of the
// reference's tests: testGnc.cpp:25-105, testLoadGraph.cpp:25-60, :224-344,
// testTemporary.cpp:25-103 (same fixtures, parameters and expected counts and
// weights), plus checks of the Pose2 path, removeLastLoopClosure and
// ignorePrefix / revivePrefix on a single robot.
#include <memory>
#include <numeric>

#include "harness.h"
#include "kimera_rpgo_amd/Dataset.h"
#include "kimera_rpgo_amd/RobustSolver.h"

using namespace kimera_rpgo_amd;

namespace {

double sum(const std::vector<double>& v, size_t a = 0, size_t n = (size_t)-1) {
  double s = 0.0;
  for (size_t i = a; i < v.size() && i - a < n; ++i) s += v[i];
  return s;
}

FactorPtr prior_a0(const Values& values) {
  const Key k = Symbol('a', 0);
  return PriorFactorPose3(k, values.atPose3(k), noiseModel::Isotropic::Variance(6, 0.01));
}

}  // namespace

TEST(RobustSolver, GncDefault) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(100.0, 100.0, Verbosity::QUIET);
  params.setGncInlierCostThresholdsAtProbability(0.01);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  const auto w = pgo.getGncWeights();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(53));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
  EXPECT(w.size() == size_t(53));
  EXPECT(sum(w, 50, 3) == 0);  // every loop closure rejected
  EXPECT(sum(w) == 50);
  EXPECT(pgo.getNumLCInliers() == 0);
}

TEST(RobustSolver, GncHighThreshold) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(100.0, 100.0, Verbosity::QUIET);
  params.setGncInlierCostThresholdsAtProbability(0.99);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  const auto w = pgo.getGncWeights();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(53));
  EXPECT(w.size() == size_t(53));
  EXPECT(sum(w, 50, 3) == 3);  // every loop closure accepted
  EXPECT(sum(w) == 53);
  EXPECT(pgo.getNumLCInliers() == 3);
}

TEST(RobustSolver, Load1) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(0.0, 10.0, Verbosity::QUIET);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  // odometry threshold 0: only the odometry and the prior remain
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(50));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
}

TEST(RobustSolver, Load1NoPrior) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(0.0, 10.0, Verbosity::QUIET);
  RobustSolver pgo(params);
  pgo.update(gv.first, gv.second);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(49));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
}

TEST(RobustSolver, LoadGraphWithPrior) {
  // README.md:114-118: loadGraph(graph, values, prior) == update(graph + prior, values)
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(0.0, 10.0, Verbosity::QUIET);
  RobustSolver pgo(params);
  pgo.loadGraph(gv.first, gv.second, prior_a0(gv.second));
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(50));
}

TEST(RobustSolver, NoRejectLoad) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setNoRejection(Verbosity::QUIET);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(53));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
}

TEST(RobustSolver, NoRejectAdd) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setNoRejection(Verbosity::QUIET);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  auto gb = load3D(DATASET_PATH + "/robot_b.g2o");
  pgo.update(gb.first, gb.second);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(96));
  EXPECT(pgo.calculateEstimate().size() == size_t(92));
  NonlinearFactorGraph nf;
  nf.add(BetweenFactorPose3(Symbol('a', 1), Symbol('b', 1), Pose3(), noiseModel::Isotropic::Variance(6, 0.01)));
  pgo.update(nf, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(97));
  EXPECT(pgo.calculateEstimate().size() == size_t(92));
}

TEST(RobustSolver, TemporaryFactors) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(100.0, 100.0, Verbosity::QUIET);
  params.setGncInlierCostThresholdsAtProbability(0.01);
  RobustSolver pgo(params);
  pgo.update(gv.first, gv.second);
  auto w = pgo.getGncWeights();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(52));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
  EXPECT(w.size() == size_t(52));
  EXPECT(sum(w, 49, 3) == 0);
  EXPECT(sum(w) == 49);
  EXPECT(pgo.getNumLCInliers() == 0);
  Values tv;
  NonlinearFactorGraph tf;
  tv.insert(Symbol('b', 0), Pose3());
  tf.add(BetweenFactorPose3(Symbol('a', 0), Symbol('b', 0), Pose3(), noiseModel::Isotropic::Variance(6, 0.01)));
  pgo.updateTempFactorsValues(tf, tv);
  pgo.forceUpdate();
  w = pgo.getGncWeights();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(52));
  EXPECT(pgo.getTempFactorsUnsafe().size() == size_t(1));
  EXPECT(pgo.calculateEstimate().size() == size_t(50));
  EXPECT(w.size() == size_t(52));
  EXPECT(pgo.getNumLCInliers() == 0);
  pgo.clearTempFactorsValues();
  pgo.forceUpdate();
  w = pgo.getGncWeights();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(52));
  EXPECT(pgo.getTempFactorsUnsafe().size() == size_t(0));
  EXPECT(w.size() == size_t(52));
  EXPECT(sum(w, 49, 3) == 0);
  EXPECT(sum(w) == 49);
  EXPECT(pgo.getNumLCInliers() == 0);
}

TEST(RobustSolver, RemoveLastAndIgnorePrefixSingleRobot) {
  auto gv = load3D(DATASET_PATH + "/robot_a.g2o");
  RobustSolverParams params;
  params.setPcm3DParams(100.0, 100.0, Verbosity::QUIET);
  RobustSolver pgo(params);
  gv.first.add(prior_a0(gv.second));
  pgo.update(gv.first, gv.second);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(53));
  EXPECT(pgo.getNumLC() == 3);
  pgo.ignorePrefix('a');
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(50));
  EXPECT(pgo.getIgnoredPrefixes().size() == 1);
  pgo.revivePrefix('a');
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(53));
  EdgePtr e = pgo.removeLastLoopClosure();
  EXPECT(e != nullptr);
  // the file's last loop closure is a33 -> a27
  if (e) {
    EXPECT(e->from_key == Symbol('a', 33).key());
    EXPECT(e->to_key == Symbol('a', 27).key());
  }
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(52));
  // the loop-closure count keeps every loop closure taken in (Pcm.h:135, 299-340)
  EXPECT(pgo.getNumLC() == 3);
  pgo.removePriorFactorsWithPrefix('a', true);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(51));
}

// Pose2 chain around a square with exact measurements: the planar embedding
// must recover the ground truth from perturbed initial values.
TEST(RobustSolver, Pose2SquareRecoversGroundTruth) {
  RobustSolverParams params;
  params.setPcm2DParams(10.0, 10.0, Verbosity::QUIET);
  RobustSolver pgo(params);
  const auto noise = noiseModel::Diagonal::Sigmas({0.1, 0.1, 0.05});
  NonlinearFactorGraph g;
  Values v;
  std::vector<Pose2> truth;
  Pose2 cur;
  truth.push_back(cur);
  for (int k = 0; k < 12; ++k) {
    const Pose2 step = (k % 3 == 2) ? Pose2(1.0, 0.0, M_PI / 2) : Pose2(1.0, 0.0, 0.0);
    cur = cur.compose(step);
    truth.push_back(cur);
  }
  const Key a0 = Symbol('a', 0);
  v.insert(a0, Pose2(0.05, -0.03, 0.02));
  g.add(PriorFactorPose2(a0, Pose2(), noiseModel::Isotropic::Variance(3, 0.01)));
  for (int k = 0; k < 12; ++k) {
    const Pose2 z = truth[k].between(truth[k + 1]);
    g.add(BetweenFactorPose2(Symbol('a', k), Symbol('a', k + 1), z, noise));
    const double d = 0.02 * ((k % 5) - 2);
    v.insert(Symbol('a', k + 1), Pose2(truth[k + 1].x + d, truth[k + 1].y - d, truth[k + 1].theta + 0.5 * d));
  }
  // loop closures: pose 12 back onto pose 0, and 6 onto 2
  g.add(BetweenFactorPose2(Symbol('a', 12), a0, truth[12].between(truth[0]), noise));
  g.add(BetweenFactorPose2(Symbol('a', 6), Symbol('a', 2), truth[6].between(truth[2]), noise));
  pgo.update(g, v);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(15));
  const Values est = pgo.calculateEstimate();
  for (int k = 0; k <= 12; ++k) {
    const Pose2 p = est.atPose2(Symbol('a', k));
    EXPECT_DOUBLES_EQUAL(truth[k].x, p.x, 1e-6);
    EXPECT_DOUBLES_EQUAL(truth[k].y, p.y, 1e-6);
    EXPECT_DOUBLES_EQUAL(0.0, std::remainder(truth[k].theta - p.theta, 2 * M_PI), 1e-6);
  }
}

RPGO_TEST_MAIN
