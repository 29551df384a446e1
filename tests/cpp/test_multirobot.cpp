// Multi-robot, landmark and frame-alignment tests of the facade, following
// the reference's tests (same graphs, thresholds and expected counts / poses):
//   testMultiRobot.cpp:22-312            inter- and intra-robot groups (PCM, PcmSimple)
//   testIgnorePrefix.cpp:22-258          ignorePrefix / revivePrefix on two robots
//   testRemoveLastLoopClosure.cpp:110-377 removal per prefix pair and in order
//   testLandmark.cpp:22-430              landmark groups (PCM, PcmSimple, no rejection)
//   testMultirobotFrameAlign.cpp:21-107  GNC pose averaging of robot frames
#include <cmath>
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

Pose3 rows(std::array<double, 9> R, double x, double y, double z) { return Pose3(R, {x, y, z}); }

Pose3 tr(double x, double y, double z) { return Pose3({1, 0, 0, 0, 1, 0, 0, 0, 1}, {x, y, z}); }

// Rot3(0, 0, 0, 1): the quaternion (w, x, y, z) = (0, 0, 0, 1), pi about z
Pose3 rpi(double x, double y, double z) { return Pose3::FromQuaternion(0, 0, 0, 1, x, y, z); }

Key A(int i) { return Symbol('a', i); }
Key B(int i) { return Symbol('b', i); }

bool near(const Pose3& a, const Pose3& b, double tol) { return a.equals(b, tol); }

// testMultiRobot.cpp:34-110: robots a and b, five odometry steps each
void two_robot_chains(RobustSolver& pgo) {
  Values init;
  init.insert(A(0), Pose3());
  init.insert(B(0), tr(0, -1, 0));
  pgo.update(NonlinearFactorGraph(), init);
  const auto noise = noiseModel::Isotropic::Variance(6, 0.1);
  for (int i = 0; i < 3; ++i) {
    Values v;
    NonlinearFactorGraph f;
    v.insert(A(i + 1), tr(1, 0, 0));
    v.insert(B(i + 1), tr(1, 0, 0));
    f.add(BetweenFactorPose3(A(i), A(i + 1), tr(1, 0, 0), noise));
    f.add(BetweenFactorPose3(B(i), B(i + 1), tr(1, 0, 0), noise));
    pgo.update(f, v);
  }
  for (int i = 3; i < 5; ++i) {
    Values v;
    NonlinearFactorGraph f;
    const Pose3 o = rows({0, -1, 0, 1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v.insert(A(i + 1), o);
    f.add(BetweenFactorPose3(A(i), A(i + 1), o, noise));
    pgo.update(f, v);
  }
  for (int i = 3; i < 5; ++i) {
    Values v;
    NonlinearFactorGraph f;
    const Pose3 o = rows({0, 1, 0, -1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v.insert(B(i + 1), o);
    f.add(BetweenFactorPose3(B(i), B(i + 1), o, noise));
    pgo.update(f, v);
  }
}

void two_robot_loops(int& rpgo_failures_, RobustSolver& pgo) {
  const auto noise = noiseModel::Isotropic::Variance(6, 0.1);
  NonlinearFactorGraph lc;
  lc.add(BetweenFactorPose3(B(1), A(1), tr(0, 1, 0), noise));
  lc.add(BetweenFactorPose3(B(4), A(4), rows({-1, 0, 0, 0, -1, 0, 0, 0, 1}, -1, 0, 0), noise));
  pgo.update(lc, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(pgo.calculateEstimate().size() == size_t(12));
  lc = NonlinearFactorGraph();
  lc.add(BetweenFactorPose3(A(2), B(2), tr(0, -1, 0), noise));
  lc.add(BetweenFactorPose3(B(5), A(5), tr(0, -3.3, 0), noise));
  pgo.update(lc, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(pgo.calculateEstimate().size() == size_t(12));
}

// testIgnorePrefix.cpp:22-104 / testRemoveLastLoopClosure.cpp:22-107
void build_test_graph(NonlinearFactorGraph* f, Values* v) {
  v->insert(A(0), Pose3());
  f->add(PriorFactorPose3(A(0), Pose3(), noiseModel::Isotropic::Variance(6, 10e-8)));
  v->insert(B(0), tr(0, -1, 0));
  f->add(PriorFactorPose3(B(0), tr(0, -1, 0), noiseModel::Isotropic::Variance(6, 1)));
  const auto noise = noiseModel::Isotropic::Variance(6, 10);
  for (int i = 0; i < 3; ++i) {
    v->insert(A(i + 1), tr(1, 0, 0));
    v->insert(B(i + 1), tr(1, 0, 0));
    f->add(BetweenFactorPose3(A(i), A(i + 1), tr(1, 0, 0), noise));
    f->add(BetweenFactorPose3(B(i), B(i + 1), tr(1, 0, 0), noise));
  }
  for (int i = 3; i < 5; ++i) {
    const Pose3 o = rows({0, -1, 0, 1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v->insert(A(i + 1), o);
    f->add(BetweenFactorPose3(A(i), A(i + 1), o, noise));
  }
  for (int i = 3; i < 5; ++i) {
    const Pose3 o = rows({0, 1, 0, -1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v->insert(B(i + 1), o);
    f->add(BetweenFactorPose3(B(i), B(i + 1), o, noise));
  }
}

NonlinearFactorGraph test_graph_loops() {
  const auto noise = noiseModel::Isotropic::Variance(6, 10e-8);
  NonlinearFactorGraph f;
  f.add(BetweenFactorPose3(B(0), A(0), Pose3(), noise));
  f.add(BetweenFactorPose3(A(5), A(3), rpi(1, 0, 0), noise));
  return f;
}

// shared start of the ignore / remove tests: 12 factors, then the two loops
void ignore_remove_start(int& rpgo_failures_, RobustSolver& pgo) {
  NonlinearFactorGraph f;
  Values v;
  build_test_graph(&f, &v);
  pgo.update(f, v);
  Values est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(est.size() == size_t(12));
  EXPECT(near(tr(0, -1, 0), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 1, 0), est.atPose3(A(5)), 10e-6));
  pgo.update(test_graph_loops());
  est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(14));
  EXPECT(est.size() == size_t(12));
  EXPECT(near(Pose3(), est.atPose3(B(0)), 10e-6));
  EXPECT(near(rpi(4, 0, 0), est.atPose3(A(5)), 10e-6));
}

// testLandmark.cpp:34-102: robot a, loop closures a3 -> a2 (good), a4 -> a1 (bad)
void landmark_chain(int& rpgo_failures_, RobustSolver& pgo) {
  Values init;
  init.insert(A(0), Pose3());
  pgo.update(NonlinearFactorGraph(), init);
  const auto noise = noiseModel::Isotropic::Variance(6, 0.1);
  for (int i = 0; i < 2; ++i) {
    Values v;
    NonlinearFactorGraph f;
    v.insert(A(i + 1), tr(1, 0, 0));
    f.add(BetweenFactorPose3(A(i), A(i + 1), tr(1, 0, 0), noise));
    pgo.update(f, v);
  }
  for (int i = 2; i < 5; ++i) {
    Values v;
    NonlinearFactorGraph f;
    const Pose3 o = rows({0, -1, 0, 1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v.insert(A(i + 1), o);
    f.add(BetweenFactorPose3(A(i), A(i + 1), o, noise));
    pgo.update(f, v);
  }
  NonlinearFactorGraph lc;
  lc.add(BetweenFactorPose3(A(3), A(2), rz(-1.57, 0, 0.9, 0), noise));
  lc.add(BetweenFactorPose3(A(4), A(1), rz(3.14, 2.1, 1.1, 2.5), noise));
  pgo.update(lc, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(6));
  EXPECT(pgo.calculateEstimate().size() == size_t(6));
}

// testLandmark.cpp:104-182: l0 seen from a1, a5 (consistent), a4 (not); l1
// from a2 and a5.  Landmark noise: precisions 0 (rotation) and 25.
void landmark_observations(int& rpgo_failures_, RobustSolver& pgo, size_t n9_after_l1_first) {
  const auto lmk = noiseModel::Diagonal::Precisions({0, 0, 0, 25, 25, 25});
  const Key l0 = Symbol('l', 0), l1 = Symbol('l', 1);
  NonlinearFactorGraph f;
  Values v;
  f.add(BetweenFactorPose3(A(1), l0, tr(0, 1, 0), lmk));
  v.insert(l0, tr(1, 1, 0));
  pgo.update(f, v);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(7));
  EXPECT(pgo.calculateEstimate().size() == size_t(7));
  f = NonlinearFactorGraph();
  f.add(BetweenFactorPose3(A(5), l0, tr(0, -1, 0), lmk));
  pgo.update(f, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(8));
  EXPECT(pgo.calculateEstimate().size() == size_t(7));
  f = NonlinearFactorGraph();
  f.add(BetweenFactorPose3(A(4), l0, tr(1, 0, 0), lmk));
  pgo.update(f, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(8));
  EXPECT(pgo.calculateEstimate().size() == size_t(7));
  f = NonlinearFactorGraph();
  v = Values();
  f.add(BetweenFactorPose3(A(2), l1, tr(0, -1, 0), lmk));
  v.insert(l1, tr(1, 1, 0));
  pgo.update(f, v);
  EXPECT(pgo.getFactorsUnsafe().size() == n9_after_l1_first);
  EXPECT(pgo.calculateEstimate().size() == size_t(8));
  f = NonlinearFactorGraph();
  f.add(BetweenFactorPose3(A(5), l1, tr(2, 0, 0), lmk));
  pgo.update(f, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == n9_after_l1_first + 1);
  EXPECT(pgo.calculateEstimate().size() == size_t(8));
}

}  // namespace

TEST(MultiRobot, Pcm) {
  RobustSolverParams params;
  params.setPcm3DParams(3.0, 0.05, Verbosity::QUIET);
  params.setMultiRobotAlignMethod(MultiRobotAlignMethod::GNC);
  RobustSolver pgo(params);
  two_robot_chains(pgo);
  two_robot_loops(rpgo_failures_, pgo);
  // intra-robot: a4 -> a1 good, a4 -> a2 bad
  const auto noise = noiseModel::Isotropic::Variance(6, 0.1);
  NonlinearFactorGraph lc;
  lc.add(BetweenFactorPose3(A(4), A(1), rows({0, 1, 0, -1, 0, 0, 0, 0, 1}, 0, 3, 0), noise));
  lc.add(BetweenFactorPose3(A(4), A(2), rz(-1.54, 0, 2, 0), noise));
  pgo.update(lc, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(13));
  EXPECT(pgo.calculateEstimate().size() == size_t(12));
}

TEST(MultiRobot, PcmSimple) {
  RobustSolverParams params;
  params.setPcmSimple3DParams(0.04, 0.01, Verbosity::QUIET);
  params.setMultiRobotAlignMethod(MultiRobotAlignMethod::GNC);
  RobustSolver pgo(params);
  two_robot_chains(pgo);
  two_robot_loops(rpgo_failures_, pgo);
}

TEST(IgnorePrefix, SinglePrefixA) {
  RobustSolverParams params;
  params.setPcm3DParams(100, 100, Verbosity::QUIET);
  RobustSolver pgo(params);
  ignore_remove_start(rpgo_failures_, pgo);
  pgo.ignorePrefix('a');
  Values est = pgo.calculateEstimate();
  std::vector<char> ig = pgo.getIgnoredPrefixes();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(est.size() == size_t(12));
  EXPECT(near(tr(0, -1, 0), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 1, 0), est.atPose3(A(5)), 10e-6));
  EXPECT(ig.size() == size_t(1) && ig[0] == 'a');
  pgo.revivePrefix('a');
  est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(14));
  EXPECT(near(Pose3(), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 0, 0), est.atPose3(A(5)), 10e-6));
  EXPECT(pgo.getIgnoredPrefixes().empty());
}

TEST(IgnorePrefix, SinglePrefixB) {
  RobustSolverParams params;
  params.setPcm3DParams(100, 100, Verbosity::QUIET);
  RobustSolver pgo(params);
  ignore_remove_start(rpgo_failures_, pgo);
  pgo.ignorePrefix('b');
  Values est = pgo.calculateEstimate();
  std::vector<char> ig = pgo.getIgnoredPrefixes();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(13));
  EXPECT(near(tr(0, -1, 0), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 0, 0), est.atPose3(A(5)), 10e-6));
  EXPECT(ig.size() == size_t(1) && ig[0] == 'b');
  pgo.revivePrefix('b');
  est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(14));
  EXPECT(near(Pose3(), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 0, 0), est.atPose3(A(5)), 10e-6));
}

TEST(RemoveLastLoopClosure, Pcm) {
  RobustSolverParams params;
  params.setPcm3DParams(100, 100, Verbosity::QUIET);
  RobustSolver pgo(params);
  ignore_remove_start(rpgo_failures_, pgo);
  EdgePtr e1 = pgo.removeLastLoopClosure('a', 'a');
  Values est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(13));
  EXPECT(e1 && e1->from_key == A(5) && e1->to_key == A(3));
  EXPECT(near(Pose3(), est.atPose3(B(0)), 10e-6));
  EXPECT(near(rpi(4, 1, 0), est.atPose3(A(5)), 10e-6));
  EdgePtr e2 = pgo.removeLastLoopClosure('a', 'b');
  est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(e2 && e2->from_key == B(0) && e2->to_key == A(0));
  EXPECT(near(tr(0, -1, 0), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 1, 0), est.atPose3(A(5)), 10e-6));
}

TEST(RemoveLastLoopClosure, PcmInOrder) {
  RobustSolverParams params;
  params.setPcm3DParams(100, 100, Verbosity::QUIET);
  RobustSolver pgo(params);
  ignore_remove_start(rpgo_failures_, pgo);
  EdgePtr e1 = pgo.removeLastLoopClosure();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(13));
  EXPECT(e1 && e1->from_key == A(5) && e1->to_key == A(3));
  EdgePtr e2 = pgo.removeLastLoopClosure();
  Values est = pgo.calculateEstimate();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(e2 && e2->from_key == B(0) && e2->to_key == A(0));
  EXPECT(near(tr(0, -1, 0), est.atPose3(B(0)), 0.01));
  EXPECT(near(rpi(4, 1, 0), est.atPose3(A(5)), 10e-6));
  EdgePtr e3 = pgo.removeLastLoopClosure();
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(12));
  EXPECT(e3 == nullptr);
}

TEST(Landmark, Pcm) {
  RobustSolverParams params;
  params.setPcm3DParams(5.0, 2.5, Verbosity::QUIET);
  params.specialSymbols = {'l'};
  RobustSolver pgo(params);
  landmark_chain(rpgo_failures_, pgo);
  landmark_observations(rpgo_failures_, pgo, 9);
}

TEST(Landmark, PcmSimple) {
  RobustSolverParams params;
  params.setPcmSimple3DParams(0.3, 0.05, Verbosity::QUIET);
  params.specialSymbols = {'l'};
  RobustSolver pgo(params);
  landmark_chain(rpgo_failures_, pgo);
  landmark_observations(rpgo_failures_, pgo, 9);
}

// testLandmark.cpp:347-430: without outlier rejection every factor stays
TEST(Landmark, NoReject) {
  RobustSolverParams params;
  params.setNoRejection(Verbosity::QUIET);
  params.specialSymbols = {'l'};
  RobustSolver pgo(params);
  Values init;
  init.insert(A(0), Pose3());
  pgo.update(NonlinearFactorGraph(), init);
  const auto noise = noiseModel::Isotropic::Variance(6, 0.1);
  for (int i = 0; i < 5; ++i) {
    Values v;
    NonlinearFactorGraph f;
    const Pose3 o = i < 2 ? tr(1, 0, 0) : rows({0, -1, 0, 1, 0, 0, 0, 0, 1}, 1, 0, 0);
    v.insert(A(i + 1), o);
    f.add(BetweenFactorPose3(A(i), A(i + 1), o, noise));
    pgo.update(f, v);
  }
  NonlinearFactorGraph lc;
  lc.add(BetweenFactorPose3(A(3), A(2), rz(-1.57, 0, 0.9, 0), noise));
  lc.add(BetweenFactorPose3(A(4), A(1), rz(3.14, 2.1, 1.1, 2.5), noise));
  pgo.update(lc, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(7));
  const auto lmk = noiseModel::Diagonal::Precisions({0, 0, 0, 25, 25, 25});
  const Key l0 = Symbol('l', 0);
  NonlinearFactorGraph f;
  Values v;
  f.add(BetweenFactorPose3(A(1), l0, tr(0, 1, 0), lmk));
  v.insert(l0, tr(1, 1, 0));
  pgo.update(f, v);
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(8));
  EXPECT(pgo.calculateEstimate().size() == size_t(7));
  f = NonlinearFactorGraph();
  f.add(BetweenFactorPose3(A(4), l0, tr(1, 0, 0), lmk));
  pgo.update(f, Values());
  EXPECT(pgo.getFactorsUnsafe().size() == size_t(9));
}

TEST(Pcm, MultirobotFrameAlign) {
  PcmParams params;
  params.odom_trans_threshold = -1;
  params.odom_rot_threshold = -1;
  params.dist_trans_threshold = 0.3;
  params.dist_rot_threshold = 100.0;
  std::unique_ptr<OutlierRemoval> pcm(new Pcm3D(params, MultiRobotAlignMethod::GNC));
  pcm->setQuiet();
  const auto noise = noiseModel::Isotropic::Variance(6, 0.01);
  const Key C0 = Symbol('c', 0), C1 = Symbol('c', 1);
  NonlinearFactorGraph nfg;
  Values est;
  est.insert(A(0), Pose3());
  nfg.add(BetweenFactorPose3(A(0), A(1), Pose3(), noise));
  est.insert(A(1), Pose3());
  est.insert(B(0), tr(1, 0, 0));
  nfg.add(BetweenFactorPose3(B(0), B(1), tr(1, 0, 0), noise));
  est.insert(B(1), tr(2, 0, 0));
  est.insert(C0, rpi(0, 0, 0));
  nfg.add(BetweenFactorPose3(C0, C1, tr(1, 0, 0), noise));
  est.insert(C1, rpi(-1, 0, 0));
  nfg.add(BetweenFactorPose3(A(0), B(0), Pose3(), noise));
  nfg.add(BetweenFactorPose3(A(1), B(1), tr(1, 0, 0), noise));
  nfg.add(BetweenFactorPose3(A(0), C0, Pose3(), noise));
  nfg.add(BetweenFactorPose3(A(1), C1, tr(1, 0, 0), noise));
  NonlinearFactorGraph out_nfg;
  Values out;
  pcm->removeOutliers(nfg, est, &out_nfg, &out);
  EXPECT(near(Pose3(), out.atPose3(A(0)), 1e-6));
  EXPECT(near(Pose3(), out.atPose3(A(1)), 1e-6));
  EXPECT(near(Pose3(), out.atPose3(B(0)), 1e-6));
  EXPECT(near(tr(1, 0, 0), out.atPose3(B(1)), 1e-6));
  EXPECT(near(Pose3(), out.atPose3(C0), 1e-6));
  EXPECT(near(tr(1, 0, 0), out.atPose3(C1), 1e-6));
}

RPGO_TEST_MAIN
