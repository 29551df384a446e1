"""Pin the landmark part of the FP64 oracle (orc_landmark_pair_dist,
orc_simple_landmark_pair, orc_pcm_adjacency_landmark).

Reference restated (paths relative to the reference tree):
  include/KimeraRPGO/outlier/Pcm.h:775-844   incrementLandmarkAdjMatrix: loop =
                                             (between(i, j) o Z_jl)^-1 o Z_il
  tests/testLandmark.cpp:22-183              PCM 3D (odom 5.0, lc 2.5): l0 seen
                                             from a1, a5 (consistent) and a4 (not);
                                             l1 from a2 and a5 (consistent)
  tests/testLandmark.cpp:186-345             the same with PcmSimple3D (0.3, 0.05)
plus the identities that fix the formula: it is the pose-pose pair check
(Pcm.h:670-718) with both loop closures ending on the same pose (the b -> d
odometry segment drops out), and exact observations give distance 0.
"""
import math

import numpy as np

import oracle as O

I3 = np.eye(3)
R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def landmark_cov():
    """Covariance of noiseModel::Diagonal::Precisions((0, 0, 0, 25, 25, 25)) as
    Gaussian::covariance() returns it: the unobserved rotation block is not a
    number, the translation block is 0.04 I."""
    c = np.full((6, 6), np.nan)
    c[3:, 3:] = 0.04 * np.eye(3)
    return c.reshape(36)


def chain_a():
    """a0..a5 of testLandmark.cpp:40-79: two steps +x, then three of Rz(90) and +x."""
    odom = [pose3(I3, (1, 0, 0))] * 2 + [pose3(R90, (1, 0, 0))] * 3
    cov = np.tile((0.1 * np.eye(6)).reshape(36), (5, 1))
    return O.fold(3, pose3(), odom, cov)


def _adj(traj, obs, simple=None, thr=2.5):
    frm = [k for k, _ in obs]
    zs = [z for _, z in obs]
    covs = [landmark_cov()] * len(obs)
    return O.pcm_adjacency_landmark(3, traj, frm, zs, covs, thr, simple=simple)


def test_reference_landmark_scenario():
    traj = chain_a()
    l0 = [(1, pose3(I3, (0, 1, 0))), (5, pose3(I3, (0, -1, 0))), (4, pose3(I3, (1, 0, 0)))]
    adj, dist = _adj(traj, l0)
    # a4 agrees with a1 alone (2.10 < 2.5), not with a5 (2.62): the clique
    # stays {a1, a5}, two landmark factors in the graph (testLandmark.cpp:153-163)
    assert adj.tolist() == [[0, 1, 1], [1, 0, 0], [1, 0, 0]]
    assert dist[1, 0] < 1e-9 and 2.0 < dist[2, 0] < 2.5 < dist[2, 1] < 2.7
    k, clique = O.max_clique_heu(adj)
    assert k == 2 and sorted(clique.tolist()) == [0, 1]
    l1 = [(2, pose3(I3, (0, -1, 0))), (5, pose3(I3, (2, 0, 0)))]
    adj, _ = _adj(traj, l1)
    assert adj.tolist() == [[0, 1], [1, 0]]


def test_reference_landmark_scenario_simple():
    traj = chain_a()
    l0 = [(1, pose3(I3, (0, 1, 0))), (5, pose3(I3, (0, -1, 0))), (4, pose3(I3, (1, 0, 0)))]
    adj, dist = _adj(traj, l0, simple=(0.3, 0.05))
    assert adj.tolist() == [[0, 1, 0], [1, 0, 0], [0, 0, 0]]
    # a1 vs a4: loop rotation pi about z, translation error 1 m in the plane:
    # Logmap translation (pi/2) |t|, over |4 - 1| + 2 nodes
    assert math.isclose(dist[2, 0], (math.pi / 2) / 5, rel_tol=1e-9)
    l1 = [(2, pose3(I3, (0, -1, 0))), (5, pose3(I3, (2, 0, 0)))]
    adj, _ = _adj(traj, l1, simple=(0.3, 0.05))
    assert adj.tolist() == [[0, 1], [1, 0]]


def _random_chain(rng, n):
    odom = []
    for _ in range(n - 1):
        w = rng.normal(0, 0.3, 3)
        th = np.linalg.norm(w)
        K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]) / max(th, 1e-12)
        R = I3 + math.sin(th) * K + (1 - math.cos(th)) * K @ K
        odom.append(pose3(R, rng.normal(0, 1, 3)))
    cov = np.tile(np.diag([1e-3] * 3 + [1e-2] * 3).reshape(36), (n - 1, 1))
    return O.fold(3, pose3(), odom, cov)


def test_equals_pair_check_with_shared_endpoint():
    """Pose-pose pair check (a, b), (c, d) with b = d: between(b, b) is the
    identity with zero covariance, so it equals the landmark check."""
    rng = np.random.default_rng(3)
    traj = _random_chain(rng, 40)
    for _ in range(30):
        i, j, b = (int(v) for v in rng.integers(0, 40, 3))
        zi, zj = pose3(I3, rng.normal(0, 2, 3)), pose3(I3, rng.normal(0, 2, 3))
        si = np.diag(rng.uniform(0.01, 0.2, 6)).reshape(36)
        sj = np.diag(rng.uniform(0.01, 0.2, 6)).reshape(36)
        ref = O.pair_dist(3, traj, i, b, zi, si, j, b, zj, sj)
        got = O.landmark_pair_dist(3, traj, i, zi, si, j, zj, sj)
        assert math.isclose(got, ref, rel_tol=1e-9, abs_tol=1e-12)
        t_ref, r_ref = O.simple_pair(3, traj, i, b, zi, si, j, b, zj, sj)
        t, r = O.simple_landmark_pair(3, traj, i, zi, si, j, zj, sj)
        assert math.isclose(t, t_ref, rel_tol=1e-9, abs_tol=1e-12)
        assert math.isclose(r, r_ref, rel_tol=1e-9, abs_tol=1e-12)


def test_exact_observations_have_zero_distance():
    rng = np.random.default_rng(5)
    traj = _random_chain(rng, 30)
    P = traj[:, :12]
    lw = rng.normal(0, 5, 3)  # landmark position in the world
    obs = []
    for k in (2, 9, 17, 25):
        R, t = P[k, :9].reshape(3, 3), P[k, 9:]
        obs.append((k, pose3(R.T, R.T @ (lw - t))))  # Z_kl = T_k^-1 T_l with T_l = (I, lw)
    adj, dist = _adj(traj, [(k, z) for k, z in obs], thr=1e-3)
    assert np.all(dist[~np.eye(len(obs), dtype=bool)] < 1e-6)
    assert adj.sum() == len(obs) * (len(obs) - 1)
    # a negative threshold turns the check off: empty adjacency
    adj, _ = _adj(traj, obs, thr=-1.0)
    assert adj.sum() == 0
