"""Pin the PcmSimple (PoseWithNode) part of the FP64 oracle.

Reference tests restated (paths relative to the reference tree):
  tests/testPoseWithNode.cpp:16-98   node arithmetic and the per-node norms
  tests/testPcmSimple.cpp:18-98      odometry check, translation (0.05 / 0.052)
  tests/testPcmSimple.cpp:100-172    odometry check, rotation (0.0049 / 0.0052)
  tests/testPcmSimple.cpp:175-288    pairwise check, translation threshold 0.05
  tests/testPcmSimple.cpp:290-400    pairwise check, rotation threshold 0.005
plus closed forms on straight chains that fix the node counts
(|j-i|+1 for the odometry check, |c-a|+|d-b|+2 for a pair) and the Pose2
head/tail split of Logmap (GeometryUtils.h:275-289).
"""
import math

import numpy as np

import oracle as O

I3 = np.eye(3)
R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
R180 = np.array([[-1, 0, 0], [0, -1, 0], [0, 0, 1.0]])


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def cov6(v):
    return v * np.eye(6)


def chain(odoms):
    odom = np.array([pose3(R, t) for R, t in odoms])
    cov = np.array([cov6(0.1).reshape(36)] * len(odoms))
    return O.fold(3, pose3(), odom, cov)


# ------------------------------------------------------------------ PoseWithNode
def test_norms_kat():
    """testPoseWithNode.cpp:82-98."""
    assert O.pwn_norms(3, pose3(), 1) == (0.0, 0.0)
    t, _ = O.pwn_norms(3, pose3(I3, (1, 0, 0)), 5)
    assert t == 0.2
    Ry_pi = O.quat_to_rot(0, 0, 1, 0)
    _, r = O.pwn_norms(3, pose3(Ry_pi, (0, 0, 0)), 5)
    assert abs(r - 3.1415927 / 5) < 1e-6


def test_norms_rotation_info_off():
    """avg_rot_norm is 0 without rotation information (GeometryUtils.h:284)."""
    t, r = O.pwn_norms(3, pose3(rz(0.3), (0.5, 0, 0)), 2, rot_info=False)
    assert r == 0.0 and t > 0.0
    _, r1 = O.pwn_norms(3, pose3(rz(0.3), (0.5, 0, 0)), 2, rot_info=True)
    assert abs(r1 - 0.15) < 1e-15


def test_pose2_head_tail_split():
    """Pose2 Logmap is (v_x, v_y, theta): tail(2) = (v_y, theta), head(1) = v_x."""
    th, x, y = 0.4, 1.3, -0.7
    pose = np.array([math.cos(th), math.sin(th), x, y])
    xi = O.se2_logmap(pose)
    assert abs(xi[2] - th) < 1e-15
    t, r = O.pwn_norms(2, pose, 3)
    assert abs(t - math.hypot(xi[1], xi[2]) / 3) < 1e-15
    assert abs(r - abs(xi[0]) / 3) < 1e-15
    # pure translation: closed form, x alone is "rotation"
    t, r = O.pwn_norms(2, np.array([1.0, 0.0, 0.6, 0.8]), 2)
    assert t == 0.4 and r == 0.3


# ------------------------------------------------------------------ closed forms
def test_odom_node_count_straight_chain():
    """Unit steps along x, LC j -> i measuring (i - j + delta) along x: the
    residual is a translation delta, node |j - i| + 1."""
    traj = chain([(I3, (1, 0, 0))] * 9)
    for i, j, delta in [(0, 3, 0.2), (7, 2, -0.3), (4, 9, 0.05)]:
        z = pose3(I3, (j - i + delta, 0, 0))
        t, r = O.simple_odom(3, traj, i, j, z, cov6(0.1))
        assert abs(t - abs(delta) / (abs(j - i) + 1)) < 1e-14, (i, j, t)
        assert r == 0.0


def test_pair_node_count_straight_chain():
    """Two LCs (a,b) and (c,d) on a straight chain with x offsets e1, e2: the
    loop is a translation e2 - e1 (sign by composition), node |c-a|+|d-b|+2."""
    traj = chain([(I3, (1, 0, 0))] * 12)
    for (a, b, e1), (c, d, e2) in [((0, 5, 0.1), (2, 9, 0.3)), ((10, 1, -0.2), (4, 7, 0.05)),
                                   ((3, 3 + 4, 0.0), (11, 0, 0.4))]:
        zab = pose3(I3, (b - a + e1, 0, 0))
        zcd = pose3(I3, (d - c + e2, 0, 0))
        t, r = O.simple_pair(3, traj, a, b, zab, cov6(0.1), c, d, zcd, cov6(0.1))
        node = abs(c - a) + abs(d - b) + 2
        assert abs(t - abs(e1 - e2) / node) < 1e-14, (a, b, c, d, t)
        assert r == 0.0


def test_nan_rotation_covariance_zeroes_rot_norm():
    """A loop closure with NaN rotation variances (GeometryUtils.h:233-235)
    carries no rotation information: the rotation norm drops to 0."""
    traj = chain([(I3, (1, 0, 0))] * 4)
    cov = cov6(0.1)
    cov[0, 0] = np.nan
    z = pose3(rz(0.2), (3, 0, 0))
    _, r = O.simple_odom(3, traj, 0, 3, z, cov)
    assert r == 0.0
    _, r_ok = O.simple_odom(3, traj, 0, 3, z, cov6(0.1))
    assert abs(r_ok - 0.2 / 4) < 1e-15


# ------------------------------------------------------------------ reference tests
def _run(traj, lcs, thr):
    frm = [a for a, _, _ in lcs]
    to = [b for _, b, _ in lcs]
    zs = [z for _, _, z in lcs]
    covs = [cov6(0.1)] * len(lcs)
    acc, od, adj, dist = O.pcm_simple_adjacency(3, traj, frm, to, zs, covs, *thr)
    return acc, od, adj, dist


def _inliers(acc, adj, loop_check):
    if not loop_check:
        return len(acc)
    if len(acc) == 0:
        return 0
    k, _ = O.max_clique_heu(adj)
    return k


def test_odom_trans_check():
    """testPcmSimple.cpp:18-98: trans avg 0.05 passes 0.051, 0.052 fails."""
    traj = chain([(R90, (1, 0, 0))] * 3)
    thr = (0.051, 100.0, -1.0, -1.0)
    t1, r1 = O.simple_odom(3, traj, 3, 0, pose3(R90, (0.8, 0, 0)), cov6(0.1))
    t2, _ = O.simple_odom(3, traj, 3, 0, pose3(R90, (0.792, 0, 0)), cov6(0.05))
    assert abs(t1 - 0.05) < 1e-12 and t1 < 0.051
    assert 0.051 < t2 < 0.053
    acc, _, _, _ = _run(traj, [(3, 0, pose3(R90, (0.8, 0, 0)))], thr)
    assert 3 + 1 + _inliers(acc, None, False) == 5  # odometry + prior + inliers
    acc, _, _, _ = _run(traj, [(3, 0, pose3(R90, (0.8, 0, 0))), (3, 0, pose3(R90, (0.792, 0, 0)))], thr)
    assert list(acc) == [0]


def test_odom_rot_check():
    """testPcmSimple.cpp:100-172: rot avg 0.0049 passes 0.005, Rz(1.55) fails."""
    traj = chain([(R90, (1, 0, 0))] * 3)
    thr = (100.0, 0.005, -1.0, -1.0)
    _, r1 = O.simple_odom(3, traj, 3, 0, pose3(rz(1.551), (1, 0, 0)), cov6(0.1))
    _, r2 = O.simple_odom(3, traj, 3, 0, pose3(rz(1.55), (1, 0, 0)), cov6(0.1))
    assert abs(r1 - (math.pi / 2 - 1.551) / 4) < 1e-12 and r1 < 0.005
    assert r2 > 0.005
    acc, _, _, _ = _run(traj, [(3, 0, pose3(rz(1.551), (1, 0, 0))),
                               (3, 0, pose3(rz(1.55), (1, 0, 0)))], thr)
    assert list(acc) == [0]


def _consistency_chain():
    return chain([(R90, (1, 0, 0))] * 2 + [(I3, (1, 0, 0))] * 4)


def test_consistency_trans_check():
    """testPcmSimple.cpp:175-288: LC1..LC3 consistent (nfg 9, 10), LC4 stays out."""
    traj = _consistency_chain()
    thr = (-1.0, -1.0, 0.05, 100.0)
    lcs = [(3, 0, pose3(R180, (0, 0.9, 0))), (4, 0, pose3(rz(3.1416), (-0.9, 0.9, 0))),
           (5, 0, pose3(rz(3.1416), (-1.9, 0.8, 0))), (6, 0, pose3(rz(3.1416), (-2.8, 0.75, 0)))]
    sizes = []
    for m in (1, 2, 3, 4):
        acc, _, adj, _ = _run(traj, lcs[:m], thr)
        assert len(acc) == m  # odometry check disabled
        sizes.append(7 + _inliers(acc, adj, True))
    assert sizes[1:] == [9, 10, 10]


def test_consistency_rot_check():
    """testPcmSimple.cpp:290-400: rotation threshold 0.005."""
    traj = _consistency_chain()
    thr = (-1.0, -1.0, 100.0, 0.005)
    lcs = [(3, 0, pose3(R180, (0, 1.0, 0))), (4, 0, pose3(rz(3.141), (-1.0, 1.0, 0))),
           (5, 0, pose3(rz(3.13), (-2.0, 1.0, 0))), (6, 0, pose3(rz(3.12), (-3.0, 1.0, 0)))]
    sizes = []
    for m in (1, 2, 3, 4):
        acc, _, adj, _ = _run(traj, lcs[:m], thr)
        sizes.append(7 + _inliers(acc, adj, True))
    assert sizes[1:] == [9, 10, 10]
    # LC4 fails against the others on rotation, not translation
    _, _, adj, dist = _run(traj, lcs, thr)
    assert adj[3, :3].sum() < 3 and np.all(dist[3, :3] < 100.0)


def test_adjacency_matches_pairwise_calls():
    """The adjacency stage equals the single-pair calls it is built from,
    both decisions (translation and rotation) included."""
    rng = np.random.default_rng(3)
    n = 40
    odoms = [(rz(rng.normal(0, 0.1)), (1.0, rng.normal(0, 0.05), 0)) for _ in range(n - 1)]
    traj = chain(odoms)
    lcs = []
    for _ in range(25):
        a, b = rng.choice(n, 2, replace=False)
        lcs.append((int(a), int(b), pose3(rz(rng.normal(0, 0.3)), rng.normal(0, 2, 3))))
    thr = (0.8, 0.08, 0.3, 0.03)
    acc, od, adj, dist = _run(traj, lcs, thr)
    exp_acc = []
    for k, (a, b, z) in enumerate(lcs):
        t, r = O.simple_odom(3, traj, a, b, z, cov6(0.1))
        assert od[k] == t
        if t < thr[0] and r < thr[1]:
            exp_acc.append(k)
    assert list(acc) == exp_acc
    for j in range(len(acc)):
        for i in range(j):
            a, b, zab = lcs[acc[i]]
            c, d, zcd = lcs[acc[j]]
            t, r = O.simple_pair(3, traj, a, b, zab, cov6(0.1), c, d, zcd, cov6(0.1))
            assert dist[j, i] == t and dist[i, j] == t
            assert adj[j, i] == (t < thr[2] and r < thr[3])
