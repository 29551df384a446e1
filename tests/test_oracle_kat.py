"""Pin the FP64 oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, invariants, brute force).

Reference tests restated (paths relative to the reference tree):
  tests/testPoseWithCovariance.cpp:32-217   compose / between / inverse KATs
  tests/testTrajectory.cpp:35-109           getBetween == stitched compose
  tests/testPoseWithNode.cpp:84-98          Log at the identity and exactly pi
  tests/testPcm.cpp:17-192                  odometry check 0.3, pairwise 0.5
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O

I3 = np.eye(3)


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def rodrigues(w):
    w = np.asarray(w, float)
    th = np.linalg.norm(w)
    if th < 1e-15:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])


# ------------------------------------------------------------ testPoseWithCovariance
def test_pwc_compose_kat():
    """testPoseWithCovariance.cpp:51-149 (B_covar, C_covar, D_covar exact)."""
    A = O.pwc_raw(3, pose3(t=(1, 1, 1)), 0.1 * np.eye(6))
    AB = O.pwc_raw(3, pose3(t=(1, 1, 1)), 0.1 * np.eye(6))
    B = O.pwc_compose(3, A, AB)
    Bp, Bc, _ = O.pwc_split(3, B)
    np.testing.assert_allclose(Bp, pose3(t=(2, 2, 2)), atol=1e-12)
    B_cov = 0.1 * np.array([[2, 0, 0, 0, -1, 1], [0, 2, 0, 1, 0, -1], [0, 0, 2, -1, 1, 0],
                            [0, 1, -1, 4, -1, -1], [-1, 0, 1, -1, 4, -1], [1, -1, 0, -1, -1, 4]])
    np.testing.assert_allclose(Bc, B_cov, atol=1e-9)

    bc_cov = 0.1 * np.eye(6)
    bc_cov[3:, 3:] = 0.01 * np.eye(3)
    BC = O.pwc_raw(3, pose3(O.quat_to_rot(0, 0, 0, 1)), bc_cov)
    C = O.pwc_compose(3, B, BC)
    C_cov = np.array([[0.3, 0, 0, 0, -0.1, -0.1], [0, 0.3, 0, 0.1, 0, 0.1], [0, 0, 0.3, 0.1, -0.1, 0],
                      [0, 0.1, 0.1, 0.41, -0.1, 0.1], [-0.1, 0, -0.1, -0.1, 0.41, 0.1],
                      [-0.1, 0.1, 0, 0.1, 0.1, 0.41]])
    np.testing.assert_allclose(O.pwc_split(3, C)[1], C_cov, atol=1e-9)

    CD = O.pwc_raw(3, pose3(O.quat_to_rot(0, 0, 1, 0), (1, 0, 0)), 0.1 * np.eye(6))
    D = O.pwc_compose(3, C, CD)
    D_cov = np.array([[0.4, 0, 0, 0, 0.1, -0.1], [0, 0.4, 0, -0.1, 0, 0.2], [0, 0, 0.4, 0.1, -0.2, 0],
                      [0, -0.1, 0.1, 0.51, 0, 0], [0.1, 0, -0.2, 0, 0.61, -0.1],
                      [-0.1, 0.2, 0, 0, -0.1, 0.61]])
    np.testing.assert_allclose(O.pwc_split(3, D)[1], D_cov, atol=1e-9)


def test_pwc_between_kat():
    """testPoseWithCovariance.cpp:168-200: between of A and C gives 0.1*I."""
    A_cov = np.array([[0.3, 0, 0, 0, -0.1, -0.1], [0, 0.3, 0, 0.1, 0, 0.1], [0, 0, 0.3, 0.1, -0.1, 0],
                      [0, 0.1, 0.1, 0.41, -0.1, 0.1], [-0.1, 0, -0.1, -0.1, 0.41, 0.1],
                      [-0.1, 0.1, 0, 0.1, 0.1, 0.41]])
    C_cov = np.array([[0.4, 0, 0, 0, 0.1, -0.1], [0, 0.4, 0, -0.1, 0, 0.2], [0, 0, 0.4, 0.1, -0.2, 0],
                      [0, -0.1, 0.1, 0.51, 0, 0], [0.1, 0, -0.2, 0, 0.61, -0.1],
                      [-0.1, 0.2, 0, 0, -0.1, 0.61]])
    A = O.pwc_raw(3, pose3(), A_cov)
    C = O.pwc_raw(3, pose3(O.quat_to_rot(0, 0, 1, 0), (1, 0, 0)), C_cov)
    B = O.pwc_between(3, A, C)
    Bp, Bc, _ = O.pwc_split(3, B)
    np.testing.assert_allclose(Bp, pose3(O.quat_to_rot(0, 0, 1, 0), (1, 0, 0)), atol=1e-12)
    np.testing.assert_allclose(Bc, 0.1 * np.eye(6), atol=1e-9)


def test_pwc_inverse_keeps_covariance():
    """testPoseWithCovariance.cpp:32-48 and quirk Q4 (GeometryUtils.h:133-139)."""
    R = rodrigues([0.3, -0.2, 0.5])
    A = O.pwc_raw(3, pose3(R, (1, 2, 3)), np.diag([1, 2, 3, 4, 5, 6.0]))
    B = O.pwc_inverse(3, A)
    Bp, Bc, _ = O.pwc_split(3, B)
    np.testing.assert_array_equal(Bc, np.diag([1, 2, 3, 4, 5, 6.0]))
    np.testing.assert_allclose(O.se3_compose(pose3(R, (1, 2, 3)), Bp), pose3(), atol=1e-14)


# ------------------------------------------------------------ testTrajectory
def test_trajectory_between_equals_stitched():
    """testTrajectory.cpp:35-109: getBetween(2,99) == compose of odometry 2..98 (1e-9)."""
    odom = pose3(I3, (1, 1, 0))
    cov = 1e-4 * np.eye(6)
    traj = O.fold(3, pose3(), np.tile(odom, (100, 1)), np.tile(cov.reshape(36), (100, 1)))
    between = O.pwc_between(3, traj[3], traj[100])  # 97 odometry steps apart
    rebuild = O.pwc_raw(3, pose3(), np.zeros((6, 6)))
    for _ in range(97):
        rebuild = O.pwc_compose(3, rebuild, O.pwc_raw(3, odom, cov))
    np.testing.assert_allclose(between[:12], rebuild[:12], atol=1e-9)
    np.testing.assert_allclose(between[12:48], rebuild[12:48], atol=1e-9)


def test_trajectory_between_rotating_chain():
    """Segment covariance identity on a rotating chain (the between covariance
    is the covariance of the stitched segment, exact arithmetic)."""
    rng = np.random.default_rng(3)
    n = 40
    odoms, covs = [], []
    for _ in range(n - 1):
        odoms.append(pose3(rodrigues(rng.normal(0, 0.2, 3)), rng.normal(0, 1, 3)))
        covs.append(np.diag(rng.uniform(0.01, 0.1, 6)).reshape(36))
    traj = O.fold(3, pose3(), np.array(odoms), np.array(covs))
    a, c = 7, 31
    seg = O.pwc_raw(3, pose3(), np.zeros((6, 6)))
    for k in range(a, c):
        seg = O.pwc_compose(3, seg, O.pwc_raw(3, odoms[k], covs[k]))
    got = O.pwc_between(3, traj[a], traj[c])
    np.testing.assert_allclose(got[:12], seg[:12], atol=1e-10)
    np.testing.assert_allclose(got[12:48], seg[12:48], rtol=1e-7, atol=1e-9)
    # reversed order takes the LLT fallback (quirk Q5): the first attempt is
    # negative definite, the fallback recomputes with the other operand's
    # Jacobian, which again yields the segment covariance in the frame of the
    # later pose c, while the pose stays T_c^-1 T_a
    rev = O.pwc_between(3, traj[c], traj[a])
    assert not O.llt_ok(traj[a][12:48].reshape(6, 6) - O.se3_adjoint(
        O.se3_between(traj[c][:12], traj[a][:12])) @ traj[c][12:48].reshape(6, 6) @ O.se3_adjoint(
        O.se3_between(traj[c][:12], traj[a][:12])).T)
    np.testing.assert_allclose(rev[12:48], seg[12:48], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(rev[:12], O.se3_between(traj[c][:12], traj[a][:12]), atol=1e-12)


# ------------------------------------------------------------ testPoseWithNode / Logmap
def test_logmap_identity_translation():
    """testPoseWithNode.cpp:92-94: Log of (I,(1,0,0)) has translation (1,0,0) exactly."""
    xi = O.se3_logmap(pose3(I3, (1, 0, 0)))
    assert xi[3] == 1.0 and xi[4] == 0.0 and xi[5] == 0.0
    assert math.sqrt(xi[3] ** 2) / 5 == 0.2


def test_logmap_exactly_pi():
    """testPoseWithNode.cpp:96-98: Log of Ry(pi) has rotation norm pi."""
    w = O.rot3_logmap(O.quat_to_rot(0, 0, 1, 0))
    assert abs(np.linalg.norm(w) / 5 - 3.1415927 / 5) < 1e-6
    for q in [(0, 1, 0, 0), (0, 0, 0, 1), (0, 0.6, 0.8, 0)]:
        w = O.rot3_logmap(O.quat_to_rot(*q))
        assert abs(np.linalg.norm(w) - math.pi) < 1e-12
        np.testing.assert_allclose(rodrigues(w), O.quat_to_rot(*q), atol=1e-12)


def test_logmap_exp_roundtrip_and_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(200):
        xi = np.concatenate([rng.normal(0, 1, 3), rng.normal(0, 3, 3)])
        if np.linalg.norm(xi[:3]) > 3.0:
            continue
        T = O.se3_expmap(xi)
        np.testing.assert_allclose(T[:9].reshape(3, 3), rodrigues(xi[:3]), atol=1e-12)
        np.testing.assert_allclose(O.se3_logmap(T), xi, atol=1e-9)
    # near-zero rotations take the Taylor branch
    for th in [1e-5, 1e-8, 1e-12]:
        xi = np.array([th, -th, 0.5 * th, 1.0, 2.0, 3.0])
        np.testing.assert_allclose(O.se3_logmap(O.se3_expmap(xi)), xi, atol=1e-9)
    # near pi (within 1e-4 of pi): the acos branch stays accurate
    for d in [1e-3, 1e-4]:
        w = np.array([0.0, 0.0, math.pi - d])
        np.testing.assert_allclose(O.rot3_logmap(rodrigues(w)), w, atol=1e-8)


def test_se2_ops_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(100):
        th, x, y = rng.uniform(-3, 3), rng.normal(), rng.normal()
        p = np.array([math.cos(th), math.sin(th), x, y])
        xi = O.se2_logmap(p)
        assert abs(xi[2] - th) < 1e-12
        # Pose2 expmap closed form reproduces (x, y)
        w = xi[2]
        V = np.array([[math.sin(w) / w, -(1 - math.cos(w)) / w], [(1 - math.cos(w)) / w, math.sin(w) / w]])
        np.testing.assert_allclose(V @ xi[:2], [x, y], atol=1e-10)
        q = O.se2_compose(p, O.se2_inverse(p))
        np.testing.assert_allclose(q, [1, 0, 0, 0], atol=1e-12)


def test_adjoint_conjugation_identity():
    """Ad(T) xi^ = T xi^ T^-1 : checked through exp(Ad xi) = T exp(xi) T^-1."""
    rng = np.random.default_rng(2)
    T = pose3(rodrigues(rng.normal(0, 1, 3)), rng.normal(0, 2, 3))
    xi = rng.normal(0, 0.3, 6)
    lhs = O.se3_expmap(O.se3_adjoint(T) @ xi)
    rhs = O.se3_compose(O.se3_compose(T, O.se3_expmap(xi)), O.se3_inverse(T))
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)


def test_llt_and_lu():
    rng = np.random.default_rng(5)
    for n in (2, 3, 6):
        M = rng.normal(size=(n, n))
        S = M @ M.T + 0.1 * np.eye(n)
        assert O.llt_ok(S)
        assert not O.llt_ok(-S)
        np.testing.assert_allclose(O.lu_inverse(S) @ S, np.eye(n), atol=1e-10)
        np.testing.assert_allclose(O.lu_inverse(M), np.linalg.inv(M), rtol=1e-8, atol=1e-10)
    assert not O.llt_ok(np.zeros((6, 6)))  # zero pivot fails (Eigen: x <= 0)


def test_nan_rotation_covariance_uses_translation_only():
    """GeometryUtils.h:102-113 and 175-183 (rotation_info=false path)."""
    cov = np.diag([np.nan, 1, 1, 4.0, 4.0, 4.0])
    p = O.pwc(3, pose3(rz(0.3), (2, 0, 0)), cov)
    _, c, ri = O.pwc_split(3, p)
    assert not ri
    assert np.all(c[:3, :] == 0) and np.all(c[3:, 3:] == np.diag([4.0, 4, 4]))
    xi = O.se3_logmap(pose3(rz(0.3), (2, 0, 0)))
    d = O.pwc_mahalanobis(3, p)
    assert abs(d - math.sqrt(xi[3:] @ xi[3:] / 4.0)) < 1e-12


# ------------------------------------------------------------ testPcm
def _odom_chain(n_steps, R, t, var):
    odom = np.tile(pose3(R, t), (n_steps, 1))
    cov = np.tile((var * np.eye(6)).reshape(36), (n_steps, 1))
    return O.fold(3, pose3(), odom, cov)


def test_pcm_odometry_check_kat():
    """testPcm.cpp:17-87: LC1 (var 0.1) passes odom threshold 0.3, LC2 (var 0.05) fails."""
    traj = _odom_chain(3, R90, (1, 0, 0), 0.1)
    lc = pose3(rz(1.51), (0.8, 0, 0))
    d1 = O.odom_dist(3, traj, 3, 0, lc, 0.1 * np.eye(6))
    d2 = O.odom_dist(3, traj, 3, 0, lc, 0.05 * np.eye(6))
    assert 0.25 < d1 < 0.3, d1
    assert d2 > 0.3, d2
    acc, _, _, _ = O.pcm_adjacency(3, traj, [3, 3], [0, 0], [lc, lc],
                                   [0.1 * np.eye(6), 0.05 * np.eye(6)], 0.3, -1)
    assert list(acc) == [0]


def test_pcm_consistency_check_kat():
    """testPcm.cpp:90-192: pairwise threshold 0.5; LC1..LC3 form the clique,
    LC4 is consistent with LC3 only and stays out of it."""
    odom = [pose3(R90, (1, 0, 0))] * 2 + [pose3(I3, (1, 0, 0))] * 4
    cov = [(0.1 * np.eye(6)).reshape(36)] * 6
    traj = O.fold(3, pose3(), np.array(odom), np.array(cov))
    lcs = [(3, pose3(rz(3.1416), (0, 0.9, 0))), (4, pose3(rz(3.1416), (-1, 0.8, 0))),
           (5, pose3(rz(0.99 * 3.1416), (-1.8, 0.8, 0))), (6, pose3(rz(0.98 * 3.1416), (-2.6, 0.6, 0)))]
    frm = [k for k, _ in lcs]
    zs = [z for _, z in lcs]
    covs = [0.1 * np.eye(6)] * 4
    sizes = []
    for m in (2, 3, 4):
        acc, _, adj, dist = O.pcm_adjacency(3, traj, frm[:m], [0] * m, zs[:m], covs[:m], -1, 0.5)
        k, clique = O.max_clique_heu(adj)
        sizes.append(k)
    assert sizes == [2, 3, 3]
    assert 3 not in set(clique.tolist())
    assert adj[3, 2] == 1 and adj[3, 0] == 0 and adj[3, 1] == 0
    assert dist[2, 0] < 0.5 and dist[2, 1] < 0.5 and dist[3, 2] < 0.5


# ------------------------------------------------------------ max clique semantics
def test_heu_stale_buffer_counterexample():
    """findCliqueHeu.cpp:83-113: V={0,1,2,3}, E={01,02,03,13} returns [0,1,2]."""
    adj = np.zeros((4, 4), np.uint8)
    for u, v in [(0, 1), (0, 2), (0, 3), (1, 3)]:
        adj[u, v] = adj[v, u] = 1
    k, out = O.max_clique_heu(adj)
    assert k == 3 and out.tolist() == [0, 1, 2]


def _is_clique(adj, s):
    return all(adj[u, v] for u, v in itertools.combinations(s, 2))


def _brute_max_clique(adj):
    n = adj.shape[0]
    best = 0
    for mask in range(1, 1 << n):
        s = [i for i in range(n) if mask >> i & 1]
        if len(s) > best and _is_clique(adj, s):
            best = len(s)
    return best


def test_exact_clique_matches_brute_force():
    rng = np.random.default_rng(11)
    for trial in range(60):
        n = int(rng.integers(1, 14))
        p = rng.uniform(0.2, 0.9)
        A = (rng.uniform(size=(n, n)) < p).astype(np.uint8)
        A = np.triu(A, 1)
        A = A + A.T
        k, s = O.max_clique_exact(A)
        assert k == _brute_max_clique(A)
        assert len(s) == k and _is_clique(A, s)
        kh, sh = O.max_clique_heu(A)
        assert 1 <= kh <= k
        assert sh[0] == sh[0]  # buffer position 0 is the seed candidate


def test_heu_greedy_size_matches_python_restatement():
    """Independent Python restatement of the A7 semantics (list based)."""
    rng = np.random.default_rng(12)
    for trial in range(80):
        n = int(rng.integers(1, 40))
        A = (rng.uniform(size=(n, n)) < rng.uniform(0.1, 0.9)).astype(np.uint8)
        A = np.triu(A, 1); A = A + A.T
        nbr = [sorted(np.nonzero(A[v])[0].tolist()) for v in range(n)]
        deg = [len(x) for x in nbr]
        best, buf, out = -1, [0] * (max(deg) + 1), None
        for c in range(n):
            if best > deg[c]:
                continue
            S = [c] + [u for u in nbr[c] if deg[u] >= best]
            for i, v in enumerate(S):
                buf[i] = v
            size, icc = len(S), 0
            while size > 0:
                icc += 1
                m = buf[size - 1]
                S1 = [buf[j] for j in range(size) if buf[j] in nbr[m] and deg[buf[j]] >= best]
                for i, v in enumerate(S1):
                    buf[i] = v
                size = len(S1)
            if best < icc:
                best, out = icc, list(buf)
        k, got = O.max_clique_heu(A)
        assert k == best and got.tolist() == out[:best]


def test_heu_dense_matches_literal():
    """The bitset evaluation of maxCliqueHeu equals the literal CSR replay."""
    rng = np.random.default_rng(13)
    for trial in range(60):
        n = int(rng.integers(1, 120))
        A = (rng.uniform(size=(n, n)) < rng.uniform(0.05, 0.95)).astype(np.uint8)
        # plant a dense block so the greedy runs deep
        blk = rng.choice(n, size=min(n, int(rng.integers(1, 40))), replace=False)
        A[np.ix_(blk, blk)] = (rng.uniform(size=(len(blk), len(blk))) < 0.97)
        A = np.triu(A, 1); A = A + A.T
        k1, s1 = O.max_clique_heu(A)
        k2, s2 = O.max_clique_heu_dense(A)
        assert k1 == k2 and s1.tolist() == s2.tolist()
