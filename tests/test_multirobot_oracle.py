"""Pin the several-chain (multi-robot) part of the FP64 oracle.

Reference test restated (path relative to the reference tree):
  tests/testMultiRobot.cpp:22-188   two robots a and b, inter-robot loop
                                    closures b1->a1, b4->a4 kept, a2->b2 and
                                    b5->a5 rejected, intra-robot a4->a1 kept and
                                    a4->a2 rejected (PCM 3D, odom 3.0, lc 0.05)
plus the invariants that fix the chain handling: a one-chain problem equals
the single-chain functions, the fold restarts at every chain start, and an
inter-robot pair check equals the single-chain pair check with the newer loop
closure's endpoints swapped (Pcm.h:691-698).
"""
import math

import numpy as np

import oracle as O

I3 = np.eye(3)
R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
Rm90 = np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1.0]])
R180 = np.diag([-1.0, -1.0, 1.0])


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def cov6(v):
    return v * np.eye(6)


def two_robots():
    """Chains a (poses 0..5) and b (poses 6..11) of testMultiRobot.cpp:34-110."""
    odom_a = [pose3(I3, (1, 0, 0))] * 3 + [pose3(R90, (1, 0, 0))] * 2
    odom_b = [pose3(I3, (1, 0, 0))] * 3 + [pose3(Rm90, (1, 0, 0))] * 2
    chain_off = [0, 6, 12]
    pose0s = np.stack([pose3(), pose3(I3, (0, -1, 0))])
    odom = np.stack(odom_a + odom_b)
    cov = np.tile(cov6(0.1).reshape(36), (10, 1))
    return chain_off, O.fold_chains(3, chain_off, pose0s, odom, cov)


A = lambda k: k        # noqa: E731  robot a pose k
B = lambda k: 6 + k    # noqa: E731  robot b pose k


def _group(traj, chain_off, lcs, odom_thr, lc_thr):
    frm = [f for f, _, _ in lcs]
    to = [t for _, t, _ in lcs]
    zs = [z for _, _, z in lcs]
    covs = [cov6(0.1)] * len(lcs)
    return O.pcm_adjacency_chains(3, chain_off, traj, frm, to, zs, covs, odom_thr, lc_thr)


def test_fold_restarts_per_chain():
    chain_off, traj = two_robots()
    pl = 12
    # each chain starts at its own pose with zero covariance and rotation info
    assert np.array_equal(traj[6, :pl], pose3(I3, (0, -1, 0)))
    assert np.all(traj[6, pl:pl + 36] == 0.0) and traj[6, -1] == 1.0
    # and equals the single-chain fold of its own odometry
    single_b = O.fold(3, pose3(I3, (0, -1, 0)), [pose3(I3, (1, 0, 0))] * 3 + [pose3(Rm90, (1, 0, 0))] * 2,
                      np.tile(cov6(0.1).reshape(36), (5, 1)))
    assert np.array_equal(traj[6:], single_b)


def test_reference_multirobot_scenario():
    """testMultiRobot.cpp:114-188."""
    chain_off, traj = two_robots()
    good = [(B(1), A(1), pose3(I3, (0, 1, 0))), (B(4), A(4), pose3(R180, (-1, 0, 0)))]
    bad = [(A(2), B(2), pose3(I3, (0, -1, 0))), (B(5), A(5), pose3(I3, (0, -3.3, 0)))]
    acc, od, adj, _ = _group(traj, chain_off, good, 3.0, 0.05)
    assert list(acc) == [0, 1] and np.all(od == 0.0)  # inter-robot: no odometry check
    k, clique = O.max_clique_heu(adj)
    assert k == 2
    acc, _, adj, _ = _group(traj, chain_off, good + bad, 3.0, 0.05)
    assert len(acc) == 4
    k, clique = O.max_clique_heu(adj)
    assert k == 2 and sorted(clique.tolist()) == [0, 1]
    # intra-robot group (a, a): a4->a1 kept, a4->a2 rejected
    intra = [(A(4), A(1), pose3(Rm90, (0, 3, 0))),
             (A(4), A(2), pose3(np.array([[math.cos(-1.54), -math.sin(-1.54), 0],
                                          [math.sin(-1.54), math.cos(-1.54), 0], [0, 0, 1.0]]), (0, 2, 0)))]
    acc, _, adj, _ = _group(traj, chain_off, intra, 3.0, 0.05)
    inliers = O.max_clique_heu(adj)[0] if len(acc) else 0
    assert inliers == 1


def test_one_chain_equals_single_chain_functions():
    rng = np.random.default_rng(4)
    n = 60
    odom = [pose3(I3, (1, rng.normal(0, .05), 0)) for _ in range(n - 1)]
    cov = np.tile(cov6(0.01).reshape(36), (n - 1, 1))
    traj1 = O.fold(3, pose3(), odom, cov)
    trajc = O.fold_chains(3, [0, n], pose3()[None], np.stack(odom), cov)
    assert np.array_equal(traj1, trajc)
    lcs = [(int(a), int(b), pose3(I3, rng.normal(0, 3, 3))) for a, b in rng.integers(0, n, (15, 2)) if a != b]
    frm = [a for a, _, _ in lcs]
    to = [b for _, b, _ in lcs]
    zs = [z for _, _, z in lcs]
    covs = [cov6(0.1)] * len(lcs)
    ref = O.pcm_adjacency(3, traj1, frm, to, zs, covs, 40.0, 30.0)
    got = O.pcm_adjacency_chains(3, [0, n], traj1, frm, to, zs, covs, 40.0, 30.0)
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)


def test_inter_robot_pair_is_swapped_single_chain_check():
    chain_off, traj = two_robots()
    z1, z2 = pose3(I3, (0, 1, 0)), pose3(R180, (-1, 0, 0))
    # older (b1 -> a1), newer (a4 -> b4): first endpoints on different chains
    _, _, _, dist = _group(traj, chain_off, [(B(1), A(1), z1), (A(4), B(4), z2)], -1.0, 1e9)
    expect = O.pair_dist(3, traj, B(1), A(1), z1, cov6(0.1), B(4), A(4), z2, cov6(0.1))
    assert dist[1, 0] == expect


def test_cross_robot_between_kat():
    """testTrajectory.cpp:112-162: getBetween(a1, b1) across two robots via the
    root poses a0, b0 is the pose (Rz(90), (2, -1, 0))."""
    cov0 = np.zeros(36)
    a0 = O.pwc_raw(3, pose3(I3, (1, 1, 1)), cov0)
    b0 = O.pwc_compose(3, a0, O.pwc_raw(3, pose3(I3, (1, 0, 0)), 0.01 * np.eye(6).reshape(36)))
    a1 = O.pwc_compose(3, a0, O.pwc_raw(3, pose3(I3, (0, 1, 0)), 0.1 * np.eye(6).reshape(36)))
    b1 = O.pwc_compose(3, b0, O.pwc_raw(3, pose3(R90, (1, 0, 0)), 0.2 * np.eye(6).reshape(36)))
    r = O.cross_between(3, a0, a1, b0, b1)
    np.testing.assert_allclose(r[:12], pose3(R90, (2, -1, 0)), atol=1e-12)
    # same robot: the root-chain route equals the direct between (pose part)
    direct = O.pwc_between(3, a1, b1)
    np.testing.assert_allclose(r[:12], direct[:12], atol=1e-12)
