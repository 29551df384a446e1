"""CPU checks of helpers the measurements rely on: the oracle's restatement of
incrementAdjMatrix's storage pattern (timed as the reference's O(m^3) CPU
cost) decides the same pairs as the oracle's adjacency; the quaternion
storage of the solve fixtures round-trips rotations at every angle."""
import os

import numpy as np

import oracle as O
from rpgo_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


def test_adj_growth_counts_the_oracle_adjacency():
    g = synth.se3_helix(n_poses=600, n_lc=150, outlier_frac=0.3, seed=21)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    acc, _, adj, _ = O.pcm_adjacency(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov, -1.0, 5.0)
    assert len(acc) == g.n_lc
    n = O.adj_growth(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov, 5.0, True)
    assert n == int(np.triu(adj, 1).sum())
    # without the pair checks every pair is stored as consistent (lc_threshold < 0, Pcm.h:673)
    m = g.n_lc
    assert O.adj_growth(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov, 5.0, False) == m * (m - 1) // 2


def test_quaternion_roundtrip_at_every_angle():
    rng = np.random.default_rng(4)
    R = synth.random_rotations(rng, 20000)
    assert np.abs(synth.quat_to_rot(synth.rot_to_quat(R)) - R).max() < 1e-14
    # rotations by exactly pi and within 1e-9 of pi (where trace-based logs lose digits)
    axes = rng.normal(size=(200, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    for ang in (np.pi, np.pi - 1e-9, 1e-12):
        Rp = synth.rot_exp(axes * ang)
        assert np.abs(synth.quat_to_rot(synth.rot_to_quat(Rp)) - Rp).max() < 1e-14
    q = synth.rot_to_quat(R)
    assert np.all(q[:, 0] >= 0) and np.allclose(np.linalg.norm(q, axis=1), 1.0)


def test_solve_fixture_format():
    z = np.load(os.path.join(HERE, "golden", "solve_se3_5k.npz"), allow_pickle=False)
    n = 5000
    assert z["q"].shape == (n, 4) and z["t"].shape == (n, 3)
    assert np.allclose(np.linalg.norm(z["q"], axis=1), 1.0, atol=1e-12)
    nk = n  # odometry n - 1 plus the prior
    w = z["weights"]
    assert w.shape == (nk + len(z["inliers"]),)
    assert np.all(w[:nk] == 1.0)
    assert np.all(np.abs(w - np.rint(w)) <= 1e-4)  # converged GNC: binary weights
