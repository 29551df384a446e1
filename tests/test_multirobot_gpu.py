"""GPU parity of several-robot PCM groups (rpgo_pcm_problem with n_chains > 1)
against the FP64 oracle (orc_fold_chains / orc_pcm_adjacency_chains).

Bar: trajectory records, accepted set and adjacency bits bit-exact; distances
within 1e-12 relative.  Cases: the two-robot scenario of
tests/testMultiRobot.cpp:22-188 (inter-robot group, intra-robot group, clique
sizes 2 / 2 / 1), seeded three-robot graphs split from one helix with
inter-robot inliers and outliers (Mahalanobis and PcmSimple checks, FP32
requested: several chains take the FP64 replay), and the upload checks.
"""
import math

import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-12
I3 = np.eye(3)


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def _robots(n_robots, n_each, seed):
    """Robots = consecutive pieces of one helix; each chain starts at its true pose."""
    g = synth.se3_helix(n_poses=n_robots * n_each, n_lc=10, outlier_frac=0.0, seed=seed)
    chain_off = [r * n_each for r in range(n_robots + 1)]
    keep = [k for k in range(g.n_poses - 1) if (k + 1) % n_each != 0]
    pose0s = np.stack([g.gt[c] for c in chain_off[:-1]])
    return g, chain_off, pose0s, g.odom[keep], g.odom_cov[keep]


def _lcs(g, chain_off, ra, rb, n_in, n_out, seed):
    """Loop closures between robots ra and rb: true relative poses with small
    noise (inliers) and random ones (outliers), random direction."""
    rng = np.random.default_rng(seed)
    Rg, tg = synth.se3_unpack(g.gt)
    frm, to, zs = [], [], []
    for q in range(n_in + n_out):
        i = int(rng.integers(chain_off[ra], chain_off[ra + 1]))
        j = int(rng.integers(chain_off[rb], chain_off[rb + 1]))
        if i == j:
            continue
        if rng.uniform() < 0.5:
            i, j = j, i
        if q < n_in:
            Rr, tr = synth.se3_rel(Rg[i], tg[i], Rg[j], tg[j])
            Rr = Rr @ synth.rot_exp(rng.normal(0, 0.003, (1, 3)))[0]
            tr = tr + rng.normal(0, 0.01, 3)
        else:
            Rr = synth.random_rotations(rng, 1)[0]
            tr = rng.uniform(-20, 20, 3)
        frm.append(i)
        to.append(j)
        zs.append(synth.se3_pack(Rr, tr))
    cov = np.tile(np.diag([0.005 ** 2] * 3 + [0.02 ** 2] * 3).reshape(36), (len(frm), 1))
    order = np.random.default_rng(seed + 1).permutation(len(frm))
    return (np.array(frm)[order], np.array(to)[order], np.array(zs)[order], cov)


def _check(ctx, chain_off, pose0s, odom, ocov, lcs, thr, simple=None, precision=_capi.PCM_FP32_RECHECK):
    frm, to, zs, cov = lcs
    print(f"[multirobot] chains={len(chain_off) - 1} n_lc={len(frm)} thr={thr} simple={simple}", flush=True)
    ctx.pcm_upload(3, pose0s, odom, ocov, frm, to, zs, cov, chain_off=chain_off)
    if simple is None:
        st = ctx.pcm_run(*thr, precision=precision, want_dist=True)
    else:
        st = ctx.pcm_simple_run(*simple, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    traj = O.fold_chains(3, chain_off, pose0s, odom, ocov)
    assert np.array_equal(ctx.pcm_traj(), traj)
    acc, od, adj, dist = O.pcm_adjacency_chains(3, chain_off, traj, frm, to, zs, cov, *thr, simple=simple)
    assert np.array_equal(got["accepted"], acc)
    np.testing.assert_allclose(got["odom_dist"], od, rtol=DIST_RTOL, atol=0)
    m = len(acc)
    A = _capi.unpack_bits(got["adj_bits"], m) if m else np.zeros((0, 0), np.uint8)
    assert np.array_equal(A, adj), np.argwhere(A != adj)[:10]
    if m:
        np.testing.assert_allclose(got["dist"], dist, rtol=DIST_RTOL, atol=1e-300)
    assert st.n_consistent == int(adj.sum()) // 2
    return st, acc, adj


def test_reference_two_robot_scenario(ctx):
    """testMultiRobot.cpp:22-188 through the C ABI."""
    R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    Rm90 = R90.T
    R180 = np.diag([-1.0, -1.0, 1.0])
    odom = np.stack([pose3(I3, (1, 0, 0))] * 3 + [pose3(R90, (1, 0, 0))] * 2 +
                    [pose3(I3, (1, 0, 0))] * 3 + [pose3(Rm90, (1, 0, 0))] * 2)
    ocov = np.tile((0.1 * np.eye(6)).reshape(36), (10, 1))
    pose0s = np.stack([pose3(), pose3(I3, (0, -1, 0))])
    chain_off = [0, 6, 12]
    lc = lambda f, t, z: (f, t, z)  # noqa: E731
    inter = [lc(7, 1, pose3(I3, (0, 1, 0))), lc(10, 4, pose3(R180, (-1, 0, 0))),
             lc(2, 8, pose3(I3, (0, -1, 0))), lc(11, 5, pose3(I3, (0, -3.3, 0)))]
    c, s = math.cos(-1.54), math.sin(-1.54)
    intra = [lc(4, 1, pose3(Rm90, (0, 3, 0))), lc(4, 2, pose3(np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]]), (0, 2, 0)))]

    def arrays(lcs):
        return (np.array([a for a, _, _ in lcs]), np.array([b for _, b, _ in lcs]),
                np.array([z for _, _, z in lcs]), np.tile((0.1 * np.eye(6)).reshape(36), (len(lcs), 1)))

    for lcs, want in ((inter[:2], 2), (inter, 2), (intra, 1)):
        _check(ctx, chain_off, pose0s, odom, ocov, arrays(lcs), (3.0, 0.05))
        k, _, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
        assert k == want


@pytest.mark.parametrize("seed", [0, 1])
def test_three_robot_groups(ctx, seed):
    g, chain_off, pose0s, odom, ocov = _robots(3, 300, seed)
    for ra, rb in ((0, 1), (1, 2), (0, 0), (2, 0)):
        lcs = _lcs(g, chain_off, ra, rb, 40, 15, seed * 10 + ra * 3 + rb)
        st, acc, adj = _check(ctx, chain_off, pose0s, odom, ocov, lcs, (12.0, 5.0))
        assert len(acc) > 0
        if ra != rb:
            assert len(acc) == len(lcs[0])  # inter-robot: no odometry check


def test_three_robot_simple(ctx):
    g, chain_off, pose0s, odom, ocov = _robots(3, 250, 5)
    lcs = _lcs(g, chain_off, 0, 2, 30, 10, 77)
    _check(ctx, chain_off, pose0s, odom, ocov, lcs, (0.0, 0.0), simple=(0.5, 0.05, 0.02, 0.002))


def test_upload_checks(ctx):
    g, chain_off, pose0s, odom, ocov = _robots(3, 100, 2)
    frm, to, zs, cov = _lcs(g, chain_off, 0, 1, 5, 0, 3)
    frm2, to2, zs2, cov2 = _lcs(g, chain_off, 1, 2, 5, 0, 4)
    with pytest.raises(RuntimeError):  # two groups in one problem
        ctx.pcm_upload(3, pose0s, odom, ocov, np.r_[frm, frm2], np.r_[to, to2], np.r_[zs, zs2],
                       np.r_[cov, cov2], chain_off=chain_off)
    with pytest.raises(RuntimeError):  # chain offsets that do not cover the poses
        ctx.pcm_upload(3, pose0s, odom, ocov, frm, to, zs, cov, chain_off=[0, 100, 150, 300])
    ctx.pcm_upload(3, pose0s, odom, ocov, frm, to, zs, cov, chain_off=chain_off)
    ctx.pcm_run(-1.0, 5.0)
    with pytest.raises(RuntimeError):  # streaming takes one chain
        ctx.pcm_append(-1.0, 5.0, odom=odom[:1], odom_cov=ocov[:1])
