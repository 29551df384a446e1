"""GPU robust solve (PCM + GNC-TLS over LM) against the FP64 oracle's solution
at the benchmark sizes.

The fixtures (tests/golden/solve_se3_5k.npz, solve_se3_50k.npz) hold the FP64
oracle's result on the seeded helix graphs: the reference heuristic clique of
the oracle's PCM, then oracle/solver_oracle.gnc_optimize (GTSAM GncOptimizer
semantics, src/RobustSolver.cpp:130-180) on odometry + prior + those loop
closures; tests/golden/make_solve_fixtures.py wrote them.  Bar (north star):
the GPU clique equals the oracle's, the GNC weight vector is the same, and the
poses are within 1e-5 m / 1e-5 rad of the oracle's.
"""
import os

import numpy as np
import pytest

from rpgo_amd import _capi, learned, synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
POSE_TOL = 1e-5
CASES = {"se3_5k": (5000, 1000, 0.2), "se3_50k": (50000, 10000, 0.5)}


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _rot_err(Ra, Rb):
    D = np.swapaxes(Ra, -1, -2) @ Rb
    s = 0.5 * np.linalg.norm(np.stack([D[:, 2, 1] - D[:, 1, 2], D[:, 0, 2] - D[:, 2, 0],
                                       D[:, 1, 0] - D[:, 0, 1]], -1), axis=-1)
    c = 0.5 * (D[:, 0, 0] + D[:, 1, 1] + D[:, 2, 2] - 1.0)
    return float(np.arctan2(s, c).max())


@pytest.mark.parametrize("name", sorted(CASES))
def test_robust_solve_matches_fp64_oracle(ctx, name):
    path = os.path.join(HERE, "golden", f"solve_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not generated (tests/golden/make_solve_fixtures.py {name})")
    z = np.load(path, allow_pickle=False)
    n, nlc, of = CASES[name]
    g = synth.se3_helix(n_poses=n, n_lc=nlc, outlier_frac=of, seed=0)
    print(f"[fixture] {name}: PCM", flush=True)
    pcm = learned.pcm_stage(ctx, g)
    assert np.array_equal(pcm.inliers, z["inliers"]), "clique differs from the oracle's"
    print(f"[fixture] {name}: plain GNC on {len(pcm.inliers)} loop closures", flush=True)
    poses, w, rep, nk = learned.solve_plain(ctx, g, pcm.inliers)
    dt = float(np.abs(poses[:, 9:] - z["t"]).max())
    dr = _rot_err(synth.quat_to_rot(z["q"]), poses[:, :9].reshape(-1, 3, 3))
    wd = int((np.rint(w) != np.rint(z["weights"])).sum())
    print(f"[fixture] {name}: gnc it {rep.gnc_iterations} (oracle {int(z['iterations'])}) lm {rep.lm_iterations} "
          f"pcg {rep.pcg_iterations} ms {rep.ms_total:.0f}; weights differing {wd}; dt {dt:.2e} m dr {dr:.2e} rad",
          flush=True)
    assert wd == 0
    # fractional weights (a cost-test stop) agree to the LM stopping tolerance
    np.testing.assert_allclose(w, z["weights"], atol=1e-3)
    assert dt <= POSE_TOL and dr <= POSE_TOL
