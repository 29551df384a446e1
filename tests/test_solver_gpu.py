"""GPU robust solve (rpgo_solve) against the FP64 oracle.

Bar: optimized poses within 1e-5 m / 1e-5 rad of the oracle's
converged FP64 solution; GNC weight vectors equal; reference KATs of
tests/testGnc.cpp:25-103 (robot_a) reproduced.
"""
import numpy as np
import pytest

import solver_oracle as S
from graphs import flat, from_robot_a, from_synth, pose_error
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu
POSE_TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _gpu(ctx, g, **kw):
    poses, kind, ki, kj, meas, info = flat(g)
    return ctx.solve(poses, kind, ki, kj, meas, info, **kw)


def test_lm_matches_oracle_least_squares(ctx):
    gr = synth.se3_helix(n_poses=600, n_lc=80, outlier_frac=0.0, seed=4)
    g, nk = from_synth(gr)
    print(f"[solve] LM N={g.N} F={g.F}", flush=True)
    out, w, rep = _gpu(ctx, g, polish_iterations=10, pcg_rel_tol=1e-6)
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    dt, dr = pose_error(out, R, t)
    print(f"[solve] lm it={rep.lm_iterations} pcg={rep.pcg_iterations} cost={rep.final_cost:.6f} "
          f"oracle={err:.6f} dt={dt:.2e} dr={dr:.2e} ms={rep.ms_total:.1f}", flush=True)
    assert dt < POSE_TOL and dr < POSE_TOL
    assert abs(rep.final_cost - err) <= 1e-6 * err


def test_lm_matches_oracle_20k(ctx):
    """Above 16,384 poses the PCG SpMV runs one thread per pose instead of one
    lane per row: the same sums, checked at the same bar."""
    gr = synth.se3_helix(n_poses=20000, n_lc=2000, outlier_frac=0.0, seed=6)
    g, nk = from_synth(gr)
    print(f"[solve] LM N={g.N} F={g.F}", flush=True)
    out, w, rep = _gpu(ctx, g, polish_iterations=10, pcg_rel_tol=1e-6)
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    dt, dr = pose_error(out, R, t)
    print(f"[solve] lm it={rep.lm_iterations} pcg={rep.pcg_iterations} cost={rep.final_cost:.6f} "
          f"oracle={err:.6f} dt={dt:.2e} dr={dr:.2e} ms={rep.ms_total:.1f}", flush=True)
    assert dt < POSE_TOL and dr < POSE_TOL
    assert abs(rep.final_cost - err) <= 1e-6 * err


def test_lm_segment_lengths_match_oracle(ctx):
    """Every preconditioner segment length reaches the oracle's optimum: S = 4,
    8, 16 and 32 run the fixed-S kernel (one workgroup per 8 segments, LDS
    staged, DPP sweeps), S = 6 and 1 the generic one.  2,003 poses: the last workgroup and
    the last segment are partial."""
    gr = synth.se3_helix(n_poses=2003, n_lc=250, outlier_frac=0.0, seed=8)
    g, nk = from_synth(gr)
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    for seg in (4, 8, 16, 32, 6, 1):
        print(f"[solve] LM N={g.N} pcg_segment={seg}", flush=True)
        out, w, rep = _gpu(ctx, g, polish_iterations=10, pcg_rel_tol=1e-6, pcg_segment=seg)
        dt, dr = pose_error(out, R, t)
        print(f"[solve] S={seg} lm it={rep.lm_iterations} pcg={rep.pcg_iterations} cost={rep.final_cost:.6f} "
              f"oracle={err:.6f} dt={dt:.2e} dr={dr:.2e} ms={rep.ms_total:.1f}", flush=True)
        assert dt < POSE_TOL and dr < POSE_TOL
        assert abs(rep.final_cost - err) <= 1e-6 * err


def test_gn_matches_oracle(ctx):
    gr = synth.se3_helix(n_poses=300, n_lc=40, outlier_frac=0.0, seed=5)
    g, nk = from_synth(gr)
    out, w, rep = _gpu(ctx, g, solver=_capi.SOLVER_GN, polish_iterations=10, pcg_rel_tol=1e-6)
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    dt, dr = pose_error(out, R, t)
    print(f"[solve] GN it={rep.lm_iterations} dt={dt:.2e} dr={dr:.2e}", flush=True)
    assert dt < POSE_TOL and dr < POSE_TOL


@pytest.mark.parametrize("alpha,expect", [(0.01, 0), (0.99, 3)])
def test_gnc_robot_a_kat(ctx, alpha, expect):
    """tests/testGnc.cpp:25-103 through the GPU solve."""
    g, nk = from_robot_a()
    out, w, rep = _gpu(ctx, g, use_gnc=1, known_inliers=nk, gnc_inlier_threshold=alpha)
    assert w[50:].sum() == expect and w.sum() == 50 + expect
    assert rep.n_inliers == expect
    ref = S.gnc_optimize(g, S.GncParams(inlier_threshold=alpha, known_inliers=np.arange(nk)))
    np.testing.assert_allclose(w, ref["weights"], atol=1e-3)


def test_gnc_matches_oracle_weights_and_poses(ctx):
    gr = synth.se3_helix(n_poses=400, n_lc=60, outlier_frac=0.3, seed=7)
    g, nk = from_synth(gr)
    print(f"[solve] GNC N={g.N} F={g.F}", flush=True)
    out, w, rep = _gpu(ctx, g, use_gnc=1, known_inliers=nk, pcg_rel_tol=1e-6, polish_iterations=10)
    ref = S.gnc_optimize(g, S.GncParams(inlier_threshold=0.9, known_inliers=np.arange(nk)))
    print(f"[solve] gnc it={rep.gnc_iterations} (oracle {ref['iterations']}) mu0={rep.mu_init:.3e} "
          f"lm={rep.lm_iterations} pcg={rep.pcg_iterations} ms={rep.ms_total:.1f}", flush=True)
    # weights of the converged GNC: same decisions; fractional values (if the
    # cost test stopped GNC) agree to the LM stopping tolerance
    assert np.array_equal(w > 0.5, ref["weights"] > 0.5)
    np.testing.assert_allclose(w, ref["weights"], atol=1e-3)
    assert abs(rep.gnc_iterations - ref["iterations"]) <= 1
    R, t, err, _ = S.least_squares(g, w, ref["R"], ref["t"])
    dt, dr = pose_error(out, R, t)
    print(f"[solve] gnc dt={dt:.2e} dr={dr:.2e}", flush=True)
    assert dt < POSE_TOL and dr < POSE_TOL


def test_bias_odom_and_cost_threshold(ctx):
    g, nk = from_robot_a()
    for mode, thr in [(_capi.GNC_COST, 1.0), (_capi.GNC_COST, 100.0)]:
        out, w, rep = _gpu(ctx, g, use_gnc=1, known_inliers=nk, gnc_threshold_mode=mode,
                           gnc_inlier_threshold=thr, gnc_bias_odom=1)
        init = np.zeros(g.F); init[:nk] = 1.0
        ref = S.gnc_optimize(g, S.GncParams(threshold_mode="cost", inlier_threshold=thr,
                                            known_inliers=np.arange(nk), init_weights=init))
        assert np.array_equal(w > 0.5, ref["weights"] > 0.5), (thr, w[nk:], ref["weights"][nk:])
        np.testing.assert_allclose(w, ref["weights"], atol=1e-3)


def test_empty_and_invalid(ctx):
    poses = np.tile(np.concatenate([np.eye(3).reshape(9), np.zeros(3)]), (3, 1))
    out, w, rep = ctx.solve(poses, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros((0, 12)),
                            np.zeros((0, 36)))
    assert np.array_equal(out, poses)
    with pytest.raises(_capi.RpgoError):
        ctx.solve(poses, [0], [0], [5], np.zeros((1, 12)), np.eye(6).reshape(1, 36))
