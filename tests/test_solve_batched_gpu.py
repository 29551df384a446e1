"""rpgo_solve_batched: a batch of uneven robust solves (the C4 shape after PCM:
independent SE(3) graphs of different sizes and outlier rates, each solved by
GNC-TLS over LM on odometry + prior + its PCM clique's loop closures, as
RobustSolver::optimize does per graph, src/RobustSolver.cpp:115-306).

Bar: every problem's GNC weights equal rpgo_solve's on that problem alone, its
poses within 1e-5 m / 1e-5 rad of that solve (the same kernels; the device sums
may round differently run to run), and its iteration counts equal; one and
four lanes agree to the same bar; the smallest problem matches the FP64 oracle
(same weights, poses within 1e-5 m / 1e-5 rad).
"""
import numpy as np
import pytest

import solver_oracle as S
from graphs import pose_error
from rpgo_amd import _capi, learned, problem, synth

pytestmark = pytest.mark.gpu
TOL = 1e-5
# (poses, outlier fraction); loop closures are a fifth of the poses
SIZES = [(1500, 0.3), (300, 0.2), (800, 0.4), (2500, 0.3), (200, 0.1), (600, 0.25)]


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def batch(ctx):
    probs, known = [], []
    for k, (n, out) in enumerate(SIZES):
        gr = synth.se3_helix(n_poses=n, n_lc=n // 5, outlier_frac=out, seed=100 + k)
        pcm = learned.pcm_stage(ctx, gr)
        arrs = problem.solve_arrays(gr, pcm.inliers)
        probs.append(arrs[:-1])
        known.append(arrs[-1])
        print(f"[batch-solve] problem {k}: {n} poses, {len(pcm.inliers)} loop closures after PCM", flush=True)
    return probs, known


def _compare(P, W, Rp, out, w, rep, g):
    dt, dr = pose_error(P, out[:, :9].reshape(-1, 3, 3), out[:, 9:])
    same_w = np.array_equal(np.rint(W), np.rint(w))
    print(f"[batch-solve] problem {g}: gnc {Rp.gnc_iterations}/{rep.gnc_iterations} lm "
          f"{Rp.lm_iterations}/{rep.lm_iterations} dt {dt:.2e} dr {dr:.2e} same weights {same_w}", flush=True)
    assert same_w
    assert dt <= TOL and dr <= TOL
    assert Rp.gnc_iterations == rep.gnc_iterations and Rp.lm_iterations == rep.lm_iterations


def test_batched_equals_single_solves(ctx, batch):
    probs, known = batch
    P, W, R, B = ctx.solve_batched(probs, known_inliers=known, lanes=4, use_gnc=1)
    print(f"[batch-solve] 4 lanes: wall {B.ms_wall:.1f} ms, device sum {B.ms_device_sum:.1f} ms", flush=True)
    assert B.lanes == 4 and B.failed_problem == -1
    for g, (p, k) in enumerate(zip(probs, known)):
        out, w, rep = ctx.solve(*p, use_gnc=1, known_inliers=k)
        _compare(P[g], W[g], R[g], out, w, rep, g)
        assert R[g].n_inliers == rep.n_inliers


def test_one_lane_and_four_lanes_agree(ctx, batch):
    probs, known = batch
    P1, W1, R1, B1 = ctx.solve_batched(probs, known_inliers=known, lanes=1, use_gnc=1)
    P4, W4, R4, B4 = ctx.solve_batched(probs, known_inliers=known, lanes=4, use_gnc=1)
    print(f"[batch-solve] wall 1 lane {B1.ms_wall:.1f} ms, 4 lanes {B4.ms_wall:.1f} ms", flush=True)
    assert B1.lanes == 1 and B4.lanes == 4
    for g in range(len(probs)):
        _compare(P4[g], W4[g], R4[g], P1[g], W1[g], R1[g], g)


def test_smallest_problem_matches_fp64_oracle(ctx, batch):
    probs, known = batch
    g = int(np.argmin([p[0].shape[0] for p in probs]))
    P, W, R, _ = ctx.solve_batched(probs, known_inliers=known, lanes=4, use_gnc=1)
    G = S.graph_from_arrays(*probs[g])
    ref = S.gnc_optimize(G, S.GncParams(inlier_threshold=0.9, known_inliers=np.arange(known[g])))
    dt, dr = pose_error(P[g], ref["R"], ref["t"])
    print(f"[batch-solve] problem {g} vs oracle: gnc {R[g].gnc_iterations} (oracle {ref['iterations']}) "
          f"dt {dt:.2e} dr {dr:.2e}", flush=True)
    assert np.array_equal(np.rint(W[g]), np.rint(ref["weights"]))
    assert dt <= TOL and dr <= TOL


def test_empty_batch_and_lane_cap(ctx):
    P, W, R, B = ctx.solve_batched([], lanes=4)
    assert P == [] and W == [] and B.failed_problem == -1
    eye = np.concatenate([np.eye(3).reshape(9), np.zeros(3)])
    one = (np.tile(eye, (2, 1)), [1], [0], [0], eye[None], np.eye(6).reshape(1, 36))
    P, W, R, B = ctx.solve_batched([one], lanes=8)
    assert B.lanes == 1  # never more lanes than problems
    assert np.allclose(P[0], np.tile(eye, (2, 1)))
