"""Learned GNC start against plain GNC on the same GPU (5,000-pose helices,
40 % outliers), with plain GNC as the judge.

What the network gains is measured, not assumed: for each graph the learned
start (features -> MLP -> initial weights and mu -> GNC) is compared with
plain GNC; the cost check accepts it only at a GNC fixed point whose TLS
objective does not exceed the first solve's, with plain's inlier set and poses
within 1e-5 m of plain's.  Whatever the check decides, the returned result has
plain's inlier set; an accepted start took fewer GNC iterations."""
import numpy as np
import pytest

from rpgo_amd import _capi, learned, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", [101, 102, 103])
def test_learned_start_judged_by_plain_gnc(ctx, seed):
    g = synth.se3_helix(n_poses=5000, n_lc=1000, outlier_frac=0.4, seed=seed)
    pcm = learned.pcm_stage(ctx, g)
    plain = learned.solve_plain(ctx, g, pcm.inliers)
    pp, wp, rp, nk = plain
    model = learned.LearnedModel.load()
    out = learned.solve_learned(ctx, g, pcm, model, plain=(pp, wp, rp))
    c = out.check
    print(f"[learned] seed {seed}: plain gnc {rp.gnc_iterations} lm {rp.lm_iterations} {rp.ms_total:.0f} ms; "
          f"learned accepted={out.accepted} gnc {out.report.gnc_iterations} {out.ms:.0f} ms; "
          f"same set {c['same_inlier_set_as_plain']} dt {c['max_translation_diff_vs_plain_m']:.2e} "
          f"tls learned {c['tls_cost_learned']:.3f} plain {c['tls_cost_plain']:.3f} "
          f"first {c['tls_cost_first_solve']:.3f}", flush=True)
    # the returned result always carries plain GNC's inlier set and poses
    assert np.array_equal(np.rint(out.weights), np.rint(wp))
    assert np.abs(out.poses[:, 9:] - pp[:, 9:]).max() <= learned.POSE_TOL
    if out.accepted:
        assert c["tls_fixed_point"] and c["binary_weights"] and c["tls_not_above_first_solve"]
        assert out.report.gnc_iterations < rp.gnc_iterations
    else:
        # a fallback costs the learned attempt plus the plain solve
        assert out.ms >= rp.ms_total
