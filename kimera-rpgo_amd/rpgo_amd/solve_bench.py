"""Time one robust solve on the synthetic SE(3) helix (the kernel-trace driver
of the solver profiles).

  python -m rpgo_amd.solve_bench [--poses 50000] [--gnc 0|1] [--pcg-graphs 0|1]

PCM (FP32 pairs + FP64 recheck, reference heuristic clique) on the helix with
loop closures 1/5 of the poses at 50 % outliers, then rpgo_solve on odometry +
prior + the clique's loop closures: GNC-TLS over LM (--gnc 1) or LM alone.
Prints one line with the device time and the iteration counts.
"""
from __future__ import annotations

import argparse
import sys

from . import _capi, learned, problem, synth


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--poses", type=int, default=50000)
    ap.add_argument("--outliers", type=float, default=0.5)
    ap.add_argument("--gnc", type=int, default=0, choices=[0, 1])
    ap.add_argument("--pcg-graphs", type=int, default=0, choices=[0, 1])
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    g = synth.se3_helix(n_poses=a.poses, n_lc=a.poses // 5, outlier_frac=a.outliers, seed=a.seed)
    with _capi.Context(0) as ctx:
        pcm = learned.pcm_stage(ctx, g)
        arrs = problem.solve_arrays(g, pcm.inliers)
        kw = dict(use_gnc=a.gnc, known_inliers=arrs[-1]) if a.gnc else {}
        _, _, rep = ctx.solve(*arrs[:-1], pcg_graphs=a.pcg_graphs, **kw)
    print(f"[solve_bench] poses {a.poses} loop closures {len(pcm.inliers)} gnc {rep.gnc_iterations} "
          f"lm {rep.lm_iterations} pcg {rep.pcg_iterations} linearizations {rep.linearizations} "
          f"ms {rep.ms_total:.1f} us_per_pcg {1e3 * rep.ms_total / max(rep.pcg_iterations, 1):.1f}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
