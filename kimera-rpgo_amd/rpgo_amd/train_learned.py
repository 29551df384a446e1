"""Train the learned GNC initialisation on the GPU.

  python -m rpgo_amd.train_learned [--graphs 60] [--epochs 40] [--out PATH]

Builds the training set with the plain method itself (PCM + GNC on seeded
graphs of 1k-4k poses, 10-60 % outliers: features of every loop closure the
solver sees, labels = its final GNC weight and log of the final mu), then
trains the MLP with the device Adam loop (rpgo_mlp_train) and writes the
parameters and the feature normalisation to an .npz file.  Prints one line
per graph and per training chunk.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import _capi, learned


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=60)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--mu-weight", type=float, default=0.1)
    ap.add_argument("--out", default=learned.MODEL_PATH)
    ap.add_argument("--dump", default="", help="also write the training rows (x, target, graph id) here")
    a = ap.parse_args(argv)
    ctx = _capi.Context(0)
    xs, ts, gid = [], [], []
    t0 = time.time()
    for g in range(a.graphs):
        gr = learned.training_graph(a.seed0 + g)
        rows = learned.graph_rows(ctx, gr)
        if rows is None:
            continue
        f, t, rep = rows
        xs.append(f)
        ts.append(t)
        gid.append(np.full(len(f), g))
        print(f"[train] graph {g} N={gr.n_poses} lc={gr.n_lc} solver_lc={len(f)} gnc={rep.gnc_iterations} "
              f"inliers={int(t[:, 0].sum())} mu_final={rep.mu_final:.3g} ({time.time() - t0:.0f}s)", flush=True)
    x = np.concatenate(xs).astype(np.float32)
    t = np.concatenate(ts).astype(np.float32)
    # balance the classes: outliers are the minority among the solver's loop closures
    pos = t[:, 0] > 0.5
    sw = np.where(pos, 0.5 / max(pos.mean(), 1e-3), 0.5 / max(1 - pos.mean(), 1e-3)).astype(np.float32)
    if a.dump:
        np.savez(a.dump, x=x, t=t, gid=np.concatenate(gid))
    norm = learned.fit_norm(x)
    model = learned.LearnedModel(learned.init_params(a.seed0), norm)
    ctx.mlp_set_params(model.params, model.norm)
    done = 0
    while done < a.epochs:
        chunk = min(10, a.epochs - done)
        hist = ctx.mlp_train(x, t, sw, a.mu_weight, batch=a.batch, epochs=chunk, lr=a.lr, seed=a.seed0 + done)
        done += chunk
        y = ctx.mlp_forward(x)
        acc = float(np.mean((y[:, 0] > 0) == pos))
        print(f"[train] epochs {done}/{a.epochs} loss {hist[-1]:.4f} train accuracy {acc:.4f} "
              f"rows {len(x)} ({time.time() - t0:.0f}s)", flush=True)
    model.params = ctx.mlp_get_params()
    model.save(a.out)
    print(f"[train] wrote {a.out}", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
