"""Shared test graphs: the same factor list as an oracle Graph and as the
flat arrays of the C ABI (rpgo_solve)."""
import os

import numpy as np

import solver_oracle as S
from rpgo_amd import g2o, synth

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")


def from_synth(gr: synth.PoseGraph, lc_mask=None, prior_var=0.01):
    """Factor order of Pcm.h:977-1005: odometry, the prior (special), loop
    closures (those with lc_mask True)."""
    n = gr.n_poses
    lc = np.arange(gr.n_lc) if lc_mask is None else np.nonzero(lc_mask)[0]
    Rv, tv = synth.se3_unpack(gr.values)
    zR, zt = synth.se3_unpack(np.concatenate([gr.odom, gr.lc_pose[lc]]))
    bi = np.concatenate([np.arange(n - 1), gr.lc_from[lc]])
    bj = np.concatenate([np.arange(1, n), gr.lc_to[lc]])
    info = np.linalg.inv(np.concatenate([gr.odom_cov, gr.lc_cov[lc]]).reshape(-1, 6, 6))
    Rg, tg = synth.se3_unpack(gr.gt[:1])
    n_odom = n - 1
    order = [("b", k) for k in range(n_odom)] + [("p", 0)] + [("b", n_odom + k) for k in range(len(lc))]
    g = S.Graph(R=Rv.copy(), t=tv.copy(), bi=bi, bj=bj, zR=zR, zt=zt, binfo=info,
                pk=np.array([0]), pR=Rg.copy(), pt=tg.copy(), pinfo=(np.eye(6) / prior_var)[None],
                order=order)
    return g, n_odom + 1


def from_robot_a():
    gg = g2o.read(os.path.join(DATA, "robot_a.g2o"))
    keys = gg.keys
    idx = {k: n for n, k in enumerate(keys)}
    R = np.stack([gg.poses[k][:9].reshape(3, 3) for k in keys])
    t = np.stack([gg.poses[k][9:] for k in keys])
    odom = [e for e in gg.edges if e[1] == e[0] + 1]
    lcs = [e for e in gg.edges if e[1] != e[0] + 1]
    edges = odom + lcs
    bi = np.array([idx[e[0]] for e in edges]); bj = np.array([idx[e[1]] for e in edges])
    zR = np.stack([e[2][:9].reshape(3, 3) for e in edges]); zt = np.stack([e[2][9:] for e in edges])
    info = np.stack([e[3] for e in edges])
    order = [("b", k) for k in range(len(odom))] + [("p", 0)] + \
            [("b", len(odom) + k) for k in range(len(lcs))]
    g = S.Graph(R=R, t=t, bi=bi, bj=bj, zR=zR, zt=zt, binfo=info, pk=np.array([0]),
                pR=R[:1].copy(), pt=t[:1].copy(), pinfo=(np.eye(6) / 0.01)[None], order=order)
    return g, len(odom) + 1


def flat(g: S.Graph):
    """(poses, kind, key_i, key_j, measured, information) in factor order."""
    F = g.F
    kind = np.zeros(F, np.int32); ki = np.zeros(F, np.int32); kj = np.zeros(F, np.int32)
    meas = np.zeros((F, 12)); info = np.zeros((F, 36))
    for q, (typ, k) in enumerate(g.order):
        if typ == "b":
            kind[q] = 0; ki[q] = g.bi[k]; kj[q] = g.bj[k]
            meas[q] = np.concatenate([g.zR[k].reshape(9), g.zt[k]])
            info[q] = g.binfo[k].reshape(36)
        else:
            kind[q] = 1; ki[q] = g.pk[k]; kj[q] = g.pk[k]
            meas[q] = np.concatenate([g.pR[k].reshape(9), g.pt[k]])
            info[q] = g.pinfo[k].reshape(36)
    poses = np.concatenate([g.R.reshape(-1, 9), g.t], 1)
    return poses, kind, ki, kj, meas, info


def pose_error(poses, R, t):
    """max translation error (m) and max rotation error (rad) against (R, t)."""
    Rp = poses[:, :9].reshape(-1, 3, 3)
    tp = poses[:, 9:]
    dt = np.abs(tp - t).max()
    dR = np.swapaxes(R, -1, -2) @ Rp
    ang = np.linalg.norm(S.rot_log(dR), axis=-1).max()
    return float(dt), float(ang)


def graph2_from_synth(gr: synth.PoseGraph, lc_idx, prior_var=0.01):
    """Pose2 oracle graph in the factor order of Pcm.h:977-1005: odometry, the
    prior on pose 0 (at the ground truth), the given loop closures in order."""
    n = gr.n_poses
    lc = np.asarray(lc_idx, dtype=np.int64)
    th = lambda p: np.arctan2(p[..., 1], p[..., 0])
    z = np.concatenate([gr.odom, gr.lc_pose[lc]])
    info = np.linalg.inv(np.concatenate([gr.odom_cov, gr.lc_cov[lc]]).reshape(-1, 3, 3))
    n_odom = n - 1
    order = [("b", k) for k in range(n_odom)] + [("p", 0)] + [("b", n_odom + k) for k in range(len(lc))]
    return S.Graph2(th=th(gr.values).copy(), xy=gr.values[:, 2:].copy(),
                    bi=np.concatenate([np.arange(n_odom), gr.lc_from[lc]]),
                    bj=np.concatenate([np.arange(1, n), gr.lc_to[lc]]), zth=th(z), zxy=z[:, 2:],
                    binfo=info, pk=np.array([0]), pth=th(gr.gt[:1]), pxy=gr.gt[:1, 2:].copy(),
                    pinfo=(np.eye(3) / prior_var)[None], order=order), n_odom + 1
