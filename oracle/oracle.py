"""ctypes front end of the FP64 CPU oracle (oracle/rpgo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, as the checker.  The product package never
imports this module.

The oracle restates Kimera-RPGO's PCM path in FP64; see the header of
rpgo_oracle.c for the reference file:line list.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librpgo_oracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, argt, rest in [
            ("orc_se3_compose", [_d, _d, _d], None),
            ("orc_se3_inverse", [_d, _d], None),
            ("orc_se3_between", [_d, _d, _d], None),
            ("orc_se3_adjoint", [_d, _d], None),
            ("orc_se3_logmap", [_d, _d], None),
            ("orc_se3_expmap", [_d, _d], None),
            ("orc_rot3_logmap", [_d, _d], None),
            ("orc_quat_to_rot", [ctypes.c_double] * 4 + [_d], None),
            ("orc_se2_compose", [_d, _d, _d], None),
            ("orc_se2_inverse", [_d, _d], None),
            ("orc_se2_between", [_d, _d, _d], None),
            ("orc_se2_adjoint", [_d, _d], None),
            ("orc_se2_logmap", [_d, _d], None),
            ("orc_pwc_size", [ctypes.c_int], ctypes.c_int),
            ("orc_pwc_from_measurement", [ctypes.c_int, _d, _d, _d], None),
            ("orc_pwc_compose", [ctypes.c_int, _d, _d, _d], None),
            ("orc_pwc_inverse", [ctypes.c_int, _d, _d], None),
            ("orc_pwc_between", [ctypes.c_int, _d, _d, _d], None),
            ("orc_pwc_mahalanobis", [ctypes.c_int, _d], ctypes.c_double),
            ("orc_llt_ok", [ctypes.c_int, _d], ctypes.c_int),
            ("orc_lu_inverse", [ctypes.c_int, _d, _d], None),
            ("orc_fold", [ctypes.c_int, ctypes.c_int, _d, _d, _d, _d], None),
            ("orc_odom_dist", [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d, _d],
             ctypes.c_double),
            ("orc_pair_dist", [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d, _d,
                               ctypes.c_int, ctypes.c_int, _d, _d], ctypes.c_double),
            ("orc_pcm_adjacency", [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _i32, _i32,
                                   _d, _d, ctypes.c_double, ctypes.c_double, _i32, _d,
                                   _u8, _d], ctypes.c_int),
            ("orc_pcm_dist_rows", [ctypes.c_int, _d, ctypes.c_int, _i32, _i32, _d, _d,
                                   ctypes.c_int, ctypes.c_int, _d], None),
            ("orc_csr_from_dense", [ctypes.c_int, _u8, _i32, _i32, ctypes.c_int], ctypes.c_int),
            ("orc_max_clique_heu", [ctypes.c_int, _i32, _i32, _i32], ctypes.c_int),
            ("orc_max_clique_heu_incremental", [ctypes.c_int, _i32, _i32, ctypes.c_int,
                                                ctypes.c_int, _i32], ctypes.c_int),
            ("orc_max_clique_exact", [ctypes.c_int, _i32, _i32, _i32], ctypes.c_int),
            ("orc_max_clique_heu_dense", [ctypes.c_int, _u8, _i32], ctypes.c_int),
            ("orc_pwn_norms", [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d, _d], None),
            ("orc_simple_odom", [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d, _d, _d],
             ctypes.c_double),
            ("orc_simple_pair", [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d, _d,
                                 ctypes.c_int, ctypes.c_int, _d, _d, _d], ctypes.c_double),
            ("orc_pcm_simple_adjacency", [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _i32, _i32,
                                          _d, _d] + [ctypes.c_double] * 4 + [_i32, _d, _u8, _d],
             ctypes.c_int),
            ("orc_fold_chains", [ctypes.c_int, ctypes.c_int, _i32, _d, _d, _d, _d], None),
            ("orc_pcm_adjacency_chains", [ctypes.c_int, ctypes.c_int, _i32, _d, ctypes.c_int, _i32, _i32, _d, _d,
                                          ctypes.c_int] + [ctypes.c_double] * 4 + [_i32, _d, _u8, _d],
             ctypes.c_int),
            ("orc_landmark_pair_dist", [ctypes.c_int, _d, ctypes.c_int, _d, _d, ctypes.c_int, _d, _d],
             ctypes.c_double),
            ("orc_simple_landmark_pair", [ctypes.c_int, _d, ctypes.c_int, _d, _d, ctypes.c_int, _d, _d, _d],
             ctypes.c_double),
            ("orc_pcm_adjacency_landmark", [ctypes.c_int, _d, ctypes.c_int, _i32, _d, _d, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, _u8, _d], None),
            ("orc_cross_between", [ctypes.c_int, _d, _d, _d, _d, _d], None),
            ("orc_adj_growth", [ctypes.c_int, _d, ctypes.c_int, _i32, _i32, _d, _d, ctypes.c_double,
                                ctypes.c_int], ctypes.c_longlong),
            ("orc_set_threads", [ctypes.c_int], None),
            ("orc_get_threads", [], ctypes.c_int),
            ("orc_pair_dist_list", [ctypes.c_int, _d, _i32, _i32, _d, _d, ctypes.c_int, _i32, _i32, _d],
             None),
        ]:
            f = getattr(L, name)
            f.argtypes = argt
            f.restype = rest
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def pose_len(dim):
    return 12 if dim == 3 else 4


def tdim(dim):
    return 6 if dim == 3 else 3


# ---------------------------------------------------------------- geometry
def se3_compose(a, b):
    o = np.zeros(12); lib().orc_se3_compose(_p(_f64(a)), _p(_f64(b)), _p(o)); return o


def se3_inverse(a):
    o = np.zeros(12); lib().orc_se3_inverse(_p(_f64(a)), _p(o)); return o


def se3_between(a, b):
    o = np.zeros(12); lib().orc_se3_between(_p(_f64(a)), _p(_f64(b)), _p(o)); return o


def se3_adjoint(a):
    o = np.zeros(36); lib().orc_se3_adjoint(_p(_f64(a)), _p(o)); return o.reshape(6, 6)


def se3_logmap(a):
    o = np.zeros(6); lib().orc_se3_logmap(_p(_f64(a)), _p(o)); return o


def se3_expmap(xi):
    o = np.zeros(12); lib().orc_se3_expmap(_p(_f64(xi)), _p(o)); return o


def rot3_logmap(R):
    o = np.zeros(3); lib().orc_rot3_logmap(_p(_f64(np.asarray(R).reshape(9))), _p(o)); return o


def quat_to_rot(w, x, y, z):
    o = np.zeros(9); lib().orc_quat_to_rot(w, x, y, z, _p(o)); return o.reshape(3, 3)


def se2_compose(a, b):
    o = np.zeros(4); lib().orc_se2_compose(_p(_f64(a)), _p(_f64(b)), _p(o)); return o


def se2_inverse(a):
    o = np.zeros(4); lib().orc_se2_inverse(_p(_f64(a)), _p(o)); return o


def se2_logmap(a):
    o = np.zeros(3); lib().orc_se2_logmap(_p(_f64(a)), _p(o)); return o


def pwc(dim, pose, cov):
    o = np.zeros(lib().orc_pwc_size(dim))
    lib().orc_pwc_from_measurement(dim, _p(_f64(pose)), _p(_f64(np.asarray(cov).reshape(-1))), _p(o))
    return o


def pwc_raw(dim, pose, cov, rot_info=True):
    """PoseWithCovariance built field by field (no NaN handling)."""
    return np.concatenate([_f64(pose), _f64(np.asarray(cov).reshape(-1)),
                           [1.0 if rot_info else 0.0]])


def pwc_split(dim, p):
    pl, d = pose_len(dim), tdim(dim)
    return p[:pl].copy(), p[pl:pl + d * d].reshape(d, d).copy(), bool(p[pl + d * d] != 0.0)


def pwc_compose(dim, a, b):
    o = np.zeros(lib().orc_pwc_size(dim)); lib().orc_pwc_compose(dim, _p(_f64(a)), _p(_f64(b)), _p(o)); return o


def pwc_inverse(dim, a):
    o = np.zeros(lib().orc_pwc_size(dim)); lib().orc_pwc_inverse(dim, _p(_f64(a)), _p(o)); return o


def pwc_between(dim, a, b):
    o = np.zeros(lib().orc_pwc_size(dim)); lib().orc_pwc_between(dim, _p(_f64(a)), _p(_f64(b)), _p(o)); return o


def pwc_mahalanobis(dim, a):
    return lib().orc_pwc_mahalanobis(dim, _p(_f64(a)))


def llt_ok(A):
    A = _f64(A); return bool(lib().orc_llt_ok(A.shape[0], _p(A)))


def lu_inverse(A):
    A = _f64(A); o = np.zeros_like(A); lib().orc_lu_inverse(A.shape[0], _p(A), _p(o)); return o


# ---------------------------------------------------------------- pipeline
def fold(dim, pose0, odom, odom_cov):
    """Trajectory fold: returns (N, pwc_size) array of PoseWithCovariance records."""
    odom = _f64(odom).reshape(-1, pose_len(dim))
    odom_cov = _f64(odom_cov).reshape(-1, tdim(dim) ** 2)
    n = odom.shape[0] + 1
    out = np.zeros((n, lib().orc_pwc_size(dim)))
    lib().orc_fold(dim, n, _p(_f64(pose0)), _p(odom), _p(odom_cov), _p(out))
    return out


def odom_dist(dim, traj, frm, to, lc_pose, lc_cov):
    return lib().orc_odom_dist(dim, _p(_f64(traj)), int(frm), int(to), _p(_f64(lc_pose)),
                               _p(_f64(np.asarray(lc_cov).reshape(-1))))


def pair_dist(dim, traj, a, b, zab, sab, c, d, zcd, scd):
    return lib().orc_pair_dist(dim, _p(_f64(traj)), int(a), int(b), _p(_f64(zab)),
                               _p(_f64(np.asarray(sab).reshape(-1))), int(c), int(d),
                               _p(_f64(zcd)), _p(_f64(np.asarray(scd).reshape(-1))))


def pcm_adjacency(dim, traj, lc_from, lc_to, lc_pose, lc_cov, odom_thr, lc_thr, want_dist=True):
    """Reference adjacency stage for one group.

    Returns (accepted_idx, odom_dist, adj (m,m) uint8, dist (m,m) float64 or None)."""
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    n = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(n, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(n, tdim(dim) ** 2)
    acc = np.zeros(max(n, 1), dtype=np.int32)
    od = np.zeros(max(n, 1))
    adj = np.zeros((max(n, 1), max(n, 1)), dtype=np.uint8)
    dist = np.zeros((max(n, 1), max(n, 1))) if want_dist else None
    m = lib().orc_pcm_adjacency(dim, traj.shape[0], _p(traj), n, _p(lc_from, _i32), _p(lc_to, _i32),
                                _p(lc_pose), _p(lc_cov), float(odom_thr), float(lc_thr),
                                _p(acc, _i32), _p(od),
                                _p(adj, _u8), _p(dist) if dist is not None else None)
    adj = adj.reshape(-1)[: m * m].reshape(m, m).copy()
    if dist is not None:
        dist = dist.reshape(-1)[: m * m].reshape(m, m).copy()
    return acc[:m].copy(), od[:n].copy(), adj, dist


# ---------------------------------------------------------------- PcmSimple
def pwn_norms(dim, pose, node, rot_info=True):
    """(avg_trans_norm, avg_rot_norm) of a PoseWithNode (GeometryUtils.h:275-289)."""
    t = ctypes.c_double()
    r = ctypes.c_double()
    lib().orc_pwn_norms(dim, _p(_f64(pose)), int(node), int(bool(rot_info)), ctypes.byref(t),
                        ctypes.byref(r))
    return t.value, r.value


def simple_odom(dim, traj, frm, to, lc_pose, lc_cov):
    """(translation, rotation) norms of the PcmSimple odometry check."""
    r = ctypes.c_double()
    t = lib().orc_simple_odom(dim, _p(_f64(traj)), int(frm), int(to), _p(_f64(lc_pose)),
                              _p(_f64(np.asarray(lc_cov).reshape(-1))), ctypes.byref(r))
    return t, r.value


def simple_pair(dim, traj, a, b, zab, sab, c, d, zcd, scd):
    r = ctypes.c_double()
    t = lib().orc_simple_pair(dim, _p(_f64(traj)), int(a), int(b), _p(_f64(zab)),
                              _p(_f64(np.asarray(sab).reshape(-1))), int(c), int(d),
                              _p(_f64(zcd)), _p(_f64(np.asarray(scd).reshape(-1))), ctypes.byref(r))
    return t, r.value


def pcm_simple_adjacency(dim, traj, lc_from, lc_to, lc_pose, lc_cov, odom_trans, odom_rot,
                         dist_trans, dist_rot, want_dist=True):
    """PcmSimple adjacency stage; same return layout as pcm_adjacency (distances are
    the translation norms)."""
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    n = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(n, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(n, tdim(dim) ** 2)
    acc = np.zeros(max(n, 1), dtype=np.int32)
    od = np.zeros(max(n, 1))
    adj = np.zeros((max(n, 1), max(n, 1)), dtype=np.uint8)
    dist = np.zeros((max(n, 1), max(n, 1))) if want_dist else None
    m = lib().orc_pcm_simple_adjacency(dim, traj.shape[0], _p(traj), n, _p(lc_from, _i32),
                                       _p(lc_to, _i32), _p(lc_pose), _p(lc_cov), float(odom_trans),
                                       float(odom_rot), float(dist_trans), float(dist_rot),
                                       _p(acc, _i32), _p(od), _p(adj, _u8),
                                       _p(dist) if dist is not None else None)
    adj = adj.reshape(-1)[: m * m].reshape(m, m).copy()
    if dist is not None:
        dist = dist.reshape(-1)[: m * m].reshape(m, m).copy()
    return acc[:m].copy(), od[:n].copy(), adj, dist


# ---------------------------------------------------------------- several chains
def fold_chains(dim, chain_off, pose0s, odom, odom_cov):
    chain_off = np.ascontiguousarray(chain_off, dtype=np.int32)
    nc = chain_off.shape[0] - 1
    odom = _f64(odom).reshape(-1, pose_len(dim))
    odom_cov = _f64(odom_cov).reshape(-1, tdim(dim) ** 2)
    n = int(chain_off[-1])
    assert odom.shape[0] == n - nc
    out = np.zeros((n, lib().orc_pwc_size(dim)))
    lib().orc_fold_chains(dim, nc, _p(chain_off, _i32), _p(_f64(pose0s)), _p(odom), _p(odom_cov), _p(out))
    return out


def pcm_adjacency_chains(dim, chain_off, traj, lc_from, lc_to, lc_pose, lc_cov, odom_thr, lc_thr,
                         simple=None, want_dist=True):
    """Adjacency stage of one loop-closure group over several chains.  simple:
    None (Mahalanobis, odom_thr / lc_thr) or (odom_trans, odom_rot, dist_trans,
    dist_rot).  Same return layout as pcm_adjacency."""
    chain_off = np.ascontiguousarray(chain_off, dtype=np.int32)
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    n = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(n, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(n, tdim(dim) ** 2)
    acc = np.zeros(max(n, 1), dtype=np.int32)
    od = np.zeros(max(n, 1))
    adj = np.zeros((max(n, 1), max(n, 1)), dtype=np.uint8)
    dist = np.zeros((max(n, 1), max(n, 1))) if want_dist else None
    thr = (odom_thr, 0.0, lc_thr, 0.0) if simple is None else tuple(simple)
    m = lib().orc_pcm_adjacency_chains(dim, chain_off.shape[0] - 1, _p(chain_off, _i32), _p(traj), n,
                                       _p(lc_from, _i32), _p(lc_to, _i32), _p(lc_pose), _p(lc_cov),
                                       0 if simple is None else 1, *[float(v) for v in thr], _p(acc, _i32),
                                       _p(od), _p(adj, _u8), _p(dist) if dist is not None else None)
    adj = adj.reshape(-1)[: m * m].reshape(m, m).copy()
    if dist is not None:
        dist = dist.reshape(-1)[: m * m].reshape(m, m).copy()
    return acc[:m].copy(), od[:n].copy(), adj, dist


def cross_between(dim, a0, a, b0, b):
    """Trajectory::getBetween of keys on two robots (PoseWithCovariance records)."""
    out = np.zeros(lib().orc_pwc_size(dim))
    lib().orc_cross_between(dim, _p(_f64(a0)), _p(_f64(a)), _p(_f64(b0)), _p(_f64(b)), _p(out))
    return out


def landmark_pair_dist(dim, traj, i, zil, sil, j, zjl, sjl):
    return lib().orc_landmark_pair_dist(dim, _p(_f64(traj)), int(i), _p(_f64(zil)), _p(_f64(sil)), int(j),
                                        _p(_f64(zjl)), _p(_f64(sjl)))


def simple_landmark_pair(dim, traj, i, zil, sil, j, zjl, sjl):
    """(translation norm, rotation norm)"""
    r = ctypes.c_double(0.0)
    t = lib().orc_simple_landmark_pair(dim, _p(_f64(traj)), int(i), _p(_f64(zil)), _p(_f64(sil)), int(j),
                                       _p(_f64(zjl)), _p(_f64(sjl)), ctypes.byref(r))
    return t, r.value


def pcm_adjacency_landmark(dim, traj, lc_from, lc_pose, lc_cov, lc_thr, simple=None, want_dist=True):
    """Adjacency of one landmark's observations from poses lc_from (Pcm.h:775-844).
    simple: None (Mahalanobis, lc_thr) or (dist_trans, dist_rot).  Returns (adj, dist)."""
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    n = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(n, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(n, tdim(dim) ** 2)
    adj = np.zeros((max(n, 1), max(n, 1)), dtype=np.uint8)
    dist = np.zeros((max(n, 1), max(n, 1))) if want_dist else None
    thr = (lc_thr, 0.0) if simple is None else tuple(simple)
    lib().orc_pcm_adjacency_landmark(dim, _p(traj), n, _p(lc_from, _i32), _p(lc_pose), _p(lc_cov),
                                     0 if simple is None else 1, float(thr[0]), float(thr[1]), _p(adj, _u8),
                                     _p(dist) if dist is not None else None)
    adj = adj.reshape(-1)[: n * n].reshape(n, n).copy()
    if dist is not None:
        dist = dist.reshape(-1)[: n * n].reshape(n, n).copy()
    return adj, dist


def pcm_dist_rows(dim, traj, lc_from, lc_to, lc_pose, lc_cov, j0=None, j1=None):
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    m = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(m, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(m, tdim(dim) ** 2)
    j0 = 0 if j0 is None else j0
    j1 = m if j1 is None else j1
    dist = np.zeros((m, m))
    lib().orc_pcm_dist_rows(dim, _p(traj), m, _p(lc_from, _i32), _p(lc_to, _i32), _p(lc_pose),
                            _p(lc_cov), j0, j1, _p(dist))
    return dist


def adj_growth(dim, traj, lc_from, lc_to, lc_pose, lc_cov, lc_thr, do_pairs=True):
    """incrementAdjMatrix's allocate/copy/fill/assign pattern (Pcm.h:737-767)
    over the given loop closures, one thread; returns the consistent pairs."""
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    m = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(m, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(m, tdim(dim) ** 2)
    n = lib().orc_adj_growth(dim, _p(traj), m, _p(lc_from, _i32), _p(lc_to, _i32), _p(lc_pose),
                             _p(lc_cov), float(lc_thr), 1 if do_pairs else 0)
    if n < 0:
        raise MemoryError("orc_adj_growth: allocation failed")
    return int(n)


# ---------------------------------------------------------------- cliques
def csr_from_dense(adj):
    adj = np.ascontiguousarray(adj, dtype=np.uint8)
    n = adj.shape[0]
    off = np.zeros(n + 1, dtype=np.int32)
    cap = int(adj.sum()) + 1
    edg = np.zeros(cap, dtype=np.int32)
    e = lib().orc_csr_from_dense(n, _p(adj, _u8), _p(off, _i32), _p(edg, _i32), cap)
    return off, edg[:e].copy()


def max_clique_heu(adj):
    """findMaxCliqueHeu: returns the first `size` entries of the reference buffer."""
    off, edg = csr_from_dense(adj)
    n = adj.shape[0]
    out = np.zeros(max(n, 1), dtype=np.int32)
    k = lib().orc_max_clique_heu(n, _p(off, _i32), _p(np.ascontiguousarray(edg) if edg.size else np.zeros(1, np.int32), _i32), _p(out, _i32))
    return int(k), out[: max(k, 0)].copy()


def max_clique_heu_incremental(adj, num_new, prev_size):
    off, edg = csr_from_dense(adj)
    n = adj.shape[0]
    out = np.zeros(max(n, 1), dtype=np.int32)
    e = np.ascontiguousarray(edg) if edg.size else np.zeros(1, np.int32)
    k = lib().orc_max_clique_heu_incremental(n, _p(off, _i32), _p(e, _i32), int(num_new),
                                             int(prev_size), _p(out, _i32))
    return int(k), out[: max(k, 0)].copy()


def max_clique_exact(adj):
    off, edg = csr_from_dense(adj)
    n = adj.shape[0]
    out = np.zeros(max(n, 1), dtype=np.int32)
    e = np.ascontiguousarray(edg) if edg.size else np.zeros(1, np.int32)
    k = lib().orc_max_clique_exact(n, _p(off, _i32), _p(e, _i32), _p(out, _i32))
    return int(k), out[: max(k, 0)].copy()


def max_clique_heu_dense(adj):
    """maxCliqueHeu semantics evaluated with bitset neighbourhood tests."""
    adj = np.ascontiguousarray(adj, dtype=np.uint8)
    n = adj.shape[0]
    out = np.zeros(max(n, 1), dtype=np.int32)
    k = lib().orc_max_clique_heu_dense(n, _p(adj, _u8), _p(out, _i32))
    return int(k), out[: max(k, 0)].copy()


def pair_dist_list(dim, traj, lc_from, lc_to, lc_pose, lc_cov, pi, pj):
    """Distances of explicit (older i, newer j) pairs of an accepted LC list."""
    traj = _f64(traj)
    lc_from = np.ascontiguousarray(lc_from, dtype=np.int32)
    lc_to = np.ascontiguousarray(lc_to, dtype=np.int32)
    m = lc_from.shape[0]
    lc_pose = _f64(lc_pose).reshape(m, pose_len(dim))
    lc_cov = _f64(lc_cov).reshape(m, tdim(dim) ** 2)
    pi = np.ascontiguousarray(pi, dtype=np.int32)
    pj = np.ascontiguousarray(pj, dtype=np.int32)
    out = np.zeros(max(len(pi), 1))
    lib().orc_pair_dist_list(dim, _p(traj), _p(lc_from, _i32), _p(lc_to, _i32), _p(lc_pose),
                             _p(lc_cov), len(pi), _p(pi, _i32), _p(pj, _i32), _p(out))
    return out[: len(pi)].copy()


def set_threads(n):
    """OpenMP threads of the distance loops (per-pair arithmetic is unchanged)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())
