"""Seeded synthetic pose graphs for the PCM / robust-solve workloads.

Layouts match the C-ABI (include/rpgo_hip.h):
  SE(3) pose   12 doubles: R row-major (9) then t (3)
  SE(2) pose    4 doubles: cos, sin, x, y
  SE(3) cov    36 doubles, tangent order (omega, v)
  SE(2) cov     9 doubles, tangent order (x, y, theta)

Workloads (sizes and noise as stated in each function's docstring):
  se2_random_walk   500-pose 2D random walk, 60 inlier + 15 outlier LCs
  se3_helix         helical trajectory, N poses, n LCs, outlier fraction
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# ----------------------------------------------------------------- SO(3) helpers
def _skew(w):
    w = np.asarray(w)
    z = np.zeros(w.shape[:-1])
    return np.stack([np.stack([z, -w[..., 2], w[..., 1]], -1),
                     np.stack([w[..., 2], z, -w[..., 0]], -1),
                     np.stack([-w[..., 1], w[..., 0], z], -1)], -2)


def rot_exp(w):
    """Batched Rodrigues formula, w (..., 3) -> (..., 3, 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1)[..., None, None]
    K = _skew(w)
    small = th < 1e-12
    ths = np.where(small, 1.0, th)
    A = np.where(small, 1.0, np.sin(ths) / ths)
    B = np.where(small, 0.5, (1.0 - np.cos(ths)) / (ths * ths))
    return np.eye(3) + A * K + B * (K @ K)


def random_rotations(rng, n):
    """Uniform random rotations (uniform unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def quat_to_rot(q):
    """Unit quaternions (w, x, y, z) (..., 4) -> rotation matrices (..., 3, 3)."""
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def rot_to_quat(R):
    """Rotation matrices (..., 3, 3) -> unit quaternions (w, x, y, z), w >= 0.
    Shepperd's method: each rotation divides by its largest of 4|w|^2,
    4|x|^2, 4|y|^2, 4|z|^2, so the result is accurate to rounding at every
    angle, pi included."""
    R = np.asarray(R, dtype=np.float64)
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    cand = np.stack([tr, R[..., 0, 0], R[..., 1, 1], R[..., 2, 2]], -1)
    k = np.argmax(cand, -1)
    q = np.zeros(R.shape[:-2] + (4,))
    for case in range(4):
        sel = k == case
        if not np.any(sel):
            continue
        M = R[sel]
        if case == 0:
            s = 2.0 * np.sqrt(1.0 + tr[sel])
            v = [0.25 * s, (M[:, 2, 1] - M[:, 1, 2]) / s, (M[:, 0, 2] - M[:, 2, 0]) / s, (M[:, 1, 0] - M[:, 0, 1]) / s]
        elif case == 1:
            s = 2.0 * np.sqrt(1.0 + M[:, 0, 0] - M[:, 1, 1] - M[:, 2, 2])
            v = [(M[:, 2, 1] - M[:, 1, 2]) / s, 0.25 * s, (M[:, 0, 1] + M[:, 1, 0]) / s, (M[:, 0, 2] + M[:, 2, 0]) / s]
        elif case == 2:
            s = 2.0 * np.sqrt(1.0 + M[:, 1, 1] - M[:, 0, 0] - M[:, 2, 2])
            v = [(M[:, 0, 2] - M[:, 2, 0]) / s, (M[:, 0, 1] + M[:, 1, 0]) / s, 0.25 * s, (M[:, 1, 2] + M[:, 2, 1]) / s]
        else:
            s = 2.0 * np.sqrt(1.0 + M[:, 2, 2] - M[:, 0, 0] - M[:, 1, 1])
            v = [(M[:, 1, 0] - M[:, 0, 1]) / s, (M[:, 0, 2] + M[:, 2, 0]) / s, (M[:, 1, 2] + M[:, 2, 1]) / s, 0.25 * s]
        q[sel] = np.stack(v, -1)
    q *= np.where(q[..., :1] < 0, -1.0, 1.0)
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def se3_pack(R, t):
    R = np.asarray(R); t = np.asarray(t)
    return np.concatenate([R.reshape(R.shape[:-2] + (9,)), t], axis=-1)


def se3_unpack(p):
    p = np.asarray(p)
    return p[..., :9].reshape(p.shape[:-1] + (3, 3)), p[..., 9:12]


def se3_rel(Ra, ta, Rb, tb):
    """a^-1 b for batches."""
    RaT = np.swapaxes(Ra, -1, -2)
    return RaT @ Rb, (RaT @ (tb - ta)[..., None])[..., 0]


def se3_mul(Ra, ta, Rb, tb):
    return Ra @ Rb, ta + (Ra @ tb[..., None])[..., 0]


@dataclass
class PoseGraph:
    """A single-prefix pose graph: odometry chain 0..N-1 plus loop closures."""
    dim: int
    gt: np.ndarray            # (N, 12|4) ground-truth poses
    values: np.ndarray        # (N, 12|4) initial estimates (odometry dead reckoning)
    odom: np.ndarray          # (N-1, 12|4) odometry measurements k -> k+1
    odom_cov: np.ndarray      # (N-1, 36|9)
    lc_from: np.ndarray       # (n,) int32
    lc_to: np.ndarray         # (n,) int32
    lc_pose: np.ndarray       # (n, 12|4)
    lc_cov: np.ndarray        # (n, 36|9)
    is_outlier: np.ndarray    # (n,) bool
    meta: dict = field(default_factory=dict)

    @property
    def n_poses(self):
        return int(self.gt.shape[0])

    @property
    def n_lc(self):
        return int(self.lc_from.shape[0])

    @property
    def pose0(self):
        return self.values[0]


def _dead_reckon3(R0, t0, Rz, tz):
    n = Rz.shape[0] + 1
    R = np.empty((n, 3, 3)); t = np.empty((n, 3))
    R[0], t[0] = R0, t0
    for k in range(n - 1):
        R[k + 1] = R[k] @ Rz[k]
        t[k + 1] = t[k] + R[k] @ tz[k]
    return R, t


def se3_helix(n_poses=5000, n_lc=1000, outlier_frac=0.2, seed=0, sigma_t=0.02,
              sigma_r=0.005, period=100, pitch=2.0, jitter=0.1, lc_radius=3.0,
              outlier_range=20.0, init="odometry"):
    """Helical SE(3) trajectory with seeded loop closures.

    Ground truth: pose k sits on a helix of `period` poses per turn (radius
    period/2pi so consecutive poses are ~1 m apart, `pitch` m climb per turn),
    heading along the helix, with a per-pose rotation jitter N(0, jitter) per
    axis (not accumulated).  Odometry and inlier loop closures are the true
    relative poses perturbed on the right by Exp(N(0, sigma_r) x3, N(0, sigma_t) x3).
    Inlier LCs join poses one turn apart (+-2 poses) that are within
    `lc_radius` m; outliers join random poses |i-j| >= 2 with translation
    U[-outlier_range, outlier_range]^3 and a uniformly random rotation.
    LC direction (later->earlier or earlier->later) and order are shuffled.
    Initial values: init="odometry" dead-reckons the odometry from pose 0 (what
    a batch load sees); init="perturbed" perturbs every true pose by
    Exp(N(0, sigma_r) x3, N(0, sigma_t) x3) (an estimate carried over from an
    earlier optimisation).  The random stream of the graph is the same for both.
    """
    rng = np.random.default_rng(seed)
    k = np.arange(n_poses, dtype=np.float64)
    ang = 2 * math.pi * k / period
    r = period / (2 * math.pi)
    pos = np.stack([r * np.cos(ang), r * np.sin(ang), pitch * k / period], -1)
    yaw = ang + math.pi / 2
    Ryaw = np.zeros((n_poses, 3, 3))
    Ryaw[:, 0, 0] = np.cos(yaw); Ryaw[:, 0, 1] = -np.sin(yaw)
    Ryaw[:, 1, 0] = np.sin(yaw); Ryaw[:, 1, 1] = np.cos(yaw); Ryaw[:, 2, 2] = 1
    Rgt = Ryaw @ rot_exp(rng.normal(0, jitter, (n_poses, 3)))
    tgt = pos

    def noisy_rel(i, j):
        Rr, tr = se3_rel(Rgt[i], tgt[i], Rgt[j], tgt[j])
        m = len(np.atleast_1d(i))
        Rn = rot_exp(rng.normal(0, sigma_r, (m, 3)))
        tn = rng.normal(0, sigma_t, (m, 3))
        return se3_mul(Rr, tr, Rn, tn)

    i0 = np.arange(n_poses - 1)
    Rz, tz = noisy_rel(i0, i0 + 1)
    cov = np.diag([sigma_r ** 2] * 3 + [sigma_t ** 2] * 3).reshape(36)
    odom_cov = np.tile(cov, (n_poses - 1, 1))

    n_out = int(round(n_lc * outlier_frac))
    n_in = n_lc - n_out
    # inlier candidates: one turn apart, +-2 poses, within lc_radius
    cand = []
    for d in range(-2, 3):
        a = np.arange(0, n_poses - period - d) if period + d > 0 else np.arange(0)
        a = a[(a + period + d) < n_poses]
        b = a + period + d
        dist = np.linalg.norm(tgt[a] - tgt[b], axis=1)
        ok = dist <= lc_radius
        cand.append(np.stack([a[ok], b[ok]], 1))
    cand = np.concatenate(cand, 0)
    if cand.shape[0] < n_in:
        raise ValueError(f"only {cand.shape[0]} inlier candidates for {n_in} loop closures")
    pick = cand[rng.choice(cand.shape[0], n_in, replace=False)]
    ia, ib = pick[:, 0], pick[:, 1]
    Rin, tin = noisy_rel(ia, ib)
    # outliers
    oa = rng.integers(0, n_poses, n_out)
    ob = rng.integers(0, n_poses, n_out)
    bad = np.abs(oa - ob) < 2
    while bad.any():
        ob[bad] = rng.integers(0, n_poses, bad.sum())
        bad = np.abs(oa - ob) < 2
    Rout = random_rotations(rng, n_out)
    tout = rng.uniform(-outlier_range, outlier_range, (n_out, 3))

    frm = np.concatenate([ia, oa]).astype(np.int32)
    to = np.concatenate([ib, ob]).astype(np.int32)
    Rl = np.concatenate([Rin, Rout]); tl = np.concatenate([tin, tout])
    is_out = np.concatenate([np.zeros(n_in, bool), np.ones(n_out, bool)])
    # random direction: flip (a,b,Z) -> (b,a,Z^-1)
    flip = rng.uniform(size=n_lc) < 0.5
    RlT = np.swapaxes(Rl, -1, -2)
    tinv = -(RlT @ tl[..., None])[..., 0]
    Rl = np.where(flip[:, None, None], RlT, Rl)
    tl = np.where(flip[:, None], tinv, tl)
    frm2 = np.where(flip, to, frm).astype(np.int32)
    to2 = np.where(flip, frm, to).astype(np.int32)
    perm = rng.permutation(n_lc)
    lc_cov = np.tile(cov, (n_lc, 1))

    if init == "perturbed":
        rng2 = np.random.default_rng([seed, 1])
        Rv = Rgt @ rot_exp(rng2.normal(0, sigma_r, (n_poses, 3)))
        tv = tgt + rng2.normal(0, sigma_t, (n_poses, 3))
        Rv[0], tv[0] = Rgt[0], tgt[0]
    else:
        Rv, tv = _dead_reckon3(Rgt[0], tgt[0], Rz, tz)
    return PoseGraph(
        dim=3, gt=se3_pack(Rgt, tgt), values=se3_pack(Rv, tv), odom=se3_pack(Rz, tz),
        odom_cov=odom_cov, lc_from=frm2[perm], lc_to=to2[perm],
        lc_pose=se3_pack(Rl, tl)[perm], lc_cov=lc_cov[perm], is_outlier=is_out[perm],
        meta=dict(kind="se3_helix", n_poses=n_poses, n_lc=n_lc, outlier_frac=outlier_frac,
                  seed=seed, sigma_t=sigma_t, sigma_r=sigma_r, period=period, pitch=pitch))


def se2_random_walk(n_poses=500, n_inlier=60, n_outlier=15, seed=0, step=1.0,
                    heading_sigma=0.3, sigma_t=0.05, sigma_theta=0.01, lc_radius=2.0,
                    outlier_range=10.0):
    """2D random walk: 1 m forward steps, heading increments N(0, 0.3); odometry
    noise sigma_t=0.05 m (x, y), sigma_theta=0.01 rad; inlier LCs between poses
    within 2 m with |i-j| >= 2; outliers with relative pose uniform in
    [-10,10]^2 x [-pi,pi].  Covariance diag(sigma_t^2, sigma_t^2, sigma_theta^2)."""
    rng = np.random.default_rng(seed)
    th = np.concatenate([[0.0], np.cumsum(rng.normal(0, heading_sigma, n_poses - 1))])
    xy = np.zeros((n_poses, 2))
    for k in range(1, n_poses):
        xy[k] = xy[k - 1] + step * np.array([math.cos(th[k - 1]), math.sin(th[k - 1])])

    def rel(i, j):
        c, s = np.cos(th[i]), np.sin(th[i])
        dx, dy = xy[j, 0] - xy[i, 0], xy[j, 1] - xy[i, 1]
        return c * dx + s * dy, -s * dx + c * dy, th[j] - th[i]

    def pack(x, y, t):
        return np.stack([np.cos(t), np.sin(t), x, y], -1)

    def noisy(i, j):
        x, y, t = rel(i, j)
        m = len(np.atleast_1d(i))
        nx, ny = rng.normal(0, sigma_t, m), rng.normal(0, sigma_t, m)
        nt = rng.normal(0, sigma_theta, m)
        # z = rel * (n)  : compose with a small noise pose on the right
        c, s = np.cos(t), np.sin(t)
        return x + c * nx - s * ny, y + s * nx + c * ny, t + nt

    i0 = np.arange(n_poses - 1)
    ox, oy, ot = noisy(i0, i0 + 1)
    cov = np.diag([sigma_t ** 2, sigma_t ** 2, sigma_theta ** 2]).reshape(9)
    ii, jj = np.triu_indices(n_poses, 2)
    d = np.linalg.norm(xy[ii] - xy[jj], axis=1)
    cand = np.stack([ii[d <= lc_radius], jj[d <= lc_radius]], 1)
    pick = cand[rng.choice(cand.shape[0], n_inlier, replace=False)]
    lx, ly, lt = noisy(pick[:, 0], pick[:, 1])
    oa = rng.integers(0, n_poses, n_outlier)
    ob = rng.integers(0, n_poses, n_outlier)
    bad = np.abs(oa - ob) < 2
    while bad.any():
        ob[bad] = rng.integers(0, n_poses, bad.sum())
        bad = np.abs(oa - ob) < 2
    qx = rng.uniform(-outlier_range, outlier_range, n_outlier)
    qy = rng.uniform(-outlier_range, outlier_range, n_outlier)
    qt = rng.uniform(-math.pi, math.pi, n_outlier)
    frm = np.concatenate([pick[:, 1], oa]).astype(np.int32)  # later -> earlier, as in testPcm
    to = np.concatenate([pick[:, 0], ob]).astype(np.int32)
    # inlier measurement is for (i -> j) = (pick0 -> pick1); invert for (pick1 -> pick0)
    c, s = np.cos(lt), np.sin(lt)
    ix, iy = -(c * lx + s * ly), -(-s * lx + c * ly)
    lp = np.concatenate([pack(ix, iy, -lt), pack(qx, qy, qt)])
    is_out = np.concatenate([np.zeros(n_inlier, bool), np.ones(n_outlier, bool)])
    perm = rng.permutation(n_inlier + n_outlier)
    vals = np.zeros((n_poses, 4)); vals[0] = [1, 0, 0, 0]
    for k in range(n_poses - 1):
        cv, sv, x, y = vals[k]
        cz, sz = math.cos(ot[k]), math.sin(ot[k])
        vals[k + 1] = [cv * cz - sv * sz, sv * cz + cv * sz, x + cv * ox[k] - sv * oy[k],
                       y + sv * ox[k] + cv * oy[k]]
    return PoseGraph(
        dim=2, gt=pack(xy[:, 0], xy[:, 1], th), values=vals, odom=pack(ox, oy, ot),
        odom_cov=np.tile(cov, (n_poses - 1, 1)), lc_from=frm[perm], lc_to=to[perm],
        lc_pose=lp[perm], lc_cov=np.tile(cov, (n_inlier + n_outlier, 1)), is_outlier=is_out[perm],
        meta=dict(kind="se2_random_walk", n_poses=n_poses, seed=seed))


def c4_batch(seed=0, n_graphs=256):
    """The batched workload: n_graphs independent SE(3) helices with pose
    counts log-uniform in [200, 20000], loop closures 0.1-0.3 of the poses and
    outlier fractions uniform in [5 %, 40 %] (graph k seeded seed * 1000 + k)."""
    rng = np.random.default_rng(seed)
    graphs = []
    for k in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        graphs.append(se3_helix(n_poses=n, n_lc=max(1, int(n * rng.uniform(0.1, 0.3))),
                                outlier_frac=float(rng.uniform(0.05, 0.4)), seed=seed * 1000 + k))
    return graphs


# Named workloads with their PCM thresholds.  odom < 0 disables the odometry
# check so that every pair (i<j) is evaluated; lc = 5.0 keeps P(chi^2_6 < 25)
# ~ 0.9997 so the inlier block is nearly complete.
WORKLOADS = {
    "se2_500": dict(fn=se2_random_walk, kw=dict(), odom_thr=1.0, lc_thr=1.0),
    "se3_5k": dict(fn=se3_helix, kw=dict(n_poses=5000, n_lc=1000, outlier_frac=0.2),
                   odom_thr=-1.0, lc_thr=5.0),
    "se3_50k": dict(fn=se3_helix, kw=dict(n_poses=50000, n_lc=10000, outlier_frac=0.5),
                    odom_thr=-1.0, lc_thr=5.0),
}


def workload(name, seed=0):
    w = WORKLOADS[name]
    g = w["fn"](seed=seed, **w["kw"])
    return g, w["odom_thr"], w["lc_thr"]
