"""FP64 CPU restatement of Kimera-RPGO's robust solve (GNC-TLS over LM/GN).

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/, smoke() and
bench.py's cpu_baseline leg, never by the product.

Restated (paths relative to the reference tree; GTSAM @ 686e16aa is not
vendored, its published algorithms are restated):
  src/RobustSolver.cpp:115-180, 225-275   GNC wiring: known inliers = the first
                                          n_odom + n_special factors (:105-113),
                                          maxIter / muStep / tolerances (:137-140),
                                          bias_odom initial weights (:144-151),
                                          COST vs PROBABILITY thresholds (:152-163),
                                          gnc_num_inliers truncation (:170-171)
  src/RobustSolver.cpp:181-218, 276-296   plain LM / GN without GNC
  gtsam::GncOptimizer (TLS)               mu init from the residuals at the
      initial values (Remark 5 of the GNC paper, floored at 1e-6, -1 when no
      residual exceeds the threshold), weights
      w = sqrt(barcSq mu (mu+1) / r) - mu clipped to {0,1} outside
      [mu/(mu+1), (mu+1)/mu] barcSq, mu <- mu * muStep, stop on relative cost
      change < tol, binary weights within weightsTol, or maxIterations; every
      inner solve starts from the initial values.
  gtsam::BetweenFactor<Pose3> with SLOW_BUT_CORRECT_BETWEENFACTOR and the
      EXPMAP retraction: e = Log(z^-1 T_i^-1 T_j), J_j = Jr^-1(e),
      J_i = -Jr^-1(e) Ad((T_i^-1 T_j)^-1); retract T Exp(delta).
  gtsam::PriorFactor<Pose3>: e = -Log(x^-1 prior), Jacobian I.
  gtsam::LevenbergMarquardtOptimizer defaults: lambda0 1e-5, factor 10,
      upper bound 1e5, min model fidelity 1e-3, diagonal damping with the
      Hessian diagonal clamped to [1e-6, 1e32], relative/absolute error
      tolerance 1e-5, 100 iterations.
  gtsam::BetweenFactor<Pose2> / PriorFactor<Pose2> (Graph2): Pose2 keeps
      GTSAM's default chart (SLOW_BUT_CORRECT_EXPMAP not defined; the
      reference's build flags set only the Pose3/Rot3 EXPMAP options,
      README.md:52): Local(r) = (r.x, r.y, r.theta) and Retract(v) =
      Pose2(v[0], v[1], v[2]) (Pose2::ChartAtOrigin), so a between factor's
      error is the (x, y, theta) of z^-1 (T_i^-1 T_j), with Jacobians
      H_r * (-Ad(hx^-1)) and H_r, H_r = [[R(r.theta), 0], [0, 1]]
      (Pose2::between, LieGroup::localCoordinates); a prior's error is
      -Local(x^-1 prior) with Jacobian I (PriorFactor at 686e16aa); values
      move by x * Pose2(dx).
Last-bit behaviour (Eigen summation order, CHOLMOD vs SuperLU, Boost chi^2)
is parity unpinned; only converged results are compared.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.stats import chi2


# ------------------------------------------------------------------ SE(3), batched
def skew(w):
    w = np.asarray(w)
    z = np.zeros(w.shape[:-1])
    return np.stack([np.stack([z, -w[..., 2], w[..., 1]], -1),
                     np.stack([w[..., 2], z, -w[..., 0]], -1),
                     np.stack([-w[..., 1], w[..., 0], z], -1)], -2)


def rot_exp(w):
    th2 = np.sum(w * w, -1)[..., None, None]
    th = np.sqrt(th2)
    W = skew(w)
    small = th2 < 1e-20
    ths = np.where(small, 1.0, th)
    A = np.where(small, 1.0 - th2 / 6.0, np.sin(ths) / ths)
    B = np.where(small, 0.5 - th2 / 24.0, (1.0 - np.cos(ths)) / (ths * ths))
    return np.eye(3) + A * W + B * (W @ W)


def rot_log(R):
    """SO3::Logmap (special branch at trace ~ -1, Taylor magnitude near 0)."""
    R = np.asarray(R)
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    out = np.zeros(R.shape[:-2] + (3,))
    ax = np.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0],
                   R[..., 1, 0] - R[..., 0, 1]], -1)
    tr_3 = tr - 3.0
    cosv = np.clip((tr - 1.0) / 2.0, -1.0, 1.0)
    theta = np.arccos(cosv)
    with np.errstate(divide="ignore", invalid="ignore"):
        mag = np.where(tr_3 < -1e-7, theta / (2.0 * np.sin(theta)), 0.5 - tr_3 / 12.0)
    out = mag[..., None] * ax
    sp_ = tr + 1.0 < 1e-10
    if np.any(sp_):
        for idx in zip(*np.nonzero(sp_)) if R.ndim > 2 else [()]:
            Ri = R[idx]
            if abs(Ri[2, 2] + 1.0) > 1e-5:
                v = (math.pi / math.sqrt(2.0 + 2.0 * Ri[2, 2])) * np.array([Ri[0, 2], Ri[1, 2], 1.0 + Ri[2, 2]])
            elif abs(Ri[1, 1] + 1.0) > 1e-5:
                v = (math.pi / math.sqrt(2.0 + 2.0 * Ri[1, 1])) * np.array([Ri[0, 1], 1.0 + Ri[1, 1], Ri[2, 1]])
            else:
                v = (math.pi / math.sqrt(2.0 + 2.0 * Ri[0, 0])) * np.array([1.0 + Ri[0, 0], Ri[1, 0], Ri[2, 0]])
            out[idx] = v
    return out


def se3_log(R, t):
    w = rot_log(R)
    th = np.linalg.norm(w, axis=-1)
    small = th < 1e-10
    ths = np.where(small, 1.0, th)
    n = w / ths[..., None]
    WT = np.cross(n, t)
    WWT = np.cross(n, WT)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(small, 0.0, 1.0 - ths / (2.0 * np.tan(0.5 * ths)))
    u = t - (0.5 * ths)[..., None] * WT + c[..., None] * WWT
    u = np.where(small[..., None], t, u)
    return np.concatenate([w, u], -1)


def se3_exp(xi):
    w, v = xi[..., :3], xi[..., 3:]
    R = rot_exp(w)
    th2 = np.sum(w * w, -1)
    big = th2 > 2.220446049250313e-16
    wxv = np.cross(w, v)
    Rwxv = (R @ wxv[..., None])[..., 0]
    par = w * np.sum(w * v, -1)[..., None]
    th2s = np.where(big, th2, 1.0)
    t = np.where(big[..., None], (wxv - Rwxv + par) / th2s[..., None], v)
    return R, t


def so3_logmap_derivative(w):
    th2 = np.sum(w * w, -1)
    th = np.sqrt(th2)
    W = skew(w)
    small = th2 <= 2.220446049250313e-16
    ths = np.where(small, 1.0, th)
    coef = np.where(small, 0.0, 1.0 / (ths * ths) - (1.0 + np.cos(ths)) / (2.0 * ths * np.sin(ths)))
    J = np.eye(3) + 0.5 * W + coef[..., None, None] * (W @ W)
    return np.where(small[..., None, None], np.eye(3), J)


def q_expmap_derivative(xi, near_zero=1e-5):
    """Barfoot14tro eq. (102), sign convention of Pose3::ComputeQforExpmapDerivative."""
    w, v = xi[..., :3], xi[..., 3:]
    V, W = skew(v), skew(w)
    WVW = W @ V @ W
    phi = np.linalg.norm(w, axis=-1)
    big = np.abs(phi) > near_zero
    p = np.where(big, phi, 1.0)
    s, c = np.sin(p), np.cos(p)
    p2, p3 = p * p, p * p * p
    p4, p5 = p3 * p, p3 * p * p
    a = np.where(big, (p - s) / p3, 1.0 / 6.0)
    b = np.where(big, (1 - p2 / 2 - c) / p4, -1.0 / 24.0)
    cc = np.where(big, -0.5 * ((1 - p2 / 2 - c) / p4 - 3 * (p - s - p3 / 6.0) / p5), 1.0 / 120.0)
    A = a[..., None, None]; B = b[..., None, None]; C = cc[..., None, None]
    return (-0.5 * V + A * (W @ V + V @ W - WVW) + B * (W @ W @ V + V @ W @ W - 3 * WVW)
            + C * (WVW @ W + W @ WVW))


def se3_logmap_derivative(xi):
    w = xi[..., :3]
    Jw = so3_logmap_derivative(w)
    Q = q_expmap_derivative(xi)
    Q2 = -Jw @ Q @ Jw
    z = np.zeros_like(Jw)
    top = np.concatenate([Jw, z], -1)
    bot = np.concatenate([Q2, Jw], -1)
    return np.concatenate([top, bot], -2)


def adjoint(R, t):
    A = skew(t) @ R
    z = np.zeros_like(R)
    return np.concatenate([np.concatenate([R, z], -1), np.concatenate([A, R], -1)], -2)


# ------------------------------------------------------------------ graph
@dataclass
class Graph:
    """Pose graph: N SE(3) poses; between factors (i -> j, z, information);
    priors (k, prior, information).  Factor order = between factors then
    priors ... as given by `order` (list of ("b", idx) / ("p", idx))."""
    R: np.ndarray                 # (N,3,3) initial values
    t: np.ndarray                 # (N,3)
    bi: np.ndarray                # (F,) int
    bj: np.ndarray
    zR: np.ndarray                # (F,3,3)
    zt: np.ndarray                # (F,3)
    binfo: np.ndarray             # (F,6,6) information (omega, v)
    pk: np.ndarray = field(default_factory=lambda: np.zeros(0, int))
    pR: np.ndarray = field(default_factory=lambda: np.zeros((0, 3, 3)))
    pt: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    pinfo: np.ndarray = field(default_factory=lambda: np.zeros((0, 6, 6)))
    order: list = None            # factor order; default: between then priors

    def __post_init__(self):
        if self.order is None:
            self.order = [("b", k) for k in range(len(self.bi))] + [("p", k) for k in range(len(self.pk))]
        self.F = len(self.order)
        self.N = self.R.shape[0]
        self._bpos = np.array([q for q, (kind, _) in enumerate(self.order) if kind == "b"], int)
        self._ppos = np.array([q for q, (kind, _) in enumerate(self.order) if kind == "p"], int)
        self._bidx = np.array([k for kind, k in self.order if kind == "b"], int)
        self._pidx = np.array([k for kind, k in self.order if kind == "p"], int)


def graph_from_arrays(poses, kind, key_i, key_j, measured, information):
    """Graph from the flat factor list of the rpgo_solve C ABI (factor order
    kept): kind 0 = BetweenFactor<Pose3>(key_i, key_j, measured), kind 1 =
    PriorFactor<Pose3>(key_i, measured); poses n x 12 (R row-major, t);
    information F x 36 in (omega, v) order."""
    poses = np.asarray(poses, float).reshape(-1, 12)
    kind = np.asarray(kind)
    meas = np.asarray(measured, float).reshape(-1, 12)
    info = np.asarray(information, float).reshape(-1, 6, 6)
    b = np.nonzero(kind == 0)[0]
    p = np.nonzero(kind == 1)[0]
    bpos = {q: n for n, q in enumerate(b)}
    ppos = {q: n for n, q in enumerate(p)}
    order = [("b", bpos[q]) if kind[q] == 0 else ("p", ppos[q]) for q in range(len(kind))]
    ki, kj = np.asarray(key_i, int), np.asarray(key_j, int)
    return Graph(R=poses[:, :9].reshape(-1, 3, 3).copy(), t=poses[:, 9:].copy(), bi=ki[b], bj=kj[b],
                 zR=meas[b, :9].reshape(-1, 3, 3), zt=meas[b, 9:], binfo=info[b], pk=ki[p],
                 pR=meas[p, :9].reshape(-1, 3, 3), pt=meas[p, 9:], pinfo=info[p], order=order)


# ------------------------------------------------------------------ SE(2), batched
def se2_compose(ta, xa, tb, xb):
    """(theta, xy) of a * b; theta wrapped by atan2 as Rot2 keeps (c, s)."""
    ca, sa = np.cos(ta), np.sin(ta)
    x = xa[..., 0] + ca * xb[..., 0] - sa * xb[..., 1]
    y = xa[..., 1] + sa * xb[..., 0] + ca * xb[..., 1]
    cb, sb = np.cos(tb), np.sin(tb)
    th = np.arctan2(sa * cb + ca * sb, ca * cb - sa * sb)
    return th, np.stack([x, y], -1)


def se2_between(ta, xa, tb, xb):
    """Pose2::between (a^-1 b) and its first Jacobian -Ad(between(b, a))."""
    c1, s1, c2, s2 = np.cos(ta), np.sin(ta), np.cos(tb), np.sin(tb)
    c = c1 * c2 + s1 * s2
    s = -s1 * c2 + c1 * s2
    dx = xb[..., 0] - xa[..., 0]
    dy = xb[..., 1] - xa[..., 1]
    th = np.arctan2(s, c)
    t = np.stack([c1 * dx + s1 * dy, -s1 * dx + c1 * dy], -1)
    z = np.zeros_like(c)
    H1 = np.stack([np.stack([-c, -s, -s2 * dx + c2 * dy], -1),
                   np.stack([s, -c, -c2 * dx - s2 * dy], -1),
                   np.stack([z, z, z - 1.0], -1)], -2)
    return th, t, H1


def _chart_H(th):
    """Jacobian of Pose2::ChartAtOrigin::Local at r: [[R(r.theta), 0], [0, 1]]."""
    c, s = np.cos(th), np.sin(th)
    z, o = np.zeros_like(c), np.ones_like(c)
    return np.stack([np.stack([c, -s, z], -1), np.stack([s, c, z], -1), np.stack([z, z, o], -1)], -2)


@dataclass
class Graph2:
    """Pose2 graph: N poses (theta, xy) initial values; between factors
    (i -> j, z = (theta, xy), information 3x3 in (x, y, theta) order); priors;
    factor order as in Graph."""
    th: np.ndarray
    xy: np.ndarray
    bi: np.ndarray
    bj: np.ndarray
    zth: np.ndarray
    zxy: np.ndarray
    binfo: np.ndarray
    pk: np.ndarray = field(default_factory=lambda: np.zeros(0, int))
    pth: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pxy: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    pinfo: np.ndarray = field(default_factory=lambda: np.zeros((0, 3, 3)))
    order: list = None

    __post_init__ = Graph.__post_init__

    @property
    def R(self):  # the generic loops carry (R, t) = (theta, xy)
        return self.th

    @property
    def t(self):
        return self.xy


def _errors2(g: Graph2, th, xy, jac=False):
    """Per-factor (x, y, theta) errors in factor order, optionally the Jacobians."""
    F = g.F
    E = np.zeros((F, 3))
    Ji = np.zeros((F, 3, 3))
    Jj = np.zeros((F, 3, 3))
    if len(g._bidx):
        k = g._bidx
        I, J = g.bi[k], g.bj[k]
        hth, ht, Hb1 = se2_between(th[I], xy[I], th[J], xy[J])
        rth, rt, _ = se2_between(g.zth[k], g.zxy[k], hth, ht)
        E[g._bpos] = np.stack([rt[:, 0], rt[:, 1], rth], -1)
        if jac:
            Hl = _chart_H(rth)
            Ji[g._bpos] = Hl @ Hb1
            Jj[g._bpos] = Hl
    if len(g._pidx):
        k = g._pidx
        X = g.pk[k]
        rth, rt, _ = se2_between(th[X], xy[X], g.pth[k], g.pxy[k])
        E[g._ppos] = -np.stack([rt[:, 0], rt[:, 1], rth], -1)
        Ji[g._ppos] = np.eye(3)
    return E, Ji, Jj


def _info2(g: Graph2):
    L = np.zeros((g.F, 3, 3))
    if len(g._bidx):
        L[g._bpos] = g.binfo[g._bidx]
    if len(g._pidx):
        L[g._ppos] = g.pinfo[g._pidx]
    return L


def factor_errors(g: Graph, R, t):
    """Whitened error e and 0.5 e^T Lambda e per factor, in factor order."""
    if isinstance(g, Graph2):
        E, _, _ = _errors2(g, R, t)
        return 0.5 * np.einsum("fi,fij,fj->f", E, _info2(g), E)
    dims = np.zeros(g.F)
    err = np.zeros(g.F)
    if len(g._bidx):
        k = g._bidx
        Ri, ti, Rj, tj = R[g.bi[k]], t[g.bi[k]], R[g.bj[k]], t[g.bj[k]]
        RiT = np.swapaxes(Ri, -1, -2)
        Rr = RiT @ Rj
        tr = (RiT @ (tj - ti)[..., None])[..., 0]
        zRT = np.swapaxes(g.zR[k], -1, -2)
        RE = zRT @ Rr
        tE = (zRT @ (tr - g.zt[k])[..., None])[..., 0]
        e = se3_log(RE, tE)
        err[g._bpos] = 0.5 * np.einsum("fi,fij,fj->f", e, g.binfo[k], e)
    if len(g._pidx):
        k = g._pidx
        Rx, tx = R[g.pk[k]], t[g.pk[k]]
        RxT = np.swapaxes(Rx, -1, -2)
        e = -se3_log(RxT @ g.pR[k], (RxT @ (g.pt[k] - tx)[..., None])[..., 0])
        err[g._ppos] = 0.5 * np.einsum("fi,fij,fj->f", e, g.pinfo[k], e)
    return err


def _linearize2(g: Graph2, th, xy, w):
    N = g.N
    E, Ji, Jj = _errors2(g, th, xy, jac=True)
    L = w[:, None, None] * _info2(g)
    I = np.zeros(g.F, int)
    J = np.zeros(g.F, int)
    if len(g._bidx):
        I[g._bpos], J[g._bpos] = g.bi[g._bidx], g.bj[g._bidx]
    if len(g._pidx):
        I[g._ppos] = g.pk[g._pidx]
    isb = np.zeros(g.F, bool)
    isb[g._bpos] = True
    JiT, JjT = np.swapaxes(Ji, -1, -2), np.swapaxes(Jj, -1, -2)
    rows, cols, vals = [], [], []

    def add_block(P, Q, M):
        ii = (3 * P)[:, None, None] + np.arange(3)[None, :, None]
        jj = (3 * Q)[:, None, None] + np.arange(3)[None, None, :]
        rows.append(np.broadcast_to(ii, M.shape).ravel())
        cols.append(np.broadcast_to(jj, M.shape).ravel())
        vals.append(M.ravel())

    add_block(I, I, JiT @ L @ Ji)
    b = np.zeros(3 * N)
    Le = (L @ E[..., None])[..., 0]
    np.add.at(b, (3 * I)[:, None] + np.arange(3), -(JiT @ Le[..., None])[..., 0])
    if isb.any():
        Bi, Bj = I[isb], J[isb]
        add_block(Bj, Bj, (JjT @ L @ Jj)[isb])
        add_block(Bi, Bj, (JiT @ L @ Jj)[isb])
        add_block(Bj, Bi, (JjT @ L @ Ji)[isb])
        np.add.at(b, (3 * Bj)[:, None] + np.arange(3), -(JjT @ Le[..., None])[..., 0][isb])
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * N, 3 * N))
    return H, b


def linearize(g: Graph, R, t, w):
    """Weighted Gauss-Newton system H (scipy CSR, 6N), gradient b = -J^T W e."""
    if isinstance(g, Graph2):
        return _linearize2(g, R, t, w)
    N = g.N
    rows, cols, vals = [], [], []
    b = np.zeros(6 * N)

    def add_block(I, J, M):
        ii = (6 * I)[:, None, None] + np.arange(6)[None, :, None]
        jj = (6 * J)[:, None, None] + np.arange(6)[None, None, :]
        rows.append(np.broadcast_to(ii, M.shape).ravel())
        cols.append(np.broadcast_to(jj, M.shape).ravel())
        vals.append(M.ravel())

    if len(g._bidx):
        k = g._bidx
        wk = w[g._bpos]
        I, J = g.bi[k], g.bj[k]
        Ri, ti, Rj, tj = R[I], t[I], R[J], t[J]
        RiT = np.swapaxes(Ri, -1, -2)
        Rr = RiT @ Rj
        tr = (RiT @ (tj - ti)[..., None])[..., 0]
        zRT = np.swapaxes(g.zR[k], -1, -2)
        RE = zRT @ Rr
        tE = (zRT @ (tr - g.zt[k])[..., None])[..., 0]
        e = se3_log(RE, tE)
        Jr = se3_logmap_derivative(e)
        RrT = np.swapaxes(Rr, -1, -2)
        Ad_inv = adjoint(RrT, -(RrT @ tr[..., None])[..., 0])
        Ji = -Jr @ Ad_inv
        Jj = Jr
        L = wk[:, None, None] * g.binfo[k]
        JiT, JjT = np.swapaxes(Ji, -1, -2), np.swapaxes(Jj, -1, -2)
        add_block(I, I, JiT @ L @ Ji)
        add_block(J, J, JjT @ L @ Jj)
        add_block(I, J, JiT @ L @ Jj)
        add_block(J, I, JjT @ L @ Ji)
        Le = (L @ e[..., None])[..., 0]
        np.add.at(b, (6 * I)[:, None] + np.arange(6), -(JiT @ Le[..., None])[..., 0])
        np.add.at(b, (6 * J)[:, None] + np.arange(6), -(JjT @ Le[..., None])[..., 0])
    if len(g._pidx):
        k = g._pidx
        wk = w[g._ppos]
        X = g.pk[k]
        Rx, tx = R[X], t[X]
        RxT = np.swapaxes(Rx, -1, -2)
        e = -se3_log(RxT @ g.pR[k], (RxT @ (g.pt[k] - tx)[..., None])[..., 0])
        L = wk[:, None, None] * g.pinfo[k]
        add_block(X, X, L)
        np.add.at(b, (6 * X)[:, None] + np.arange(6), -(L @ e[..., None])[..., 0])
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(6 * N, 6 * N))
    return H, b


def _retract(g, R, t, dx):
    """values.retract(dx): Expmap for Pose3, x * Pose2(dx) for Pose2."""
    if isinstance(g, Graph2):
        d = dx.reshape(-1, 3)
        return se2_compose(R, t, d[:, 2], d[:, :2])
    return retract(R, t, dx)


def retract(R, t, dx):
    dR, dt = se3_exp(dx.reshape(-1, 6))
    Rn = R @ dR
    tn = t + (R @ dt[..., None])[..., 0]
    return Rn, tn


@dataclass
class LMParams:
    lambda_initial: float = 1e-5
    lambda_factor: float = 10.0
    lambda_upper: float = 1e5
    lambda_lower: float = 0.0
    min_model_fidelity: float = 1e-3
    diagonal_damping: bool = True
    min_diagonal: float = 1e-6
    max_diagonal: float = 1e32
    max_iterations: int = 100
    relative_error_tol: float = 1e-5
    absolute_error_tol: float = 1e-5


def _solve(H, b):
    return spla.spsolve(H.tocsc(), b)


def lm_optimize(g: Graph, R0, t0, w, p: LMParams = LMParams(), gauss_newton=False, trace=None):
    """LevenbergMarquardtOptimizer / GaussNewtonOptimizer on the weighted graph."""
    R, t = R0.copy(), t0.copy()
    err = float(np.dot(w, factor_errors(g, R, t)))
    lam = p.lambda_initial
    it = 0
    while it < p.max_iterations:
        H, b = linearize(g, R, t, w)
        if gauss_newton:
            dx = _solve(H, b)
            Rn, tn = _retract(g, R, t, dx)
            new_err = float(np.dot(w, factor_errors(g, Rn, tn)))
            R, t = Rn, tn
            it += 1
            conv = _converged(p, err, new_err)
            err = new_err
            if trace is not None:
                trace.append(err)
            if conv:
                break
            continue
        diag = H.diagonal()
        if p.diagonal_damping:
            D = np.clip(diag, p.min_diagonal, p.max_diagonal)
        else:
            D = np.ones_like(diag)
        accepted = False
        while True:
            Hd = H + sp.diags(lam * D)
            dx = _solve(Hd, b)
            # linearized cost change: b^T dx - 0.5 dx^T H dx  (b = -g)
            lin = float(b @ dx - 0.5 * dx @ (H @ dx))
            Rn, tn = _retract(g, R, t, dx)
            new_err = float(np.dot(w, factor_errors(g, Rn, tn)))
            nonlin = err - new_err
            fid = nonlin / lin if lin > 0 else -1.0
            if lin >= 0 and fid >= p.min_model_fidelity:
                accepted = True
                lam = max(lam / p.lambda_factor, p.lambda_lower)
                break
            lam *= p.lambda_factor
            if lam >= p.lambda_upper:
                break
        it += 1
        if not accepted:
            break
        conv = _converged(p, err, new_err)
        R, t, err = Rn, tn, new_err
        if trace is not None:
            trace.append(err)
        if conv:
            break
    return R, t, err, it


def _converged(p, cur, new):
    if new <= 0.0:
        return True
    dec = cur - new
    return (dec / cur <= p.relative_error_tol) or (dec <= p.absolute_error_tol)


@dataclass
class GncParams:
    threshold_mode: str = "probability"   # or "cost"
    inlier_threshold: float = 0.9
    max_iterations: int = 100
    mu_step: float = 1.4
    relative_cost_tol: float = 1e-5
    weights_tol: float = 1e-4
    known_inliers: np.ndarray = None
    init_weights: np.ndarray = None       # GncOptimizer::setWeights
    init_mu: float = None                 # learned-init extension (not in GTSAM)


def barc_sq(g: Graph, p: GncParams):
    dims = np.full(g.F, 3.0 if isinstance(g, Graph2) else 6.0)
    if p.threshold_mode == "probability":
        return 0.5 * chi2.ppf(p.inlier_threshold, dims)
    return np.full(g.F, p.inlier_threshold)


def tls_mu_init(r0, barc):
    """GncOptimizer::initializeMu for TLS (Remark 5 of the GNC paper, per-factor
    thresholds): the smallest barc_k / (2 r_k - barc_k) over factors with
    r_k > barc_k / 2 at the initial values, floored at 1e-6; -1 when no factor
    qualifies (GNC then stops before its first iteration)."""
    mu = math.inf
    for k in range(len(r0)):
        if 2 * r0[k] - barc[k] > 0:
            mu = min(mu, barc[k] / (2 * r0[k] - barc[k]))
    if 0 <= mu < 1e-6:
        mu = 1e-6
    return mu if (mu > 0 and not math.isinf(mu)) else -1.0


def tls_weights(r, mu, barc, known):
    """GncOptimizer::calculateWeights for TLS: w = sqrt(barc mu (mu+1) / r) - mu
    inside the band (mu/(mu+1) barc, (mu+1)/mu barc), 1 at or below it, 0 at or
    above it; known inliers keep weight 1."""
    F = len(r)
    w = np.ones(F)
    ub = (mu + 1) / mu * barc
    lb = mu / (mu + 1) * barc
    with np.errstate(divide="ignore", invalid="ignore"):
        wn = np.sqrt(barc * mu * (mu + 1) / r) - mu
    wn = np.where((r >= ub) | (wn < 0), 0.0, np.where((r <= lb) | (wn > 1), 1.0, wn))
    unk = ~known
    w[unk] = wn[unk]
    return w


def gnc_converged(cost, prev_cost, w, p: GncParams):
    """GncOptimizer::checkConvergence for TLS: relative cost change below
    relativeCostTol, or every weight within weightsTol of 0 or 1 (the mu test
    applies to GM only)."""
    cost_conv = abs(cost - prev_cost) / max(prev_cost, 1e-7) < p.relative_cost_tol
    w_conv = bool(np.all(np.abs(w - np.round(w)) <= p.weights_tol))
    return cost_conv or w_conv


def gnc_optimize(g: Graph, p: GncParams, lm: LMParams = LMParams(), gauss_newton=False,
                 polish=None):
    """Returns dict(R, t, weights, iterations, mu_trace, cost)."""
    F = g.F
    known = np.zeros(F, bool)
    if p.known_inliers is not None:
        known[np.asarray(p.known_inliers, int)] = True
    w = np.ones(F) if p.init_weights is None else np.asarray(p.init_weights, float).copy()
    barc = barc_sq(g, p)
    R0, t0 = g.R, g.t
    R, t, _, _ = lm_optimize(g, R0, t0, w, lm, gauss_newton)
    if p.init_mu is not None:
        mu = float(p.init_mu)
    else:
        mu = tls_mu_init(factor_errors(g, R0, t0), barc)
    prev_cost = float(np.dot(w, factor_errors(g, R, t)))
    n_unknown = F - int(known.sum())
    mu_trace = [mu]
    iters = 0
    if not (mu <= 0 or n_unknown == 0):
        for it in range(p.max_iterations):
            w = tls_weights(factor_errors(g, R, t), mu, barc, known)
            R, t, _, _ = lm_optimize(g, R0, t0, w, lm, gauss_newton)
            cost = float(np.dot(w, factor_errors(g, R, t)))
            iters = it + 1
            if gnc_converged(cost, prev_cost, w, p):
                break
            mu = mu * p.mu_step
            mu_trace.append(mu)
            prev_cost = cost
    if polish is not None:
        R, t, _, _ = lm_optimize(g, R, t, w, polish, gauss_newton)
    return dict(R=R, t=t, weights=w, iterations=iters, mu_trace=mu_trace,
                cost=float(np.dot(w, factor_errors(g, R, t))))


def least_squares(g: Graph, w, R0=None, t0=None, tol=1e-12, max_it=50):
    """Converged weighted least squares from (R0, t0) by Gauss-Newton (polish)."""
    p = LMParams(relative_error_tol=tol, absolute_error_tol=0.0, max_iterations=max_it)
    R0 = g.R if R0 is None else R0
    t0 = g.t if t0 is None else t0
    return lm_optimize(g, R0, t0, w, p)
