"""g2o reader/writer with gtsam::load3D / load2D semantics.

Reads VERTEX_SE3:QUAT / EDGE_SE3:QUAT and VERTEX_SE2 / EDGE_SE2 lines.
  * SE(3) vertex: id x y z qx qy qz qw; the rotation is the quaternion's
    rotation matrix (unit-quaternion formula, no renormalisation).
  * SE(3) edge: i j x y z qx qy qz qw + 21 upper-triangular information
    entries in g2o's (t, R) order; reordered to GTSAM's (R, t) tangent order
    (rotation block first), as the reference's fixtures are read in
    tests/testGnc.cpp:31 and written back by src/Logger.cpp:148-152.
  * SE(2): x y theta, information (x, y, theta) upper triangle (6 entries).
Keys are kept as the file's integers (gtsam::Symbol packing: chr << 56 | index).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def quat_to_rot(qw, qx, qy, qz):
    tx, ty, tz = 2.0 * qx, 2.0 * qy, 2.0 * qz
    twx, twy, twz = tx * qw, ty * qw, tz * qw
    txx, txy, txz = tx * qx, ty * qx, tz * qx
    tyy, tyz, tzz = ty * qy, tz * qy, tz * qz
    return np.array([[1.0 - (tyy + tzz), txy - twz, txz + twy],
                     [txy + twz, 1.0 - (txx + tzz), tyz - twx],
                     [txz - twy, tyz + twx, 1.0 - (txx + tyy)]])


def rot_to_quat(R):
    """Rotation matrix -> (qw, qx, qy, qz) (Shepperd's method)."""
    R = np.asarray(R)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = 0.5 / np.sqrt(tr + 1.0)
        return 0.25 / s, (R[2, 1] - R[1, 2]) * s, (R[0, 2] - R[2, 0]) * s, (R[1, 0] - R[0, 1]) * s
    i = int(np.argmax([R[0, 0], R[1, 1], R[2, 2]]))
    j, k = (i + 1) % 3, (i + 2) % 3
    s = 2.0 * np.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
    q = [0.0, 0.0, 0.0]
    q[i] = 0.25 * s
    q[j] = (R[j, i] + R[i, j]) / s
    q[k] = (R[k, i] + R[i, k]) / s
    w = (R[k, j] - R[j, k]) / s
    return w, q[0], q[1], q[2]


def symbol(chr_, index):
    return (ord(chr_) << 56) | int(index)


def symbol_chr(key):
    return chr((int(key) >> 56) & 0xFF)


def symbol_index(key):
    return int(key) & ((1 << 56) - 1)


@dataclass
class G2oGraph:
    dim: int
    keys: list = field(default_factory=list)        # vertex keys in file order
    poses: dict = field(default_factory=dict)       # key -> pose (12 | 4 doubles)
    edges: list = field(default_factory=list)       # (key_i, key_j, pose, information (d x d))

    def covariance(self, info):
        return np.linalg.inv(info)


def read(path) -> G2oGraph:
    g = None
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            tag = tok[0]
            if tag == "VERTEX_SE3:QUAT":
                g = g or G2oGraph(dim=3)
                k = int(tok[1])
                x, y, z, qx, qy, qz, qw = map(float, tok[2:9])
                g.keys.append(k)
                g.poses[k] = np.concatenate([quat_to_rot(qw, qx, qy, qz).reshape(9), [x, y, z]])
            elif tag == "EDGE_SE3:QUAT":
                g = g or G2oGraph(dim=3)
                i, j = int(tok[1]), int(tok[2])
                x, y, z, qx, qy, qz, qw = map(float, tok[3:10])
                vals = list(map(float, tok[10:31]))
                m = np.zeros((6, 6))
                it = iter(vals)
                for a in range(6):
                    for b in range(a, 6):
                        m[a, b] = next(it)
                m = np.triu(m) + np.triu(m, 1).T
                info = np.zeros((6, 6))
                info[:3, :3] = m[3:, 3:]
                info[3:, 3:] = m[:3, :3]
                info[:3, 3:] = m[3:, :3]
                info[3:, :3] = m[:3, 3:]
                pose = np.concatenate([quat_to_rot(qw, qx, qy, qz).reshape(9), [x, y, z]])
                g.edges.append((i, j, pose, info))
            elif tag == "VERTEX_SE2":
                g = g or G2oGraph(dim=2)
                k = int(tok[1])
                x, y, th = map(float, tok[2:5])
                g.keys.append(k)
                g.poses[k] = np.array([np.cos(th), np.sin(th), x, y])
            elif tag == "EDGE_SE2":
                g = g or G2oGraph(dim=2)
                i, j = int(tok[1]), int(tok[2])
                x, y, th = map(float, tok[3:6])
                vals = list(map(float, tok[6:12]))
                m = np.zeros((3, 3))
                it = iter(vals)
                for a in range(3):
                    for b in range(a, 3):
                        m[a, b] = next(it)
                m = np.triu(m) + np.triu(m, 1).T
                g.edges.append((i, j, np.array([np.cos(th), np.sin(th), x, y]), m))
    if g is None:
        raise ValueError(f"{path}: no SE(2)/SE(3) vertices or edges")
    return g


def write(path, dim, keys, poses, edges):
    """Write a g2o file; information converted back to g2o's (t, R) order."""
    with open(path, "w") as f:
        for k in keys:
            p = poses[k]
            if dim == 3:
                qw, qx, qy, qz = rot_to_quat(np.asarray(p[:9]).reshape(3, 3))
                f.write(f"VERTEX_SE3:QUAT {k} {p[9]} {p[10]} {p[11]} {qx} {qy} {qz} {qw}\n")
            else:
                f.write(f"VERTEX_SE2 {k} {p[2]} {p[3]} {np.arctan2(p[1], p[0])}\n")
        for i, j, p, info in edges:
            if dim == 3:
                qw, qx, qy, qz = rot_to_quat(np.asarray(p[:9]).reshape(3, 3))
                m = np.zeros((6, 6))
                m[:3, :3] = info[3:, 3:]
                m[3:, 3:] = info[:3, :3]
                m[:3, 3:] = info[3:, :3]
                m[3:, :3] = info[:3, 3:]
                tri = " ".join(str(m[a, b]) for a in range(6) for b in range(a, 6))
                f.write(f"EDGE_SE3:QUAT {i} {j} {p[9]} {p[10]} {p[11]} {qx} {qy} {qz} {qw} {tri}\n")
            else:
                tri = " ".join(str(info[a, b]) for a in range(3) for b in range(a, 3))
                f.write(f"EDGE_SE2 {i} {j} {p[2]} {p[3]} {np.arctan2(p[1], p[0])} {tri}\n")
