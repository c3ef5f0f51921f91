"""Articulated worlds: host-side forward kinematics and coordinate helpers.

An articulated model has one joint per body, in body order, parents first: REVOLUTE (one angle)
or FREE (the floating base: origin position and quaternion xyzw; velocity = COM linear velocity
and angular velocity, world frame), as Newton's ``joint_q`` / ``joint_qd``
(``newton/_src/sim/state.py:149-153``, ``articulation.py`` ``eval_fk``).  A revolute joint's
frame is ``joint_X_p`` in the parent body frame (the world for a root joint) and ``joint_X_c`` in
the child body frame; the child body pose is ``X_parent * X_p * Rot(axis, angle) * X_c^-1``.
"""

from __future__ import annotations

import math

import numpy as np

JOINT_REVOLUTE = 1  # JointType.REVOLUTE (sim/enums.py:116)
JOINT_FREE = 4  # JointType.FREE (sim/enums.py:125)


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _quat(R):
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        return np.array([(R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s])
    if R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        return np.array([0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s, (R[2, 1] - R[1, 2]) / s])
    if R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        return np.array([(R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s, (R[0, 2] - R[2, 0]) / s])
    s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
    return np.array([(R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s, (R[1, 0] - R[0, 1]) / s])


def fk_host(host: dict, joint_q: np.ndarray) -> np.ndarray:
    """Body poses [B, 7] (origin, quaternion xyzw) of an articulated model at ``joint_q``."""
    jt, par = host["joint_type"], host["joint_parent"]
    Xp, Xc, ax, qs = host["joint_X_p"], host["joint_X_c"], host["joint_axis"], host["joint_q_start"]
    B = len(jt)
    Rs, ps = [None] * B, [None] * B
    out = np.zeros((B, 7))
    for b in range(B):
        q = joint_q[qs[b]:qs[b + 1]]
        if jt[b] == JOINT_FREE:
            R, p = _rot(q[3:7]), np.asarray(q[:3], dtype=np.float64)
        else:
            Rj, pj = _rot(Xp[b, 3:]), Xp[b, :3].astype(np.float64)
            if par[b] >= 0:
                pj = ps[par[b]] + Rs[par[b]] @ pj
                Rj = Rs[par[b]] @ Rj
            a = ax[b] / np.linalg.norm(ax[b])
            s, c = math.sin(0.5 * q[0]), math.cos(0.5 * q[0])
            Ra = _rot((a[0] * s, a[1] * s, a[2] * s, c))
            Rc = _rot(Xc[b, 3:])
            R = Rj @ Ra @ Rc.T
            p = pj - R @ Xc[b, :3]
        Rs[b], ps[b] = R, p
        out[b, :3] = p
        out[b, 3:] = _quat(R)
    return out
