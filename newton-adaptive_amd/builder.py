"""Scene builder with the reference's ``ModelBuilder`` surface (``newton/_src/sim/builder.py:88``).

Only the subset the hot path needs is provided: free rigid bodies carrying sphere /
box / capsule shapes, static planes and boxes (world ``-1``), per-world replication,
and the shape material config.  Mass properties follow ``geometry/inertia.py``
(solid sphere ``:44``, capsule ``:84``, box ``:228``) and the broad-phase bounding radius
follows ``geometry/utils.py:73`` (an infinite plane reports ``1e6``).
"""

from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

from .model import GeoType, Model, SUPPORTED_SHAPES


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw,
            aw * bw - ax * bx - ay * by - az * bz)


def quat_to_matrix(q) -> np.ndarray:
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def quat_rotate(q, v):
    return quat_to_matrix(q) @ np.asarray(v, dtype=np.float64)


def transform_multiply(a, b):
    pa, qa = a[:3], a[3:]
    pb, qb = b[:3], b[3:]
    p = np.asarray(pa) + quat_rotate(qa, pb)
    return (*p, *quat_mul(qa, qb))


def quat_from_axis_angle(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(0.5 * angle)
    return (axis[0] * s, axis[1] * s, axis[2] * s, math.cos(0.5 * angle))


IDENTITY = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)


class ModelBuilder:
    """Build a :class:`Model` (``builder.py:88``)."""

    @dataclass
    class ShapeConfig:
        """Shape material / collision settings (``builder.py:231-330``)."""

        density: float = 1000.0
        ke: float = 2.5e3
        kd: float = 100.0
        kf: float = 1000.0
        mu: float = 1.0
        margin: float = 0.0
        gap: float | None = None
        collision_group: int = 1
        has_shape_collision: bool = True
        is_visible: bool = True

        def copy(self):
            return copy.copy(self)

    def __init__(self, up_axis: int = 2, gravity: float = -9.81):
        self.up_axis = up_axis
        self.up_vector = tuple(1.0 if i == up_axis else 0.0 for i in range(3))
        self.gravity = gravity
        self.rigid_gap = 0.1  # builder.py:1183
        self.default_shape_cfg = ModelBuilder.ShapeConfig()
        self.current_world = -1
        self._num_worlds = 0
        # bodies
        self.body_q: list = []
        self.body_qd: list = []
        self.body_world: list = []
        self.body_has_joint: list = []
        self.body_mass_override: list = []
        self.body_label: list = []
        # shapes
        self.shape_type: list = []
        self.shape_body: list = []
        self.shape_transform: list = []
        self.shape_scale: list = []
        self.shape_cfg: list = []
        self.shape_world: list = []
        # explicit joints (articulations); a model with none is all free bodies
        self.joints: list = []
        self._joint_children: set = set()

    # ------------------------------------------------------------------ counts
    @property
    def body_count(self) -> int:
        return len(self.body_q)

    @property
    def shape_count(self) -> int:
        return len(self.shape_type)

    @property
    def joint_q(self) -> list:
        out = []
        for q in self.body_q:
            out.extend(q)
        return out

    @joint_q.setter
    def joint_q(self, values):
        values = list(values)
        for b in range(self.body_count):
            self.body_q[b] = tuple(float(v) for v in values[7 * b:7 * b + 7])

    @property
    def joint_qd(self) -> list:
        out = []
        for qd in self.body_qd:
            out.extend(qd)
        return out

    @joint_qd.setter
    def joint_qd(self, values):
        values = list(values)
        for b in range(self.body_count):
            self.body_qd[b] = tuple(float(v) for v in values[6 * b:6 * b + 6])

    # ------------------------------------------------------------------ worlds
    def begin_world(self) -> int:
        self.current_world = self._num_worlds
        self._num_worlds += 1
        return self.current_world

    def end_world(self) -> None:
        self.current_world = -1

    # ------------------------------------------------------------------ bodies
    def add_link(self, xform=IDENTITY, mass: float | None = None, label: str | None = None) -> int:
        """Add a rigid body without a joint (``builder.py`` ``add_link``)."""
        self.body_q.append(tuple(float(v) for v in xform))
        self.body_qd.append((0.0,) * 6)
        self.body_world.append(self.current_world)
        self.body_has_joint.append(False)
        self.body_mass_override.append(mass)
        self.body_label.append(label)
        return self.body_count - 1

    def add_body(self, xform=IDENTITY, mass: float | None = None, label: str | None = None) -> int:
        """Add a rigid body with a FREE joint to the world (``builder.py:4105``)."""
        b = self.add_link(xform, mass, label)
        self.body_has_joint[b] = True
        if self.joints:
            self._add_joint(4, -1, b, IDENTITY, IDENTITY, (0.0, 0.0, 1.0), 0.0, 0.0, 0.0, 0.0, 0.0)
        return b

    def add_joint_free(self, child: int, parent: int = -1, **_kw) -> int:
        """Attach ``child`` to the world with a FREE joint (``builder.py:4696``).  In an
        articulation this is the floating base; its coordinates start at the body's xform."""
        if parent != -1:
            raise ValueError("only world-parented FREE joints are supported on this path")
        self.body_has_joint[child] = True
        if self.joints:
            return self._add_joint(4, -1, child, IDENTITY, IDENTITY, (0.0, 0.0, 1.0), 0.0, 0.0, 0.0, 0.0, 0.0)
        return child

    def add_joint_revolute(self, parent: int, child: int, parent_xform=IDENTITY, child_xform=IDENTITY,
                           axis=(0.0, 0.0, 1.0), target_pos: float = 0.0, target_ke: float = 0.0,
                           target_kd: float = 0.0, armature: float = 0.0, q0: float = 0.0, **_kw) -> int:
        """Revolute joint (``builder.py:4387``): ``child`` turns about ``axis`` (joint frame) by the
        joint angle; the joint frame is ``parent_xform`` in the parent body frame (the world for
        ``parent = -1``) and ``child_xform`` in the child body frame.  A PD drive pulls the angle
        to ``target_pos`` with stiffness ``target_ke`` and damping ``target_kd``; ``armature``
        adds to the joint-space inertia.  ``q0`` is the initial angle."""
        if self.joints == [] and any(self.body_has_joint):
            # the free bodies built so far become explicit FREE joints
            for b in range(self.body_count):
                if self.body_has_joint[b]:
                    self._add_joint(4, -1, b, IDENTITY, IDENTITY, (0.0, 0.0, 1.0), 0.0, 0.0, 0.0, 0.0, 0.0)
        self.body_has_joint[child] = True
        return self._add_joint(1, parent, child, parent_xform, child_xform, axis, target_pos, target_ke, target_kd,
                               armature, q0)

    def _add_joint(self, jtype, parent, child, xp, xc, axis, tpos, tke, tkd, arm, q0) -> int:
        if parent >= child:
            raise ValueError("a joint's parent body must be added before its child")
        if child in self._joint_children:
            raise ValueError(f"body {child} already has a joint")
        self._joint_children.add(int(child))
        self.joints.append(dict(type=int(jtype), parent=int(parent), child=int(child),
                                X_p=tuple(float(v) for v in xp), X_c=tuple(float(v) for v in xc),
                                axis=tuple(float(v) for v in axis), target_pos=float(tpos), target_ke=float(tke),
                                target_kd=float(tkd), armature=float(arm), q0=float(q0)))
        return len(self.joints) - 1

    # ------------------------------------------------------------------ shapes
    def _add_shape(self, body, gtype, xform, scale, cfg):
        cfg = (cfg or self.default_shape_cfg).copy()
        if gtype not in SUPPORTED_SHAPES:
            raise ValueError(f"shape type {gtype} is not on the supported primitive path")
        self.shape_type.append(int(gtype))
        self.shape_body.append(int(body))
        self.shape_transform.append(tuple(float(v) for v in xform))
        self.shape_scale.append(tuple(float(v) for v in scale))
        self.shape_cfg.append(cfg)
        world = self.body_world[body] if body >= 0 else self.current_world
        self.shape_world.append(world)
        return self.shape_count - 1

    def add_shape_sphere(self, body: int, xform=IDENTITY, radius: float = 1.0, cfg=None, **_kw) -> int:
        return self._add_shape(body, GeoType.SPHERE, xform, (radius, 0.0, 0.0), cfg)

    def add_shape_box(self, body: int, xform=IDENTITY, hx: float = 0.5, hy: float = 0.5, hz: float = 0.5,
                      cfg=None, **_kw) -> int:
        return self._add_shape(body, GeoType.BOX, xform, (hx, hy, hz), cfg)

    def add_shape_capsule(self, body: int, xform=IDENTITY, radius: float = 1.0, half_height: float = 0.5,
                          cfg=None, **_kw) -> int:
        """Capsule along the shape's local z axis (``builder.py:6537``)."""
        return self._add_shape(body, GeoType.CAPSULE, xform, (radius, half_height, 0.0), cfg)

    def add_shape_plane(self, plane=(0.0, 0.0, 1.0, 0.0), xform=None, width: float = 0.0, length: float = 0.0,
                        body: int = -1, cfg=None, **_kw) -> int:
        """Plane ``n . x + d = 0``; normal is the shape's local +z; width=length=0 is infinite."""
        if xform is None:
            n = np.asarray(plane[:3], dtype=np.float64)
            n = n / np.linalg.norm(n)
            p = -float(plane[3]) * n
            z = np.array([0.0, 0.0, 1.0])
            c = float(np.dot(z, n))
            if c > 1.0 - 1e-12:
                q = (0.0, 0.0, 0.0, 1.0)
            elif c < -1.0 + 1e-12:
                q = (1.0, 0.0, 0.0, 0.0)
            else:
                axis = np.cross(z, n)
                q = quat_from_axis_angle(axis, math.acos(c))
            xform = (*p, *q)
        return self._add_shape(body, GeoType.PLANE, xform, (width, length, 0.0), cfg)

    def add_ground_plane(self, height: float = 0.0, cfg=None, **_kw) -> int:
        """Infinite ground plane along the up axis (``builder.py:6342``)."""
        return self.add_shape_plane(plane=(*self.up_vector, -height), width=0.0, length=0.0, cfg=cfg)

    # ------------------------------------------------------------------ replication
    def add_builder(self, other: "ModelBuilder", xform=None, world: int | None = None) -> None:
        """Append ``other``'s bodies and shapes, assigning them to ``world``."""
        if world is None:
            world = self.begin_world()
            close = True
        else:
            close = False
        off = self.body_count
        for b in range(other.body_count):
            q = other.body_q[b]
            if xform is not None:
                q = transform_multiply(xform, q)
            self.body_q.append(tuple(float(v) for v in q))
            self.body_qd.append(other.body_qd[b])
            self.body_world.append(world)
            self.body_has_joint.append(other.body_has_joint[b])
            self.body_mass_override.append(other.body_mass_override[b])
            self.body_label.append(other.body_label[b])
        for s in range(other.shape_count):
            ob = other.shape_body[s]
            xf = other.shape_transform[s]
            if ob < 0 and xform is not None:
                xf = transform_multiply(xform, xf)
            self.shape_type.append(other.shape_type[s])
            self.shape_body.append(ob + off if ob >= 0 else -1)
            self.shape_transform.append(tuple(xf))
            self.shape_scale.append(other.shape_scale[s])
            self.shape_cfg.append(other.shape_cfg[s].copy())
            self.shape_world.append(world)
        if close:
            self.end_world()

    def replicate(self, other: "ModelBuilder", world_count: int, spacing=(0.0, 0.0, 0.0)) -> None:
        """``world_count`` copies of ``other``, one world each (``builder.py:2513``)."""
        for i in range(world_count):
            off = np.asarray(spacing, dtype=np.float64) * i
            xf = (*off, 0.0, 0.0, 0.0, 1.0) if np.any(off) else None
            self.add_builder(other, xform=xf)

    # ------------------------------------------------------------------ finalize
    @staticmethod
    def _shape_mass(gtype, scale, density):
        if gtype == GeoType.SPHERE:
            r = scale[0]
            m = density * 4.0 / 3.0 * math.pi * r ** 3
            ia = 0.4 * m * r * r
            return m, np.diag([ia, ia, ia])
        if gtype == GeoType.BOX:
            hx, hy, hz = scale
            m = density * 8.0 * hx * hy * hz
            return m, np.diag([m / 3.0 * (hy * hy + hz * hz), m / 3.0 * (hx * hx + hz * hz),
                               m / 3.0 * (hx * hx + hy * hy)])
        if gtype == GeoType.CAPSULE:
            r, hh = scale[0], scale[1]
            h = 2.0 * hh
            ms = density * 4.0 / 3.0 * math.pi * r ** 3
            mc = density * math.pi * r * r * h
            ia = mc * (0.25 * r * r + h * h / 12.0) + ms * (0.4 * r * r + 0.375 * r * h + 0.25 * h * h)
            ib = (mc * 0.5 + ms * 0.4) * r * r
            return ms + mc, np.diag([ia, ia, ib])
        return 0.0, np.zeros((3, 3))

    @staticmethod
    def _shape_radius(gtype, scale):
        if gtype == GeoType.SPHERE:
            return abs(scale[0])
        if gtype == GeoType.BOX:
            return float(np.linalg.norm(np.abs(scale)))
        if gtype == GeoType.CAPSULE:
            return abs(scale[0]) + abs(scale[1])
        if gtype == GeoType.PLANE:
            if abs(scale[0]) > 0.0 and abs(scale[1]) > 0.0:
                return float(np.linalg.norm(np.abs(scale))) * 0.5
            return 1.0e6
        return 10.0

    def finalize(self, device="cuda:0", requires_grad: bool = False) -> Model:
        B, S = self.body_count, self.shape_count
        bw = np.asarray(self.body_world, dtype=np.int64)
        sw = np.asarray(self.shape_world, dtype=np.int64)
        sb = np.asarray(self.shape_body, dtype=np.int64)
        if B and np.any(bw < 0):
            # bodies built outside any world form world 0
            if np.any(bw >= 0):
                raise ValueError("mixing bodies inside and outside worlds is not supported")
            bw[:] = 0
            for s in range(S):
                if sb[s] >= 0:
                    sw[s] = 0
        world_count = int(bw.max()) + 1 if B else max(self._num_worlds, 1)
        # contiguity of bodies / world-local shapes per world
        if B and np.any(np.diff(bw) < 0):
            raise ValueError("bodies of each world must be contiguous and ordered by world")
        local = np.where(sw >= 0)[0]
        if len(local) and np.any(np.diff(sw[local]) < 0):
            raise ValueError("world-local shapes must be ordered by world")
        if len(local) and np.any(np.diff(local) != 1):
            raise ValueError("world-local shapes must be contiguous (global shapes go before or after)")
        for s in range(S):
            if sb[s] >= 0 and sw[s] != bw[sb[s]]:
                raise ValueError("shape world differs from its body's world")
        world_body_start = np.zeros(world_count + 1, dtype=np.int32)
        world_shape_start = np.zeros(world_count + 1, dtype=np.int32)
        first_local = int(local[0]) if len(local) else 0
        for w in range(world_count):
            world_body_start[w + 1] = world_body_start[w] + int(np.sum(bw == w))
            world_shape_start[w + 1] = world_shape_start[w] + int(np.sum(sw == w))
        world_shape_start += first_local
        global_shapes = np.where(sw < 0)[0].astype(np.int32)

        # mass properties (composite of each body's shapes about the body COM)
        mass = np.zeros(B)
        com = np.zeros((B, 3))
        inertia = np.zeros((B, 3, 3))
        parts = [[] for _ in range(B)]
        for s in range(S):
            b = sb[s]
            if b < 0:
                continue
            m, I = self._shape_mass(self.shape_type[s], self.shape_scale[s], self.shape_cfg[s].density)
            xf = self.shape_transform[s]
            R = quat_to_matrix(xf[3:])
            parts[b].append((m, np.asarray(xf[:3]), R @ I @ R.T))
        for b in range(B):
            mt = sum(p[0] for p in parts[b])
            if mt > 0.0:
                c = sum(p[0] * p[1] for p in parts[b]) / mt
                I = np.zeros((3, 3))
                for m, p, Ic in parts[b]:
                    d = p - c
                    I += Ic + m * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
                mass[b], com[b], inertia[b] = mt, c, I
            override = self.body_mass_override[b]
            if override is not None and override > 0.0:
                if mass[b] > 0.0:
                    inertia[b] *= override / mass[b]
                else:
                    inertia[b] = np.eye(3) * override * 1e-3
                mass[b] = override
        inv_mass = np.where(mass > 0.0, 1.0 / np.maximum(mass, 1e-30), 0.0)
        inv_inertia = np.zeros_like(inertia)
        for b in range(B):
            if mass[b] > 0.0:
                inv_inertia[b] = np.linalg.inv(inertia[b])

        # filter pairs: shapes on the same body never collide with each other, and the static
        # shapes (body -1) count as one body (builder.py:6204-6208 with body_shapes = {-1: []}).
        # Static shapes of two different worlds can never pass the world test, so their pairs are
        # left out of the list: same-world and world-global static pairs are what remains.
        filt = []

        def _all_pairs(ss):
            for i in range(len(ss)):
                for j in range(i + 1, len(ss)):
                    filt.append((int(min(ss[i], ss[j])), int(max(ss[i], ss[j]))))

        for b in range(B):
            _all_pairs(np.where(sb == b)[0])
        static = np.where(sb < 0)[0]
        glob = static[sw[static] < 0]
        _all_pairs(glob)
        for w in np.unique(sw[static][sw[static] >= 0]):
            local_static = static[sw[static] == w]
            _all_pairs(local_static)
            for s in local_static:
                for g in glob:
                    filt.append((int(min(s, g)), int(max(s, g))))
        filt = np.asarray(sorted(filt), dtype=np.int32).reshape(-1, 2)

        cfg = self.shape_cfg
        host = {
            "world_count": world_count,
            "body_count": B,
            "shape_count": S,
            "gravity": np.asarray([0.0, 0.0, 0.0]) + np.asarray(self.up_vector) * self.gravity,
            "body_q": np.asarray(self.body_q, dtype=np.float64).reshape(B, 7),
            "body_qd": np.asarray(self.body_qd, dtype=np.float64).reshape(B, 6),
            "body_com": com,
            "body_mass": mass,
            "body_inv_mass": inv_mass,
            "body_inertia": inertia,
            "body_inv_inertia": inv_inertia,
            "body_world": bw.astype(np.int32),
            "shape_type": np.asarray(self.shape_type, dtype=np.int32),
            "shape_body": sb.astype(np.int32),
            "shape_transform": np.asarray(self.shape_transform, dtype=np.float64).reshape(S, 7),
            "shape_scale": np.asarray(self.shape_scale, dtype=np.float64).reshape(S, 3),
            "shape_margin": np.asarray([c.margin for c in cfg], dtype=np.float64),
            "shape_gap": np.asarray([c.gap if c.gap is not None else self.rigid_gap for c in cfg],
                                    dtype=np.float64),
            "shape_collision_group": np.asarray(
                [c.collision_group if c.has_shape_collision else 0 for c in cfg], dtype=np.int32),
            "shape_world": sw.astype(np.int32),
            "shape_collision_radius": np.asarray(
                [self._shape_radius(self.shape_type[s], self.shape_scale[s]) for s in range(S)]),
            "shape_material_ke": np.asarray([c.ke for c in cfg], dtype=np.float64),
            "shape_material_kd": np.asarray([c.kd for c in cfg], dtype=np.float64),
            "shape_material_mu": np.asarray([c.mu for c in cfg], dtype=np.float64),
            "shape_collision_filter_pairs": filt,
            "world_body_start": world_body_start,
            "world_shape_start": world_shape_start,
            "global_shapes": global_shapes,
        }
        if self.joints:
            host.update(self._articulation(B, bw, world_count, world_body_start, host))
        return Model(host, device)

    def _articulation(self, B, bw, W, world_body_start, host) -> dict:
        """Joint arrays of an articulated model: one joint per body, in body order, parents first."""
        by_child = {j["child"]: j for j in self.joints}
        if sorted(by_child) != list(range(B)):
            raise ValueError("in an articulated model every body needs exactly one joint (FREE or revolute)")
        J = [by_child[b] for b in range(B)]
        for b, j in enumerate(J):
            if j["type"] == 4 and j["parent"] != -1:
                raise ValueError("FREE joints must have the world as parent")
            if j["parent"] >= 0 and bw[j["parent"]] != bw[b]:
                raise ValueError("a joint connects bodies of one world")
        jtype = np.asarray([j["type"] for j in J], dtype=np.int32)
        nq = np.where(jtype == 4, 7, 1)
        nd = np.where(jtype == 4, 6, 1)
        q_start = np.concatenate([[0], np.cumsum(nq)]).astype(np.int32)
        qd_start = np.concatenate([[0], np.cumsum(nd)]).astype(np.int32)
        joint_q = np.zeros(int(q_start[-1]))
        per_dof = {k: np.zeros(int(qd_start[-1])) for k in ("target_pos", "target_ke", "target_kd", "armature")}
        for b, j in enumerate(J):
            if j["type"] == 4:
                joint_q[q_start[b]:q_start[b] + 7] = self.body_q[b]
            else:
                joint_q[q_start[b]] = j["q0"]
                for k in per_dof:
                    per_dof[k][qd_start[b]] = j[k]
        art = {
            "joint_type": jtype,
            "joint_parent": np.asarray([j["parent"] for j in J], dtype=np.int32),
            "joint_X_p": np.asarray([j["X_p"] for j in J], dtype=np.float64).reshape(B, 7),
            "joint_X_c": np.asarray([j["X_c"] for j in J], dtype=np.float64).reshape(B, 7),
            "joint_axis": np.asarray([j["axis"] for j in J], dtype=np.float64).reshape(B, 3),
            "joint_q_start": q_start, "joint_qd_start": qd_start,
            "world_joint_start": world_body_start.copy(),
            "world_q_start": q_start[world_body_start], "world_qd_start": qd_start[world_body_start],
            "joint_q": joint_q, "joint_qd": np.zeros(int(qd_start[-1])),
            "joint_target_pos": per_dof["target_pos"], "joint_target_ke": per_dof["target_ke"],
            "joint_target_kd": per_dof["target_kd"], "joint_armature": per_dof["armature"],
        }
        # initial body poses from the joint coordinates
        from .articulation import fk_host
        host_fk = dict(host, **art)
        art["body_q"] = fk_host(host_fk, joint_q)
        return art
