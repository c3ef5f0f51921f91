"""Device-resident containers with Newton's names, shapes and units.

``Model`` / ``State`` / ``Control`` keep the attribute names, shapes and dtypes of
the reference's ``newton/_src/sim/model.py:36``, ``state.py:47`` and ``control.py:31``
so host code written against Newton reads the same arrays.  The arrays are torch
tensors on the HIP device; their ``data_ptr()`` values are what cross the C ABI.

Layout rules the HIP kernels rely on (checked in ``ModelBuilder.finalize``):

* the bodies of world ``w`` are contiguous, ``[world_body_start[w], world_body_start[w+1])``;
* the world-local shapes of world ``w`` are contiguous, ``[world_shape_start[w], world_shape_start[w+1])``;
* shapes of world ``-1`` (ground plane, static walls) are listed in ``global_shapes``;
* every body carries one FREE joint, so ``joint_q`` (7 per body: world position of the
  body origin, quaternion ``xyzw``) and ``joint_qd`` (6 per body: COM linear velocity,
  angular velocity, both world frame) hold exactly the numbers of ``body_q`` / ``body_qd``
  (``state.py:110-117,149-153``).  ``State.joint_q`` / ``joint_qd`` are therefore views of
  the same device memory as ``body_q`` / ``body_qd``.
"""

from __future__ import annotations

import numpy as np
import torch


class GeoType:
    """Shape type codes (``newton/_src/geometry/types.py:76-106``)."""

    NONE = 0
    PLANE = 1
    HFIELD = 2
    SPHERE = 3
    CAPSULE = 4
    ELLIPSOID = 5
    CYLINDER = 6
    BOX = 7
    MESH = 8
    CONE = 9
    CONVEX_MESH = 10


SUPPORTED_SHAPES = (GeoType.PLANE, GeoType.SPHERE, GeoType.CAPSULE, GeoType.BOX)


def _t(arr, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(arr), dtype=dtype).to(device)


class State:
    """Per-world mutable state (``newton/_src/sim/state.py:47``).

    ``body_q``  [B, 7] float32: origin position + quaternion (x, y, z, w).
    ``body_qd`` [B, 6] float32: COM linear velocity (world), angular velocity (world).
    ``body_f``  [B, 6] float32: external wrench at the COM (world).
    ``joint_q`` [7B] / ``joint_qd`` [6B]: FREE-joint coordinates, views of the same memory.
    """

    def __init__(self, body_q: torch.Tensor, body_qd: torch.Tensor, body_f: torch.Tensor):
        self.body_q = body_q
        self.body_qd = body_qd
        self.body_f = body_f
        self.joint_q = body_q.view(-1)
        self.joint_qd = body_qd.view(-1)

    @property
    def body_count(self) -> int:
        return int(self.body_q.shape[0])

    def assign(self, other: "State") -> None:
        """Copy ``other`` into this state (``state.py:169``)."""
        self.body_q.copy_(other.body_q)
        self.body_qd.copy_(other.body_qd)
        self.body_f.copy_(other.body_f)
        if self.joint_q.data_ptr() != self.body_q.data_ptr():  # articulated: coordinates of their own
            self.joint_q.copy_(other.joint_q)
            self.joint_qd.copy_(other.joint_qd)

    def clear_forces(self) -> None:
        self.body_f.zero_()


class Control:
    """Control inputs (``newton/_src/sim/control.py:31-180``), float32 per dof of the ``joint_qd``
    layout (the reference's default ``newton.use_coord_layout_targets = False``):

    ``joint_f``          generalized forces; a FREE joint's six entries are a wrench (force; torque)
                         at the child body's COM in the world frame, added to ``State.body_f``;
    ``joint_target_q``   position targets of the joint drives (from the model's ``joint_target_pos``);
    ``joint_target_qd``  velocity targets;
    ``joint_act``        feedforward generalized forces added to the drive.

    A drive applies ``ke (q_t - q) + kd (qd_t - qd) + act`` (gains on the model), implicitly in the
    step.  Every array is read at each boundary call, so a policy acts by writing them between calls.
    """

    def __init__(self, model: "Model"):
        D, dev = model.joint_dof_count, model.device
        self.joint_f = torch.zeros(D, dtype=torch.float32, device=dev)
        tp = getattr(model, "joint_target_pos", None)
        self.joint_target_q = tp.clone() if tp is not None else torch.zeros(D, dtype=torch.float32, device=dev)
        self.joint_target_qd = torch.zeros(D, dtype=torch.float32, device=dev)
        self.joint_act = torch.zeros(D, dtype=torch.float32, device=dev)

    # deprecated aliases of the reference (dof layout)
    @property
    def joint_target_pos(self):
        return self.joint_target_q

    @joint_target_pos.setter
    def joint_target_pos(self, value):
        self.joint_target_q = value

    @property
    def joint_target_vel(self):
        return self.joint_target_qd

    @joint_target_vel.setter
    def joint_target_vel(self, value):
        self.joint_target_qd = value

    def clear(self, model: "Model | None" = None) -> None:
        """Zero every input; ``joint_target_q`` is restored from ``model`` when one is given
        (``control.py:136-166``)."""
        self.joint_f.zero_()
        self.joint_target_qd.zero_()
        self.joint_act.zero_()
        tp = getattr(model, "joint_target_pos", None) if model is not None else None
        if tp is not None:
            self.joint_target_q.copy_(tp)
        else:
            self.joint_target_q.zero_()


class MujocoShapeAttributes:
    """MuJoCo geom parameters per shape (the ``model.mujoco`` custom attributes that
    ``SolverMuJoCo.register_custom_attributes`` adds, ``solver_mujoco.py:712-790``), with the
    reference's defaults: ``condim`` 3, ``geom_priority`` 0, ``geom_solimp`` (0.9, 0.95, 0.001, 0.5,
    2), ``geom_solmix`` 1 and ``solref`` (0, 0), which means "not authored": the solver then derives
    the shape's solref from ``shape_material_ke`` / ``shape_material_kd`` (``convert_solref(ke, kd,
    1, 1)``, ``solver_mujoco.py:5224-5233``).  Edit them before the first solver is created on the
    model.  ``condim`` must be 1 (frictionless) or 3 (sliding friction)."""

    def __init__(self, shape_count: int, device):
        S = int(shape_count)
        self.condim = torch.full((S,), 3, dtype=torch.int32, device=device)
        self.geom_priority = torch.zeros(S, dtype=torch.int32, device=device)
        self.geom_solimp = torch.tensor([[0.9, 0.95, 0.001, 0.5, 2.0]] * S, dtype=torch.float32,
                                        device=device).reshape(S, 5)
        self.geom_solmix = torch.ones(S, dtype=torch.float32, device=device)
        self.solref = torch.zeros((S, 2), dtype=torch.float32, device=device)


class Model:
    """Static scene description (``newton/_src/sim/model.py:36``).

    Built by :meth:`ModelBuilder.finalize`; all arrays live on ``device``.
    """

    def __init__(self, host: dict, device):
        self.device = torch.device(device)
        self._host = host
        d = self.device
        self.world_count = int(host["world_count"])
        self.body_count = int(host["body_count"])
        self.shape_count = int(host["shape_count"])
        self.joint_count = self.body_count
        self.joint_coord_count = 7 * self.body_count
        self.joint_dof_count = 6 * self.body_count
        self.up_axis = 2
        self.gravity = _t(host["gravity"].reshape(1, 3), torch.float32, d)

        self.body_q = _t(host["body_q"], torch.float32, d)
        self.body_qd = _t(host["body_qd"], torch.float32, d)
        self.body_com = _t(host["body_com"], torch.float32, d)
        self.body_mass = _t(host["body_mass"], torch.float32, d)
        self.body_inv_mass = _t(host["body_inv_mass"], torch.float32, d)
        self.body_inertia = _t(host["body_inertia"], torch.float32, d)
        self.body_inv_inertia = _t(host["body_inv_inertia"], torch.float32, d)
        self.body_world = _t(host["body_world"], torch.int32, d)
        self.joint_q = self.body_q.view(-1)
        self.joint_qd = self.body_qd.view(-1)

        self.shape_type = _t(host["shape_type"], torch.int32, d)
        self.shape_body = _t(host["shape_body"], torch.int32, d)
        self.shape_transform = _t(host["shape_transform"], torch.float32, d)
        self.shape_scale = _t(host["shape_scale"], torch.float32, d)
        self.shape_margin = _t(host["shape_margin"], torch.float32, d)
        self.shape_gap = _t(host["shape_gap"], torch.float32, d)
        self.shape_collision_group = _t(host["shape_collision_group"], torch.int32, d)
        self.shape_world = _t(host["shape_world"], torch.int32, d)
        self.shape_collision_radius = _t(host["shape_collision_radius"], torch.float32, d)
        self.shape_material_ke = _t(host["shape_material_ke"], torch.float32, d)
        self.shape_material_kd = _t(host["shape_material_kd"], torch.float32, d)
        self.shape_material_mu = _t(host["shape_material_mu"], torch.float32, d)
        fp = host["shape_collision_filter_pairs"].reshape(-1, 2)
        self.shape_collision_filter_pairs = _t(fp if len(fp) else np.zeros((0, 2)), torch.int32, d)

        self.world_body_start = _t(host["world_body_start"], torch.int32, d)
        self.world_shape_start = _t(host["world_shape_start"], torch.int32, d)
        self.global_shapes = _t(host["global_shapes"], torch.int32, d)
        self.mujoco = MujocoShapeAttributes(self.shape_count, d)

        # articulated worlds: one joint per body (REVOLUTE / FREE root), generalized coordinates
        self.is_articulated = "joint_type" in host
        if self.is_articulated:
            for name in ("joint_type", "joint_parent", "joint_q_start", "joint_qd_start", "world_joint_start",
                         "world_q_start", "world_qd_start"):
                setattr(self, name, _t(host[name], torch.int32, d))
            for name in ("joint_X_p", "joint_X_c", "joint_axis", "joint_target_pos", "joint_target_ke",
                         "joint_target_kd", "joint_armature"):
                setattr(self, name, _t(host[name], torch.float32, d))
            self.joint_q = _t(host["joint_q"], torch.float32, d)
            self.joint_qd = _t(host["joint_qd"], torch.float32, d)
            self.joint_count = self.body_count
            self.joint_coord_count = int(host["joint_q_start"][-1])
            self.joint_dof_count = int(host["joint_qd_start"][-1])

    # --- explicit broad-phase pair list (model.shape_contact_pairs of the reference) ---
    def contact_pairs(self):
        """Candidate shape pairs for the explicit broad phase: canonical ``(min, max)`` global
        shape indices of every pair that the world, collision-group, same-body and filter rules
        let collide (``broad_phase_common.py:18-180``), grouped by world, plus ``[W+1]`` world
        ranges.  Global-global pairs are never included.  Computed on first use; returns host
        arrays ``(pairs [P, 2], world_pair_start [W+1])``."""
        if getattr(self, "_contact_pairs", None) is None:
            h = self._host
            ws, g = h["world_shape_start"], h["global_shapes"].astype(np.int64)
            grp, body, sw = h["shape_collision_group"], h["shape_body"], h["shape_world"]
            S = max(self.shape_count, 1)
            fp = h["shape_collision_filter_pairs"].reshape(-1, 2).astype(np.int64)
            fkey = fp[:, 0] * S + fp[:, 1]
            out, starts = [], [0]
            for w in range(self.world_count):
                seg = np.concatenate([np.arange(ws[w], ws[w + 1], dtype=np.int64), g])
                i, j = np.triu_indices(len(seg), 1)
                a, b = np.minimum(seg[i], seg[j]), np.maximum(seg[i], seg[j])
                ga, gb = grp[a], grp[b]
                ok = ~((sw[a] == -1) & (sw[b] == -1))
                ok &= (ga != 0) & (gb != 0) & np.where(ga > 0, (ga == gb) | (gb < 0), ga != gb)
                ok &= body[a] != body[b]  # one body; the static shapes (-1) count as one body
                if len(fkey):
                    ok &= ~np.isin(a * S + b, fkey)
                out.append(np.stack([a[ok], b[ok]], axis=1))
                starts.append(starts[-1] + int(ok.sum()))
            pairs = np.concatenate(out).astype(np.int32) if out else np.zeros((0, 2), np.int32)
            self._contact_pairs = (pairs.reshape(-1, 2), np.asarray(starts, dtype=np.int32))
        return self._contact_pairs

    # --- host mirrors (numpy) used by the oracle-side tests and the autotuner ---
    def host(self, name: str) -> np.ndarray:
        return self._host[name]

    def bodies_per_world(self) -> np.ndarray:
        s = self._host["world_body_start"]
        return s[1:] - s[:-1]

    def shapes_per_world(self) -> np.ndarray:
        s = self._host["world_shape_start"]
        return s[1:] - s[:-1] + len(self._host["global_shapes"])

    def state(self) -> State:
        """Fresh state initialised from the model's reference pose (``model.py:1572``)."""
        s = State(self.body_q.clone(), self.body_qd.clone(),
                  torch.zeros((self.body_count, 6), dtype=torch.float32, device=self.device))
        if self.is_articulated:
            s.joint_q = self.joint_q.clone()
            s.joint_qd = self.joint_qd.clone()
        return s

    def control(self) -> Control:
        return Control(self)

    def contacts(self, rigid_contact_max: int | None = None):
        from .contacts import Contacts

        cap = rigid_contact_max if rigid_contact_max is not None else 128 * self.world_count
        return Contacts(cap, self.device)

    def collide(self, state: State, contacts=None):
        """One-shot contact generation through a default pipeline (``model.py:1744``)."""
        from .contacts import CollisionPipeline

        if not hasattr(self, "_default_pipeline"):
            self._default_pipeline = CollisionPipeline(self)
        pipe = self._default_pipeline
        if contacts is None:
            contacts = pipe.contacts()
        pipe.collide(state, contacts)
        return contacts


def eval_fk(model: Model, joint_q: torch.Tensor, joint_qd: torch.Tensor, state: State) -> None:
    """Forward kinematics for FREE joints: body pose/twist are the joint coordinates."""
    state.body_q.view(-1).copy_(joint_q.view(-1))
    state.body_qd.view(-1).copy_(joint_qd.view(-1))
