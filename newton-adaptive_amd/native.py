"""Native handles: a :class:`Model` bound to ``na_model`` in ``libnadapt.so``."""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib


class ModelHandle:
    """Owns the ``na_model`` of a :class:`Model` (the device arrays stay owned by the Model)."""

    def __init__(self, model, with_pairs: bool = False):
        if model.device.type != "cuda":
            raise RuntimeError("the nadapt library runs on a HIP device; build the model with device='cuda:0'")
        L = _lib.lib()
        d = _lib.NaModelDesc()
        d.world_count = model.world_count
        d.body_count = model.body_count
        d.shape_count = model.shape_count
        d.global_count = int(model.global_shapes.numel())
        d.filter_count = int(model.shape_collision_filter_pairs.shape[0])
        self._keep = []

        def p(t):
            t = t.contiguous()
            self._keep.append(t)
            return t.data_ptr() if t.numel() else None

        # empty arrays still need a valid pointer for kernels that never read them
        dummy_i = torch.zeros(2, dtype=torch.int32, device=model.device)
        self._keep.append(dummy_i)
        d.world_body_start = p(model.world_body_start)
        d.world_shape_start = p(model.world_shape_start)
        d.global_shapes = p(model.global_shapes) or dummy_i.data_ptr()
        d.body_com = p(model.body_com)
        d.body_mass = p(model.body_mass)
        d.body_inv_mass = p(model.body_inv_mass)
        d.body_inertia = p(model.body_inertia)
        d.body_inv_inertia = p(model.body_inv_inertia)
        d.shape_type = p(model.shape_type)
        d.shape_body = p(model.shape_body)
        d.shape_collision_group = p(model.shape_collision_group)
        d.shape_world = p(model.shape_world)
        d.shape_transform = p(model.shape_transform)
        d.shape_scale = p(model.shape_scale)
        d.shape_margin = p(model.shape_margin)
        d.shape_gap = p(model.shape_gap)
        d.shape_material_ke = p(model.shape_material_ke)
        d.shape_material_mu = p(model.shape_material_mu)
        d.shape_collision_radius = p(model.shape_collision_radius)
        d.filter_pairs = p(model.shape_collision_filter_pairs) or dummy_i.data_ptr()
        g = model.gravity.view(-1).cpu().tolist()
        d.gravity = (C.c_float * 3)(*g)
        # a world pool keeps explicit [W] end arrays; a packed model leaves them NULL
        wbe, wse = getattr(model, "world_body_end", None), getattr(model, "world_shape_end", None)
        d.world_body_end = p(wbe) if wbe is not None else None
        d.world_shape_end = p(wse) if wse is not None else None
        # explicit broad-phase pairs, only when a caller asks for that broad phase
        self.with_pairs = False
        d.explicit_pair_count = 0
        if with_pairs:
            pairs, starts = model.contact_pairs()
            if len(pairs):
                d.explicit_pair_count = int(len(pairs))
                d.explicit_pairs = p(torch.as_tensor(pairs, device=model.device))
                d.world_pair_start = p(torch.as_tensor(starts, device=model.device))
                self.with_pairs = True
        # MuJoCo geom parameters (contact law 1); models without them (a serving pool) leave NULL
        mjat = getattr(model, "mujoco", None)
        if mjat is not None and model.shape_count > 0:
            from .mujoco_adapter import resolved_solref
            d.mj_solref = p(torch.as_tensor(resolved_solref(model), dtype=torch.float32, device=model.device))
            d.mj_solimp = p(mjat.geom_solimp.to(torch.float32))
            d.mj_solmix = p(mjat.geom_solmix.to(torch.float32))
            d.mj_priority = p(mjat.geom_priority.to(torch.int32))
            d.mj_condim = p(mjat.condim.to(torch.int32))
            # SolverMuJoCo zeroes every margin of a model with boxes (solver_mujoco.py:3352-3357)
            d.mj_zero_margins = int(bool((model.shape_type == 7).any().item()))
        # articulation (one joint per body); free-body models leave joint_type NULL
        if getattr(model, "is_articulated", False):
            for name in ("joint_type", "joint_parent", "joint_q_start", "joint_qd_start", "joint_X_p", "joint_X_c",
                         "joint_axis", "joint_target_pos", "joint_target_ke", "joint_target_kd", "joint_armature"):
                setattr(d, name, p(getattr(model, name)) or dummy_i.data_ptr())
        self.desc = d
        h = C.c_void_p()
        _lib.check(L.na_model_create(C.byref(d), C.byref(h)), "na_model_create")
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                _lib.lib().na_model_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def model_handle(model, with_pairs: bool = False) -> ModelHandle:
    """The model's cached handle; rebuilt once with the explicit pair list when asked for it."""
    h = getattr(model, "_na_handle", None)
    if h is None or (with_pairs and not h.with_pairs):
        h = ModelHandle(model, with_pairs=with_pairs or (h is not None and h.with_pairs))
        model._na_handle = h
    return h
