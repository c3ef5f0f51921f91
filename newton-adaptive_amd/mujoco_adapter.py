"""MuJoCo-path adapters: joint coordinate conversion and contact-parameter mixing.

These are what sits around the MuJoCo inner step of ``SolverMuJoCoAdaptive`` in the reference
(``_update_mjc_data`` -> ``mujoco_warp.step`` -> ``_update_newton_state``,
``solver_mujoco.py:4086,3465,4145``), for the joints of this build: free rigid bodies (one FREE
joint each, parent frame the world) and articulations of revolute joints under an optional FREE
root.

* :func:`to_mujoco` / :func:`from_mujoco` run the HIP conversion kernels
  (``na_coords_to_mujoco`` / ``na_coords_from_mujoco``), which replace
  ``convert_warp_coords_to_mj_kernel`` (``kernels.py:705-789``) and
  ``convert_mj_coords_to_warp_kernel`` (``kernels.py:578-666``);
* :func:`convert_solref` is ``kernels.py:193-207`` (stiffness/damping -> MuJoCo time constant and
  damping ratio), :func:`geom_solref` applies it per shape as ``solver_mujoco.py:5233`` does;
* :func:`contact_params` is the per-pair mixing of ``kernels.py:127-190`` (priority, solmix,
  friction, margins, solref, solimp).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .native import model_handle

DEFAULT_LIMIT_SOLREF = (0.02, 1.0)  # constants.py:26
MJ_MINMU = 1.0e-5  # constants.py:41
MJ_MINVAL = 2.220446049250313e-16  # constants.py:44
DEFAULT_SOLIMP = (0.9, 0.95, 0.001, 0.5, 2.0)  # MuJoCo's default geom solimp


def _shapes(model):
    """(rows, qpos width, qvel width): [B, 7] / [B, 6] for free bodies; the flat joint_q / joint_qd
    sizes for articulated models."""
    if model.is_articulated:
        return None, model.joint_coord_count, model.joint_dof_count
    return model.body_count, 7, 6


def to_mujoco(model, state, qpos: torch.Tensor | None = None, qvel: torch.Tensor | None = None):
    """MuJoCo ``qpos`` (origin, quaternion wxyz) and ``qvel`` (origin velocity in world, angular
    velocity in the body frame) of ``state``: [B, 7] / [B, 6] for free bodies, rows of one world
    contiguous; for articulated models the flat ``joint_q`` / ``joint_qd`` layout, FREE roots
    converted (``kernels.py:705-789``), revolute joints copied."""
    B, nq, nv = _shapes(model)
    shp_q, shp_v = ((B, 7), (B, 6)) if B is not None else ((nq,), (nv,))
    qpos = qpos if qpos is not None else torch.empty(shp_q, dtype=torch.float32, device=model.device)
    qvel = qvel if qvel is not None else torch.empty(shp_v, dtype=torch.float32, device=model.device)
    _check_size(qpos, B * nq if B is not None else nq, "qpos")
    _check_size(qvel, B * nv if B is not None else nv, "qvel")
    h = model_handle(model)
    _lib.check(_lib.lib().na_coords_to_mujoco(h.handle, state.joint_q.data_ptr(), state.joint_qd.data_ptr(),
                                              qpos.data_ptr(), qvel.data_ptr(), _lib.stream_handle()),
               "na_coords_to_mujoco")
    return qpos, qvel


def from_mujoco(model, qpos: torch.Tensor, qvel: torch.Tensor, state) -> None:
    """Write MuJoCo ``qpos`` / ``qvel`` back into ``state.joint_q`` / ``joint_qd``
    (``kernels.py:578-666``).  For free bodies these are ``body_q`` / ``body_qd``; an articulated
    state's body poses follow from the coordinates at its next step."""
    B, nq, nv = _shapes(model)
    _check_size(qpos, B * nq if B is not None else nq, "qpos")
    _check_size(qvel, B * nv if B is not None else nv, "qvel")
    h = model_handle(model)
    _lib.check(_lib.lib().na_coords_from_mujoco(h.handle, qpos.data_ptr(), qvel.data_ptr(), state.joint_q.data_ptr(),
                                                state.joint_qd.data_ptr(), _lib.stream_handle()),
               "na_coords_from_mujoco")


def _check_size(t, n, name):
    if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n:
        raise ValueError(f"{name} must be a contiguous float32 tensor of {n} elements")


def convert_solref(ke, kd, d_width=1.0, d_r=1.0) -> np.ndarray:
    """MuJoCo (timeconst, dampratio) from contact stiffness ``ke`` and damping ``kd``:
    ``timeconst = 2 / (kd d_width)``, ``dampratio = kd / 2 sqrt(d_r / ke)``; the default limit
    solref (0.02, 1.0) when ``ke`` or ``kd`` is not positive.  Vectorised over ``ke``/``kd``."""
    ke = np.asarray(ke, dtype=np.float64)
    kd = np.asarray(kd, dtype=np.float64)
    ke, kd = np.broadcast_arrays(ke, kd)
    ok = (ke > 0.0) & (kd > 0.0)
    tc = np.where(ok, 2.0 / np.where(ok, kd * d_width, 1.0), DEFAULT_LIMIT_SOLREF[0])
    dr = np.where(ok, kd / 2.0 * np.sqrt(d_r / np.where(ok, ke, 1.0)), DEFAULT_LIMIT_SOLREF[1])
    return np.stack([tc, dr], axis=-1)


def geom_solref(model) -> np.ndarray:
    """Per-shape solref of ``model`` with ``convert_solref(ke, kd, 1, 1)``."""
    return convert_solref(model.host("shape_material_ke"), model.host("shape_material_kd"), 1.0, 1.0)


def resolved_solref(model) -> np.ndarray:
    """The solref SolverMuJoCo gives each geom (``solver_mujoco.py:5224-5233``): the authored
    ``model.mujoco.solref`` where it is not (0, 0), else ``convert_solref(ke, kd, 1, 1)``."""
    sr = model.mujoco.solref.detach().cpu().numpy().astype(np.float64).reshape(-1, 2)
    authored = (sr[:, 0] != 0.0) | (sr[:, 1] != 0.0)
    return np.where(authored[:, None], sr, geom_solref(model))


def contact_params(pairs, *, priority, solmix, solref, solimp, friction, margin, gap, condim):
    """Per-pair contact parameters from per-shape MuJoCo geom parameters.

    ``pairs`` [P, 2] shape indices; ``priority``, ``condim`` [S] int; ``solmix``, ``margin``,
    ``gap`` [S]; ``solref`` [S, 2]; ``solimp`` [S, 5]; ``friction`` [S, 3] (sliding, torsional,
    rolling).  The higher priority shape takes everything; equal priorities mix solref/solimp by
    ``solmix_a / (solmix_a + solmix_b)`` (0.5 when both are ~0, one-sided when one is), take the
    larger condim and the elementwise max friction; a non-positive solref time constant on either
    side selects the elementwise min solref.  Returns a dict of margin, gap, condim, friction [P, 5]
    (floored at MJ_MINMU), solref [P, 2], solimp [P, 5] and mix [P]."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    a, b = pairs[:, 0], pairs[:, 1]
    priority, condim = np.asarray(priority), np.asarray(condim)
    solmix = np.asarray(solmix, dtype=np.float64)
    solref = np.asarray(solref, dtype=np.float64)
    solimp = np.asarray(solimp, dtype=np.float64)
    friction = np.asarray(friction, dtype=np.float64)
    pa, pb = priority[a], priority[b]
    s1, s2 = solmix[a], solmix[b]
    with np.errstate(divide="ignore", invalid="ignore"):
        mix = np.where(s1 + s2 != 0.0, s1 / (s1 + s2), 0.0)
    lo1, lo2 = s1 < MJ_MINVAL, s2 < MJ_MINVAL
    mix = np.where(lo1 & lo2, 0.5, mix)
    mix = np.where(lo1 & ~lo2, 0.0, mix)
    mix = np.where(~lo1 & lo2, 1.0, mix)
    mix = np.where(pa > pb, 1.0, np.where(pb > pa, 0.0, mix))
    cd = np.where(pa > pb, condim[a], np.where(pb > pa, condim[b], np.maximum(condim[a], condim[b])))
    fr = np.where((pa > pb)[:, None], friction[a],
                  np.where((pb > pa)[:, None], friction[b], np.maximum(friction[a], friction[b])))
    fr5 = np.maximum(MJ_MINMU, np.stack([fr[:, 0], fr[:, 0], fr[:, 1], fr[:, 2], fr[:, 2]], axis=1))
    both_pos = (solref[a, 0] > 0.0) & (solref[b, 0] > 0.0)
    sr = np.where(both_pos[:, None], mix[:, None] * solref[a] + (1.0 - mix[:, None]) * solref[b],
                  np.minimum(solref[a], solref[b]))
    si = mix[:, None] * solimp[a] + (1.0 - mix[:, None]) * solimp[b]
    return {"margin": np.asarray(margin)[a] + np.asarray(margin)[b], "gap": np.asarray(gap)[a] + np.asarray(gap)[b],
            "condim": cd, "friction": fr5, "solref": sr, "solimp": si, "mix": mix}
