"""ctypes binding of ``libnadapt.so`` (declarations in ``include/nadapt.h``).

The library is built in-tree by :func:`build` (``hipcc --offload-arch=gfx950``).  Loading
fails loudly when the library is missing: there is no CPU fallback on this path.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_HERE, "libnadapt.so")
HEADER = os.path.join(_REPO, "include", "nadapt.h")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # per-lane arrays (the work counters) stay private instead of taking 6 KB of each world's LDS,
    # which keeps every free-body kernel at four waves per CU
    "-mllvm", "-disable-promote-alloca-to-lds",
]

vp = C.c_void_p
i32 = C.c_int32


class NaModelDesc(C.Structure):
    _fields_ = [
        ("world_count", i32), ("body_count", i32), ("shape_count", i32), ("global_count", i32),
        ("filter_count", i32),
        ("world_body_start", vp), ("world_shape_start", vp), ("global_shapes", vp),
        ("body_com", vp), ("body_mass", vp), ("body_inv_mass", vp), ("body_inertia", vp), ("body_inv_inertia", vp),
        ("shape_type", vp), ("shape_body", vp), ("shape_collision_group", vp), ("shape_world", vp),
        ("shape_transform", vp), ("shape_scale", vp), ("shape_margin", vp), ("shape_gap", vp),
        ("shape_material_ke", vp), ("shape_material_mu", vp), ("shape_collision_radius", vp),
        ("filter_pairs", vp),
        ("gravity", C.c_float * 3),
        ("world_body_end", vp), ("world_shape_end", vp),
        ("explicit_pair_count", i32), ("explicit_pairs", vp), ("world_pair_start", vp),
        ("joint_type", vp), ("joint_parent", vp), ("joint_q_start", vp), ("joint_qd_start", vp),
        ("joint_X_p", vp), ("joint_X_c", vp), ("joint_axis", vp), ("joint_target_pos", vp),
        ("joint_target_ke", vp), ("joint_target_kd", vp), ("joint_armature", vp),
        ("mj_solref", vp), ("mj_solimp", vp), ("mj_solmix", vp), ("mj_priority", vp), ("mj_condim", vp),
        ("mj_zero_margins", i32),
    ]


class NaConfig(C.Structure):
    _fields_ = [
        ("mode", i32), ("controller", i32),
        ("tol", C.c_double), ("dt_inner_init", C.c_double), ("dt_inner_min", C.c_double),
        ("dt_inner_max", C.c_double), ("divergence_threshold", C.c_double),
        ("max_substeps", i32), ("max_iterations", i32), ("max_rigid_contact", i32), ("broad_phase", i32),
        ("optimality_rel_tol", C.c_float), ("optimality_abs_tol", C.c_float), ("contact_tau_d", C.c_float),
        ("sigma", C.c_float), ("beta", C.c_float), ("ls_rel_tol", C.c_float), ("cg_rel_tol", C.c_float),
        ("max_cg", i32), ("cg_forcing", C.c_float), ("line_search", i32), ("contact_law", i32),
        ("impratio", C.c_float), ("warm_start", i32),
    ]


class NaControlDesc(C.Structure):
    _fields_ = [(n, vp) for n in ("joint_f", "joint_target_q", "joint_target_qd", "joint_act")]


class NaTelemetry(C.Structure):
    _fields_ = [(n, vp) for n in (
        "dt", "dt_half", "ideal_dt", "sim_time", "next_time", "last_raw_error", "last_error", "accepted",
        "commit", "diverged", "substeps", "iterations", "solve_fail", "overflow", "cum_accepted", "work",
        "body_q64", "body_trend")]


class NaContactsDesc(C.Structure):
    _fields_ = [("capacity", i32)] + [(n, vp) for n in (
        "count", "shape0", "shape1", "point0", "point1", "offset0", "offset1", "normal", "margin0", "margin1",
        "world_count_out", "scratch")]


# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "na_last_error_string": (C.c_char_p, []),
    "na_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "na_model_create": (C.c_int, [C.POINTER(NaModelDesc), C.POINTER(vp)]),
    "na_model_destroy": (C.c_int, [vp]),
    "na_solver_create": (C.c_int, [vp, C.POINTER(NaConfig), vp, C.POINTER(vp)]),
    "na_solver_destroy": (C.c_int, [vp]),
    "na_solver_telemetry": (C.c_int, [vp, C.POINTER(NaTelemetry)]),
    "na_solver_set_schedule": (C.c_int, [vp, C.c_int, vp]),
    "na_step": (C.c_int, [vp, vp, vp, vp, vp, C.c_double, vp]),
    "na_step_plain": (C.c_int, [vp, vp, vp, vp, vp, C.c_double, C.POINTER(i32), vp]),
    "na_reset": (C.c_int, [vp, vp, vp]),
    "na_coords_to_mujoco": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "na_match_contacts": (C.c_int, [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, C.c_float, C.c_float, vp, vp,
                                    vp, vp, vp, vp, vp, vp]),
    "na_coords_from_mujoco": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "na_serve": (C.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_double, i32, vp, vp]),
    "na_step_articulated": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, C.c_double, vp]),
    "na_solver_set_plain_batch": (C.c_int, [vp, i32]),
    "na_solver_set_tiles": (C.c_int, [vp, i32]),
    "na_serve_articulated": (C.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                       C.c_double, i32, vp, vp]),
    "na_serve_mixed": (C.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                 C.c_double, i32, vp, vp]),
    "na_art_substep": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "na_art_attempt": (C.c_int, [vp] * 16),
    "na_collide_scratch_words": (C.c_int64, [vp, i32]),
    "na_sort_contacts_workspace_bytes": (C.c_int64, [i32]),
    "na_sort_contacts": (C.c_int, [C.POINTER(NaContactsDesc), i32, vp, vp, vp, i32, vp, vp, vp, C.c_int64, vp]),
    "na_collide": (C.c_int, [vp, vp, C.POINTER(NaContactsDesc), i32, i32, vp]),
    "na_collide_noalloc": (C.c_int, [vp, vp, C.POINTER(NaContactsDesc), i32, i32, vp]),
    "na_collide_debug": (C.c_int, [vp, vp, vp, vp, vp, vp, i32, i32, vp]),
    "na_substep": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "na_attempt": (C.c_int, [vp] * 16),
    "na_ctrl_calc_adjusted_step_f32": (C.c_int, [i32, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, vp]),
    "na_ctrl_adapt_dt_f32": (C.c_int, [i32] + [vp] * 11 + [i32, C.c_float, C.c_float, C.c_float, C.c_float, vp]),
    "na_ctrl_rebase_time_f32": (C.c_int, [i32, vp, vp, vp]),
    "na_ctrl_reset_worlds_f32": (C.c_int, [i32, vp, C.c_float, vp, vp, vp, vp, vp, vp, vp, vp]),
    "na_ctrl_clamp_dt_to_boundary_f32": (C.c_int, [i32, vp, vp, vp, vp, vp]),
    "na_ctrl_inf_norm_error_f32": (C.c_int, [i32, i32, vp, vp, vp, vp, vp]),
}


def build(verbose: bool = False) -> str:
    """Compile ``csrc/nadapt.hip`` and ``csrc/na_sort.hip`` for gfx950 into ``libnadapt.so`` next
    to this file."""
    srcs = [os.path.join(_HERE, "csrc", f) for f in ("nadapt.hip", "na_sort.hip")]
    tmp = LIB_PATH + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(" ".join(cmd))
    return LIB_PATH


_lib = None


def lib():
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        path = os.environ.get("NADAPT_LIB", LIB_PATH)  # alternative build (e.g. with phase timers)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str) -> int:
    if status < 0:
        msg = lib().na_last_error_string()
        raise RuntimeError(f"{what}: {msg.decode() if msg else 'error'}")
    return status


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream_handle(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
