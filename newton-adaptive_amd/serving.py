"""Continuous-batching rollout serving of the step-doubling rigid-body step.

A :class:`ServingEngine` keeps a device-resident pool of worlds.  Each pool slot owns an arena
range of bodies and of shapes; the global shapes (ground, bin walls) are shared by every slot.
The engine runs a host loop of *rounds*:

1. admit pending requests, oldest first, while a slot, arena space and the HBM budget allow;
   their arrays are copied into the pool behind the running launch on the same stream;
2. launch ``na_serve`` over every resident world that has not finished: one persistent launch in
   which waves take worlds from an atomic list and advance each one through its boundary frames
   at its own adaptive dt until its horizon or the round's device-time budget;
3. retire finished requests (final state copied out, latency recorded) and free their slots.

The rounds are pipelined: round k+1's launch is enqueued before the host waits for round k, and
all host<->device copies go through pinned buffers without waiting for the running launch, so
the device does not idle while the host retires and admits.

Requests that arrive while others are on the GPU join the next round's launch, so their work
overlaps on the device with the older rollouts.  A world interrupted by the budget resumes with
its exact state, and its arithmetic never depends on the slot, arena position or the other
worlds, so a request's result is bit-identical alone and in any mix.

There is no counterpart of this loop in the reference; it wraps the same per-world boundary
loop as ``SolverSAPAdaptive.step`` (``solver_sap_adaptive.py:877-975``).
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .model import _t
from .solvers import SolverMuJoCoAdaptive, SolverSAPAdaptive

_I32_MAX = 2**31 - 1


def _dev(arr, dtype, device) -> torch.Tensor:
    """Host array -> device tensor through a pinned staging buffer, enqueued on the current stream
    without waiting for the work already on it (a copy from pageable memory would first wait for
    the running launch)."""
    t = torch.as_tensor(np.ascontiguousarray(arr)).to(dtype)
    return t.pin_memory().to(device, non_blocking=True)


@dataclass
class RolloutRequest:
    """One rollout: a single world (host arrays), its horizon and tolerance.

    The world runs boundary frames of the engine's ``dt_outer`` until it has completed
    ``horizon_frames`` frames or taken ``horizon_accepted`` accepted steps, whichever comes first
    (``None`` = no limit of that kind; at least one must be set).
    """

    bodies: dict
    shapes: dict
    joints: dict | None = None  # articulated world: joint, dof and coordinate arrays (local indices)
    horizon_frames: int | None = None
    horizon_accepted: int | None = None
    tol: float | None = None
    arrival: float = 0.0
    rid: int = -1
    source_world: int = -1  # world of the source model an articulated request was taken from (scenes)
    # filled in by the engine
    admitted_at: float = math.nan
    finished_at: float = math.nan
    status: int = 0
    accepted: int = 0
    frames: int = 0
    body_q: np.ndarray | None = None
    body_qd: np.ndarray | None = None
    joint_q: np.ndarray | None = None
    joint_qd: np.ndarray | None = None
    cost_estimate: float = math.nan  # remaining work from the measured cost so far (tiles pick by it)
    slot: int = -1
    b0: int = -1
    s0: int = -1
    c0: int = -1
    d0: int = -1

    @property
    def articulated(self) -> bool:
        return self.joints is not None

    @property
    def coord_count(self) -> int:
        return int(self.joints["joint_q"].shape[0]) if self.joints is not None else 0

    @property
    def dof_count(self) -> int:
        return int(self.joints["joint_qd"].shape[0]) if self.joints is not None else 0

    @property
    def body_count(self) -> int:
        return int(self.bodies["body_q"].shape[0])

    @property
    def shape_count(self) -> int:
        return int(self.shapes["shape_type"].shape[0])

    @property
    def latency(self) -> float:
        return self.finished_at - self.arrival


_BODY_FIELDS = ("body_q", "body_qd", "body_com", "body_mass", "body_inv_mass", "body_inertia", "body_inv_inertia")
_SHAPE_FIELDS = ("shape_type", "shape_body", "shape_collision_group", "shape_transform", "shape_scale",
                 "shape_margin", "shape_gap", "shape_material_ke", "shape_material_mu", "shape_collision_radius")


_JOINT_BODY_FIELDS = ("joint_type", "joint_parent", "joint_q_start", "joint_qd_start", "joint_X_p", "joint_X_c",
                      "joint_axis")
_JOINT_DOF_FIELDS = ("joint_target_pos", "joint_target_ke", "joint_target_kd", "joint_armature")


def world_arrays(model, w: int, state=None) -> tuple[dict, dict, dict | None]:
    """Host arrays of world ``w`` of a finalized :class:`Model` (shape bodies as local indices):
    ``(bodies, shapes, joints)``, ``joints`` None for a free-body model.  With ``state`` (a dict of
    host arrays), the world starts from that state's body_q / body_qd (and joint_q / joint_qd)."""
    wb, ws = model.host("world_body_start"), model.host("world_shape_start")
    b0, b1, s0, s1 = int(wb[w]), int(wb[w + 1]), int(ws[w]), int(ws[w + 1])
    bodies = {k: np.array(model.host(k)[b0:b1]) for k in _BODY_FIELDS}
    if state is not None:
        bodies["body_q"] = state["body_q"][b0:b1].copy()
        bodies["body_qd"] = state["body_qd"][b0:b1].copy()
    shapes = {k: np.array(model.host(k)[s0:s1]) for k in _SHAPE_FIELDS}
    sb = shapes["shape_body"].astype(np.int64)
    shapes["shape_body"] = np.where(sb >= 0, sb - b0, -1).astype(np.int32)
    joints = None
    if getattr(model, "is_articulated", False):
        qs, ds = model.host("joint_q_start"), model.host("joint_qd_start")
        q0, q1, d0, d1 = int(qs[b0]), int(qs[b1]), int(ds[b0]), int(ds[b1])
        joints = {k: np.array(model.host(k)[b0:b1]) for k in _JOINT_BODY_FIELDS}
        par = joints["joint_parent"].astype(np.int64)
        joints["joint_parent"] = np.where(par >= 0, par - b0, -1).astype(np.int32)
        joints["joint_q_start"] = (joints["joint_q_start"] - q0).astype(np.int32)
        joints["joint_qd_start"] = (joints["joint_qd_start"] - d0).astype(np.int32)
        joints.update({k: np.array(model.host(k)[d0:d1]) for k in _JOINT_DOF_FIELDS})
        src = state if state is not None and "joint_q" in state else None
        joints["joint_q"] = np.array((src["joint_q"] if src else model.host("joint_q"))[q0:q1], dtype=np.float32)
        joints["joint_qd"] = np.array((src["joint_qd"] if src else model.host("joint_qd"))[d0:d1], dtype=np.float32)
    return bodies, shapes, joints


def requests_from_model(model, *, horizon_frames=None, horizon_accepted=None, tol=None, arrivals=None,
                        worlds=None, state=None) -> list:
    """One :class:`RolloutRequest` per world of ``model`` (scalars or per-world sequences).
    ``state``: start the worlds from this :class:`State` (float32 pose and velocity)."""
    worlds = range(model.world_count) if worlds is None else worlds
    host_state = None
    if state is not None:
        host_state = {"body_q": state.body_q.detach().cpu().numpy(), "body_qd": state.body_qd.detach().cpu().numpy()}
        if getattr(model, "is_articulated", False):
            host_state["joint_q"] = state.joint_q.detach().cpu().numpy()
            host_state["joint_qd"] = state.joint_qd.detach().cpu().numpy()
    out = []
    for i, w in enumerate(worlds):
        b, s, j = world_arrays(model, w, host_state)

        def pick(x):
            return x[i] if isinstance(x, (list, tuple, np.ndarray)) else x

        out.append(RolloutRequest(b, s, j, horizon_frames=pick(horizon_frames), horizon_accepted=pick(horizon_accepted),
                                  tol=pick(tol), arrival=float(pick(arrivals)) if arrivals is not None else 0.0,
                                  rid=i))
    return out


def autotune_served(run, candidates=None) -> dict:
    """Pick the schedule of a served rollout by timing it on the device.

    A candidate is a dict with ``waves_per_cu`` (the persistent grid), ``tiles`` (how many worlds
    of each launch run as two-wave tiles, the ones with the most measured work left) and
    ``round_ms`` (device time per launch; 0: one launch for the whole rollout).  ``run(grid_waves,
    tiles, round_ms)`` performs the served rollout of a fixed set of requests from the same start
    (a fresh engine each time) and returns ``(ms, signature)``, ``signature`` a tensor of the
    rollout's results.  A schedule decides only which waves run a world and when, never its
    arithmetic, so every candidate must return the same signature bit for bit; a mismatch raises.
    Returns ``{"best": row, "candidates": rows}``, each row the candidate with ``grid_waves`` and
    ``ms``."""
    if candidates is None:
        candidates = [{"waves_per_cu": w, "tiles": 0, "round_ms": 0.0} for w in (4, 3, 2)]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    rows, ref = [], None
    for c in candidates:
        grid = int(c["waves_per_cu"]) * cus
        ms, sig = run(grid, int(c.get("tiles", 0)), float(c.get("round_ms", 0.0)))
        if ref is None:
            ref = sig
        elif not torch.equal(sig, ref):
            raise RuntimeError("autotune_served: schedules gave different results (a schedule changed arithmetic)")
        rows.append(dict(c, grid_waves=grid, ms=float(ms)))
    return {"best": min(rows, key=lambda r: r["ms"]), "candidates": rows}


class _Arena:
    """First-fit allocator of contiguous index ranges in ``[0, capacity)``."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]  # sorted (start, length)

    def alloc(self, n: int):
        if n == 0:
            return 0
        for i, (s, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, l - n)
                return s
        return None

    def release(self, start: int, n: int) -> None:
        if n == 0:
            return
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((s, l))
        self.free = merged

    @property
    def free_total(self) -> int:
        return sum(l for _, l in self.free)


class WorldPool:
    """Device arrays of a world pool, laid out like a :class:`Model` so the solver binds to it.

    Bodies ``[0, body_capacity)`` and shapes ``[0, shape_capacity)`` are arena space; the global
    shapes of ``globals_model`` follow at ``[shape_capacity, shape_capacity + G)`` so that, as in a
    packed model, every global shape has a larger index than every world shape.  Slot ``w`` holds
    its bodies in ``[world_body_start[w], world_body_end[w])`` and its shapes likewise.
    """

    def __init__(self, globals_model, max_worlds: int, body_capacity: int, shape_capacity: int, device,
                 articulated: bool = False):
        self.device = torch.device(device)
        d = self.device
        g = globals_model.host("global_shapes").astype(np.int64)
        G = len(g)
        self.world_count = int(max_worlds)
        self.body_count = int(body_capacity)
        self.shape_count = int(shape_capacity) + G
        self.shape_capacity = int(shape_capacity)
        self.joint_count = self.body_count
        self.joint_coord_count = 7 * self.body_count
        self.joint_dof_count = 6 * self.body_count
        self.gravity = globals_model.gravity.clone().to(d)
        f32, i32 = dict(dtype=torch.float32, device=d), dict(dtype=torch.int32, device=d)
        Bc, Sc = self.body_count, self.shape_count
        self.body_q = torch.zeros((Bc, 7), **f32)
        self.body_q[:, 6] = 1.0
        self.body_qd = torch.zeros((Bc, 6), **f32)
        self.body_f = torch.zeros((Bc, 6), **f32)
        self.body_com = torch.zeros((Bc, 3), **f32)
        self.body_mass = torch.zeros(Bc, **f32)
        self.body_inv_mass = torch.zeros(Bc, **f32)
        self.body_inertia = torch.zeros((Bc, 3, 3), **f32)
        self.body_inv_inertia = torch.zeros((Bc, 3, 3), **f32)
        self.shape_type = torch.zeros(Sc, **i32)
        self.shape_body = torch.full((Sc,), -1, **i32)
        self.shape_collision_group = torch.zeros(Sc, **i32)
        self.shape_world = torch.full((Sc,), -1, **i32)
        self.shape_transform = torch.zeros((Sc, 7), **f32)
        self.shape_scale = torch.zeros((Sc, 3), **f32)
        self.shape_margin = torch.zeros(Sc, **f32)
        self.shape_gap = torch.zeros(Sc, **f32)
        self.shape_material_ke = torch.zeros(Sc, **f32)
        self.shape_material_mu = torch.zeros(Sc, **f32)
        self.shape_collision_radius = torch.zeros(Sc, **f32)
        self.shape_collision_filter_pairs = torch.zeros((0, 2), **i32)
        W = self.world_count
        self.world_body_start = torch.zeros(W, **i32)
        self.world_body_end = torch.zeros(W, **i32)
        self.world_shape_start = torch.zeros(W, **i32)
        self.world_shape_end = torch.zeros(W, **i32)
        self.global_shapes = torch.arange(self.shape_capacity, Sc, **i32)
        # global shapes: copied with the same float32 rounding a finalized Model applies
        gs = slice(self.shape_capacity, Sc)
        self.shape_type[gs] = _t(globals_model.host("shape_type")[g], torch.int32, d)
        self.shape_collision_group[gs] = _t(globals_model.host("shape_collision_group")[g], torch.int32, d)
        for name in ("shape_transform", "shape_scale", "shape_margin", "shape_gap", "shape_material_ke",
                     "shape_material_mu", "shape_collision_radius"):
            getattr(self, name)[gs] = _t(globals_model.host(name)[g], torch.float32, d)
        # articulated slots: joint arrays per body slot (joint_type 0 marks the bodies of free-body
        # worlds), and coordinate / dof arenas of 7 / 6 entries per body slot
        self.is_articulated = bool(articulated)
        if self.is_articulated:
            Qc, Dc = 7 * Bc, 6 * Bc
            self.coord_capacity, self.dof_capacity = Qc, Dc
            self.joint_type = torch.zeros(Bc, **i32)
            self.joint_parent = torch.full((Bc,), -1, **i32)
            self.joint_q_start = torch.zeros(Bc + 1, **i32)
            self.joint_qd_start = torch.zeros(Bc + 1, **i32)
            ident = torch.tensor([0, 0, 0, 0, 0, 0, 1], **f32)
            self.joint_X_p = ident.repeat(Bc, 1)
            self.joint_X_c = ident.repeat(Bc, 1)
            self.joint_axis = torch.zeros((Bc, 3), **f32)
            for name in _JOINT_DOF_FIELDS:
                setattr(self, name, torch.zeros(Dc, **f32))
            self.joint_q = torch.zeros(Qc, **f32)
            self.joint_qd = torch.zeros(Dc, **f32)
            self.joint_q64 = torch.zeros(Qc, dtype=torch.float64, device=d)
            self.joint_qd64 = torch.zeros(Dc, dtype=torch.float64, device=d)

    def state(self):
        from .model import State

        return State(self.body_q, self.body_qd, self.body_f)


def _bytes_per_body() -> int:
    # pool arrays (q 28, qd 24, f 24, com 12, mass 8, inertia 72) + FP64 pose shadow (56)
    return 28 + 24 + 24 + 12 + 8 + 72 + 56


def _bytes_per_shape() -> int:
    return 4 * 4 + 28 + 12 + 6 * 4


def _bytes_per_coord() -> int:
    return 4 + 8  # joint_q + FP64 shadow


def _bytes_per_dof() -> int:
    return 4 + 8 + 16  # joint_qd + FP64 shadow + drive targets, gains, armature


def _bytes_per_joint() -> int:
    return 4 * 4 + 17 * 4  # type, parent, q / qd starts; X_p, X_c, axis


def _bytes_per_slot() -> int:
    # controller vectors (7 f64, 3 u8, 4 i32), work counters (12 i64), serving vectors (5 i32, 1 f64),
    # world ranges (4 i32)
    return 7 * 8 + 3 + 4 * 4 + 12 * 8 + 5 * 4 + 8 + 4 * 4


_BODY_SHAPES = {"body_q": (7,), "body_qd": (6,), "body_com": (3,), "body_mass": (), "body_inv_mass": (),
                "body_inertia": (3, 3), "body_inv_inertia": (3, 3)}
_SHAPE_SHAPES = {"shape_type": (), "shape_body": (), "shape_collision_group": (), "shape_transform": (7,),
                 "shape_scale": (3,), "shape_margin": (), "shape_gap": (), "shape_material_ke": (),
                 "shape_material_mu": (), "shape_collision_radius": ()}
_WORLD_SHAPE_TYPES = (1, 3, 4, 7)  # plane, sphere, capsule, box (GeoType codes)
_JT_REVOLUTE, _JT_FREE = 1, 4


def _check_rows(arrays: dict, spec: dict, n: int, what: str) -> None:
    for k, tail in spec.items():
        if k not in arrays:
            raise ValueError(f"{what} array {k!r} is missing")
        a = np.asarray(arrays[k])
        if a.size != n * int(np.prod(tail, dtype=np.int64)):
            raise ValueError(f"{what} array {k!r} must hold {n} rows of shape {tail}, got shape {a.shape}")
        if a.dtype.kind == "f" and not np.all(np.isfinite(a)):
            raise ValueError(f"{what} array {k!r} holds non-finite values")


def _check_joints(r: RolloutRequest) -> None:
    """The articulation rules the device kernels index by (the host twin of the library's
    validate_articulation): one joint per body, revolute (1 coordinate / 1 dof, parent -1 or an
    earlier body of the world) or FREE (root only: parent -1, 7 coordinates / 6 dofs), contiguous
    coordinate and dof starts from 0, at most 16 bodies and 32 coordinates / dofs."""
    j, nb = r.joints, r.body_count
    for k in ("joint_type", "joint_parent", "joint_q_start", "joint_qd_start", "joint_X_p", "joint_X_c",
              "joint_axis", "joint_q", "joint_qd", "joint_target_pos", "joint_target_ke", "joint_target_kd",
              "joint_armature"):
        if k not in j:
            raise ValueError(f"joint array {k!r} is missing")
    jt = np.asarray(j["joint_type"]).reshape(-1)
    par = np.asarray(j["joint_parent"]).reshape(-1)
    qs = np.asarray(j["joint_q_start"]).reshape(-1).astype(np.int64)
    ds = np.asarray(j["joint_qd_start"]).reshape(-1).astype(np.int64)
    if jt.size != nb or par.size != nb:
        raise ValueError("joint_type and joint_parent need one entry per body")
    if qs.size != nb or ds.size != nb or qs[0] != 0 or ds[0] != 0:
        raise ValueError("joint_q_start / joint_qd_start need one entry per body, starting at 0")
    # the last joint ends at the world's coordinate / dof count
    qs = np.append(qs, r.coord_count)
    ds = np.append(ds, r.dof_count)
    for b in range(nb):
        t, p = int(jt[b]), int(par[b])
        nq, nd = int(qs[b + 1] - qs[b]), int(ds[b + 1] - ds[b])
        if t == _JT_FREE:
            if p != -1:
                raise ValueError(f"joint {b}: a FREE joint must have the world (-1) as parent")
            if (nq, nd) != (7, 6):
                raise ValueError(f"joint {b}: a FREE joint has 7 coordinates and 6 dofs")
        elif t == _JT_REVOLUTE:
            if (nq, nd) != (1, 1):
                raise ValueError(f"joint {b}: a revolute joint has 1 coordinate and 1 dof")
            if p != -1 and not 0 <= p < b:
                raise ValueError(f"joint {b}: the parent must be -1 or an earlier body of the world, got {p}")
        else:
            raise ValueError(f"joint {b}: joint types must be 1 (revolute) or 4 (FREE), got {t}")
    if nb > 16 or r.coord_count > 32 or r.dof_count > 32:
        raise ValueError("an articulated world holds at most 16 bodies, 32 coordinates and 32 dofs")
    for k, n, tail in (("joint_X_p", nb, 7), ("joint_X_c", nb, 7), ("joint_axis", nb, 3),
                       ("joint_target_pos", r.dof_count, 1), ("joint_target_ke", r.dof_count, 1),
                       ("joint_target_kd", r.dof_count, 1), ("joint_armature", r.dof_count, 1)):
        a = np.asarray(j[k])
        if a.size != n * tail or not np.all(np.isfinite(a)):
            raise ValueError(f"joint array {k!r} must hold {n} finite rows of {tail}")


def check_request(r: RolloutRequest, global_count: int, budget_bytes: int) -> None:
    """Raise ``ValueError`` for a request the pool can never hold or whose indices the device
    kernels could not use: array sizes, shape types and body references, and the joint rules."""
    if r.horizon_frames is None and r.horizon_accepted is None:
        raise ValueError("a request needs horizon_frames or horizon_accepted")
    for h in (r.horizon_frames, r.horizon_accepted):
        if h is not None and not 1 <= int(h) <= _I32_MAX - 1:
            raise ValueError(f"horizons must be in [1, 2^31 - 2], got {h!r}")
    if r.tol is not None and not float(r.tol) > 0.0:
        raise ValueError(f"tol must be > 0, got {r.tol!r}")
    if "body_q" not in r.bodies or "shape_type" not in r.shapes:
        raise ValueError("a request needs body_q and shape_type arrays")
    if r.body_count < 1 or r.body_count > 64 or r.shape_count + global_count > 128:
        raise ValueError("a world holds 1..64 bodies and at most 128 segment shapes (world + global)")
    _check_rows(r.bodies, _BODY_SHAPES, r.body_count, "body")
    _check_rows(r.shapes, _SHAPE_SHAPES, r.shape_count, "shape")
    st = np.asarray(r.shapes["shape_type"]).reshape(-1)
    if not np.all(np.isin(st, _WORLD_SHAPE_TYPES)):
        raise ValueError(f"shape types must be among {_WORLD_SHAPE_TYPES} (plane, sphere, capsule, box)")
    sb = np.asarray(r.shapes["shape_body"]).reshape(-1).astype(np.int64)
    if np.any((sb < -1) | (sb >= r.body_count)):
        raise ValueError("shape_body must be -1 (static) or a local body index in [0, body_count)")
    if r.articulated:
        _check_joints(r)
    if ServingEngine.request_bytes(r) > budget_bytes:
        raise ValueError("request larger than the HBM budget")


@dataclass
class _Snapshot:
    """Host copies of a launch's per-world counters, valid once ``event`` has completed."""

    event: torch.cuda.Event
    bufs: list  # pinned status, accepted, frames_done [W]
    live: list  # the requests listed in that launch
    work: torch.Tensor | None = None  # pinned per-slot work counters [W, NA_WORK] (engines with tiles)


@dataclass
class ServingStats:
    completed: int = 0
    failed: int = 0
    accepted_steps: int = 0
    wall_s: float = 0.0
    rounds: int = 0
    device_s: float = 0.0  # summed device time of the launches (HIP events around each)
    latencies: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.asarray(self.latencies, dtype=np.float64)
        pct = (lambda p: float(np.percentile(lat, p))) if len(lat) else (lambda p: math.nan)
        return {"completed": self.completed, "failed": self.failed, "accepted_steps": self.accepted_steps,
                "wall_s": self.wall_s, "rounds": self.rounds,
                "accepted_steps_per_s": self.accepted_steps / self.wall_s if self.wall_s > 0 else math.nan,
                "latency_p50_s": pct(50), "latency_p99_s": pct(99),
                "latency_max_s": float(lat.max()) if len(lat) else math.nan,
                "device_busy": self.device_s / self.wall_s if self.wall_s > 0 else math.nan}


class ServingEngine:
    """Persistent request loop over a device-resident world pool (see the module docstring).

    ``globals_model``: a finalized model whose world ``-1`` shapes are the shared environment.
    ``hbm_budget_bytes``: admission limit on the resident bytes of admitted requests (bodies,
    shapes and per-slot state); by default the whole pool.  ``budget_ms``: device time per
    launch (<= 0: a launch runs every resident world to its horizon).  Remaining keyword
    arguments go to the solver (``tol`` is the default per-request tolerance).

    ``tiles``: the first ``tiles`` worlds of each launch's list (the oldest requests, largest
    first) run as two-wave tiles (``SolverSAPAdaptive.set_tiles``; needs
    ``warm_start="concurrent"``).

    ``schedule="continuous"`` is the served form described in the module docstring.
    ``schedule="static"`` is the fixed-batch baseline with the reference's host loop: when the
    pool is empty it admits every arrived request that fits, then advances the whole batch in
    lock step, one launch per step-doubling iteration followed by a 4-byte "any world
    unfinished" readback, until every world of the batch has reached its horizon; results are
    returned when the batch ends.  Both schedules give every request the same arithmetic, so
    their results are bit-identical.
    """

    def __init__(self, globals_model, *, max_worlds: int = 4096, body_capacity: int | None = None,
                 shape_capacity: int | None = None, hbm_budget_bytes: int | None = None, dt_outer: float = 0.01,
                 budget_ms: float = 20.0, controller: str = "sap", schedule: str = "continuous", device="cuda:0",
                 articulated: bool = False, split_kinds: bool = False, tiles: int = 0, **solver_kw):
        if schedule not in ("continuous", "static"):
            raise ValueError(f"schedule must be 'continuous' or 'static', got {schedule!r}")
        self.schedule = schedule
        self.split_kinds = bool(split_kinds)
        if max_worlds < 1:
            raise ValueError("max_worlds must be >= 1")
        if not dt_outer > 0:
            raise ValueError("dt_outer must be > 0")
        body_capacity = int(body_capacity if body_capacity is not None else 32 * max_worlds)
        shape_capacity = int(shape_capacity if shape_capacity is not None else 32 * max_worlds)
        self.pool = WorldPool(globals_model, max_worlds, body_capacity, shape_capacity, device, articulated)
        self.G = int(self.pool.global_shapes.numel())
        self.dt_outer = float(dt_outer)
        self.budget_ms = float(budget_ms)
        cls = {"sap": SolverSAPAdaptive, "mujoco": SolverMuJoCoAdaptive}[controller]
        if controller == "mujoco":
            # requests carry no MuJoCo geom parameters: the pool solves contacts with SAP
            solver_kw.setdefault("contact_model", "sap")
        solver_kw.setdefault("max_substeps", 256)
        solver_kw.setdefault("world_order", "index")
        self.solver = cls(self.pool, **solver_kw)
        if tiles:
            self.solver.set_tiles(tiles)
        self.default_tol = self.solver._tol
        d = self.pool.device
        W = self.pool.world_count
        i32 = dict(dtype=torch.int32, device=d)
        self.horizon_frames = torch.zeros(W, **i32)
        self.horizon_accepted = torch.zeros(W, **i32)
        self.frames_done = torch.zeros(W, **i32)
        self.accepted = torch.zeros(W, **i32)
        self.status = torch.zeros(W, **i32)
        self.world_tol = torch.full((W,), self.default_tol, dtype=torch.float64, device=d)
        self.capacity_bytes = self.pool_bytes(W, body_capacity, shape_capacity, articulated)
        self.hbm_budget_bytes = int(hbm_budget_bytes) if hbm_budget_bytes is not None else self.capacity_bytes
        self.resident_bytes = 0
        self._bodies = _Arena(body_capacity)
        self._shapes = _Arena(shape_capacity)
        self._coords = _Arena(self.pool.coord_capacity if articulated else 0)
        self._dofs = _Arena(self.pool.dof_capacity if articulated else 0)
        self._free_slots = list(range(W - 1, -1, -1))
        self.pending: list[RolloutRequest] = []
        self.resident: dict[int, RolloutRequest] = {}
        self.finished: list[RolloutRequest] = []
        self.stats = ServingStats()
        self._stream = torch.cuda.current_stream(d)
        self._unfinished = torch.zeros(1, dtype=torch.int32, device=d)
        self._unfinished_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self._snap_bufs = [[torch.zeros(W, dtype=torch.int32, pin_memory=True) for _ in range(3)] for _ in range(2)]
        self._snap_work = [torch.zeros(tuple(self.solver._work.shape), dtype=torch.int64, pin_memory=True)
                           for _ in range(2)] if tiles else None
        self.tiles = int(tiles)
        self._snap_flip = 0
        self._last_live: list = []
        self._launch_events: list = []
        self._copyouts: list = []

    # ------------------------------------------------------------------ admission
    @staticmethod
    def pool_bytes(worlds: int, body_capacity: int, shape_capacity: int, articulated: bool = False) -> int:
        """Resident bytes of a full pool, the default admission budget: body, shape and slot arenas,
        plus the joint, coordinate (7 per body slot) and dof (6 per body slot) arenas of an
        articulated pool."""
        b = body_capacity * _bytes_per_body() + shape_capacity * _bytes_per_shape() + worlds * _bytes_per_slot()
        if articulated:
            b += body_capacity * (_bytes_per_joint() + 7 * _bytes_per_coord() + 6 * _bytes_per_dof())
        return b

    @staticmethod
    def request_bytes(req: RolloutRequest) -> int:
        b = req.body_count * _bytes_per_body() + req.shape_count * _bytes_per_shape() + _bytes_per_slot()
        if req.articulated:
            b += (req.body_count * _bytes_per_joint() + req.coord_count * _bytes_per_coord()
                  + req.dof_count * _bytes_per_dof())
        return b

    def submit(self, reqs) -> None:
        """Queue requests (kept in arrival order)."""
        reqs = [reqs] if isinstance(reqs, RolloutRequest) else list(reqs)
        for r in reqs:
            check_request(r, self.G, self.hbm_budget_bytes)
            if r.articulated and not self.pool.is_articulated:
                raise ValueError("an articulated request needs an engine built with articulated=True")
        self.pending.extend(reqs)
        self.pending.sort(key=lambda r: r.arrival)

    def admit(self, now: float) -> list:
        """Admit arrived requests (oldest first) while a slot, arena space and the budget allow."""
        batch = []
        while self.pending and self.pending[0].arrival <= now and self._free_slots:
            r = self.pending[0]
            need = self.request_bytes(r)
            if self.resident_bytes + need > self.hbm_budget_bytes:
                break
            b0 = self._bodies.alloc(r.body_count)
            if b0 is None:
                break
            s0 = self._shapes.alloc(r.shape_count)
            if s0 is None:
                self._bodies.release(b0, r.body_count)
                break
            c0 = d0 = 0
            if r.articulated:
                c0 = self._coords.alloc(r.coord_count)
                d0 = self._dofs.alloc(r.dof_count) if c0 is not None else None
                if c0 is None or d0 is None:
                    if c0 is not None:
                        self._coords.release(c0, r.coord_count)
                    self._bodies.release(b0, r.body_count)
                    self._shapes.release(s0, r.shape_count)
                    break
            self.pending.pop(0)
            r.slot, r.b0, r.s0, r.c0, r.d0 = self._free_slots.pop(), b0, s0, c0, d0
            r.admitted_at = now
            self.resident_bytes += need
            self.resident[r.slot] = r
            batch.append(r)
        if batch:
            self._upload(batch)
        return batch

    def _upload(self, batch) -> None:
        P, d = self.pool, self.pool.device
        bidx = np.concatenate([np.arange(r.b0, r.b0 + r.body_count) for r in batch]).astype(np.int64)
        sidx = np.concatenate([np.arange(r.s0, r.s0 + r.shape_count) for r in batch]).astype(np.int64)
        bi = _dev(bidx, torch.int64, d)
        si = _dev(sidx, torch.int64, d)
        for k in _BODY_FIELDS:
            vals = np.concatenate([r.bodies[k] for r in batch])
            dst = getattr(P, k)
            dst.index_copy_(0, bi, _dev(vals, torch.float32, d).reshape((-1,) + tuple(dst.shape[1:])))
        P.body_f.index_fill_(0, bi, 0.0)
        # FP64 pose shadow = the float32 pose, as a first boundary call would seed it
        self.solver._q64.index_copy_(0, bi, P.body_q.index_select(0, bi).double())
        for k in _SHAPE_FIELDS:
            vals = np.concatenate([r.shapes[k] for r in batch])
            if k == "shape_body":
                vals = np.concatenate([np.where(r.shapes[k] >= 0, r.shapes[k] + r.b0, -1) for r in batch])
            dt = torch.int32 if k in ("shape_type", "shape_body", "shape_collision_group") else torch.float32
            dst = getattr(P, k)
            dst.index_copy_(0, si, _dev(vals, dt, d).reshape((-1,) + tuple(dst.shape[1:])))
        P.shape_world.index_copy_(0, si, _dev(np.concatenate([np.full(r.shape_count, r.slot) for r in batch]),
                                             torch.int32, d))
        if P.is_articulated:
            self._upload_joints(batch, bi)
        slots = _dev([r.slot for r in batch], torch.int64, d)

        def put(t, vals, dtype):
            t.index_copy_(0, slots, _dev(np.asarray(vals), dtype, d))

        put(P.world_body_start, [r.b0 for r in batch], torch.int32)
        put(P.world_body_end, [r.b0 + r.body_count for r in batch], torch.int32)
        put(P.world_shape_start, [r.s0 for r in batch], torch.int32)
        put(P.world_shape_end, [r.s0 + r.shape_count for r in batch], torch.int32)
        put(self.horizon_frames, [r.horizon_frames if r.horizon_frames is not None else _I32_MAX for r in batch],
            torch.int32)
        put(self.horizon_accepted,
            [r.horizon_accepted if r.horizon_accepted is not None else _I32_MAX for r in batch], torch.int32)
        put(self.world_tol, [r.tol if r.tol is not None else self.default_tol for r in batch], torch.float64)
        zeros = [0] * len(batch)
        put(self.frames_done, zeros, torch.int32)
        put(self.accepted, zeros, torch.int32)
        put(self.status, zeros, torch.int32)
        self.solver._work.index_fill_(0, slots, 0)  # the slot's work counters count this request only
        mask = torch.zeros(P.world_count, dtype=torch.uint8, device=d)
        mask.index_fill_(0, slots, 1)
        self._reset_mask = mask  # kept alive until the asynchronous reset has run
        _lib.check(_lib.lib().na_reset(self.solver._h, mask.data_ptr(), _lib.stream_handle()), "na_reset")

    def _upload_joints(self, batch, bi) -> None:
        """Joint arrays of the batch's body slots (type 0 for free-body worlds), the articulated
        worlds' per-dof arrays, coordinates, velocities and their FP64 shadows."""
        P, d = self.pool, self.pool.device
        jt = np.concatenate([r.joints["joint_type"] if r.articulated else np.zeros(r.body_count, np.int32)
                             for r in batch]).astype(np.int32)
        P.joint_type.index_copy_(0, bi, _dev(jt, torch.int32, d))
        art = [r for r in batch if r.articulated]
        if not art:
            return
        ab = _dev(np.concatenate([np.arange(r.b0, r.b0 + r.body_count) for r in art]), torch.int64, d)
        par = np.concatenate([np.where(r.joints["joint_parent"] >= 0, r.joints["joint_parent"] + r.b0, -1)
                              for r in art])
        P.joint_parent.index_copy_(0, ab, _dev(par, torch.int32, d))
        P.joint_q_start.index_copy_(0, ab, _dev(np.concatenate([r.joints["joint_q_start"] + r.c0 for r in art]),
                                                torch.int32, d))
        P.joint_qd_start.index_copy_(0, ab, _dev(np.concatenate([r.joints["joint_qd_start"] + r.d0 for r in art]),
                                                 torch.int32, d))
        for k in ("joint_X_p", "joint_X_c", "joint_axis"):
            dst = getattr(P, k)
            dst.index_copy_(0, ab, _dev(np.concatenate([r.joints[k] for r in art]), torch.float32, d).reshape(
                (-1,) + tuple(dst.shape[1:])))
        di = _dev(np.concatenate([np.arange(r.d0, r.d0 + r.dof_count) for r in art]), torch.int64, d)
        ci = _dev(np.concatenate([np.arange(r.c0, r.c0 + r.coord_count) for r in art]), torch.int64, d)
        for k in _JOINT_DOF_FIELDS:
            getattr(P, k).index_copy_(0, di, _dev(np.concatenate([r.joints[k] for r in art]), torch.float32, d))
        P.joint_q.index_copy_(0, ci, _dev(np.concatenate([r.joints["joint_q"] for r in art]), torch.float32, d))
        P.joint_qd.index_copy_(0, di, _dev(np.concatenate([r.joints["joint_qd"] for r in art]), torch.float32, d))
        P.joint_q64.index_copy_(0, ci, P.joint_q.index_select(0, ci).double())
        P.joint_qd64.index_copy_(0, di, P.joint_qd.index_select(0, di).double())

    # ------------------------------------------------------------------ rounds
    def launch(self, max_iterations: int = 0, unfinished: torch.Tensor | None = None) -> int:
        """Enqueue one persistent launch over the resident, unfinished worlds (oldest first).
        ``max_iterations`` > 0 caps the iterations per world (1: lock step)."""
        # oldest first; among requests that arrived together the largest worlds first, so the
        # longest rollouts start early instead of forming the launch's tail (the order decides only
        # which wave runs a world, never its arithmetic)
        live = sorted((r for r in self.resident.values() if r.status == 0),
                      key=lambda r: (r.arrival, -(r.body_count + r.shape_count), r.rid))
        if not live:
            return 0
        if self.tiles:
            # the free-body worlds with the most estimated work left go first and run as two-wave
            # tiles (before any estimate exists: the largest); the tiles only change which waves run
            # a world
            free = [r for r in live if not r.articulated]
            known = [r for r in free if r.cost_estimate == r.cost_estimate]
            top = sorted(known, key=lambda r: (-r.cost_estimate, r.rid))[:self.tiles] if known else free[:self.tiles]
            ids = {id(r) for r in top}
            live = top + [r for r in live if id(r) not in ids]
            self.solver.set_tiles(len(top))
        P = self.pool
        L = _lib.lib()
        budget = self.budget_ms if max_iterations <= 0 else 0.0
        flag = unfinished.data_ptr() if unfinished is not None else None
        if P.is_articulated and self.split_kinds:
            self._last_live = live
            t_start = self._timed_event()
            self._launch_split(live, budget, max_iterations, unfinished)
            self._launch_events.append((t_start, self._timed_event()))
            self.stats.rounds += 1
            return len(live)
        self._active = _dev([r.slot for r in live], torch.int32, P.device)
        self._last_live = live
        t_start = self._timed_event()
        if P.is_articulated:
            # one launch over both kinds: every wave takes the next world of either kind
            _lib.check(L.na_serve_mixed(
                self.solver._h, self._active.data_ptr(), len(live), self.horizon_frames.data_ptr(),
                self.horizon_accepted.data_ptr(), self.frames_done.data_ptr(), self.accepted.data_ptr(),
                self.status.data_ptr(), self.world_tol.data_ptr(), P.body_q.data_ptr(), P.body_qd.data_ptr(),
                P.body_f.data_ptr(), P.joint_q.data_ptr(), P.joint_qd.data_ptr(), P.joint_q64.data_ptr(),
                P.joint_qd64.data_ptr(), self.dt_outer, budget, int(max_iterations), flag, _lib.stream_handle()),
                "na_serve_mixed")
        else:
            _lib.check(L.na_serve(self.solver._h, self._active.data_ptr(), len(live), self.horizon_frames.data_ptr(),
                                  self.horizon_accepted.data_ptr(), self.frames_done.data_ptr(),
                                  self.accepted.data_ptr(), self.status.data_ptr(), self.world_tol.data_ptr(),
                                  P.body_q.data_ptr(), P.body_qd.data_ptr(), P.body_f.data_ptr(), self.dt_outer,
                                  budget, int(max_iterations), flag, _lib.stream_handle()), "na_serve")
        self._launch_events.append((t_start, self._timed_event()))
        self.stats.rounds += 1
        return len(live)

    def _launch_split(self, live, budget, max_iterations, unfinished) -> None:
        """One launch per kind on the same stream (na_serve, then na_serve_articulated): the two
        kinds take turns instead of sharing the device."""
        P, L = self.pool, _lib.lib()
        flag = unfinished.data_ptr() if unfinished is not None else None
        free = [r for r in live if not r.articulated]
        art = [r for r in live if r.articulated]
        if free:
            self._active = _dev([r.slot for r in free], torch.int32, P.device)
            _lib.check(L.na_serve(self.solver._h, self._active.data_ptr(), len(free), self.horizon_frames.data_ptr(),
                                  self.horizon_accepted.data_ptr(), self.frames_done.data_ptr(),
                                  self.accepted.data_ptr(), self.status.data_ptr(), self.world_tol.data_ptr(),
                                  P.body_q.data_ptr(), P.body_qd.data_ptr(), P.body_f.data_ptr(), self.dt_outer,
                                  budget, int(max_iterations), flag, _lib.stream_handle()), "na_serve")
        if art:
            self._active_art = _dev([r.slot for r in art], torch.int32, P.device)
            if free and unfinished is not None:  # the second launch starts by clearing the flag
                self._flag_free = unfinished.clone()
            _lib.check(L.na_serve_articulated(
                self.solver._h, self._active_art.data_ptr(), len(art), self.horizon_frames.data_ptr(),
                self.horizon_accepted.data_ptr(), self.frames_done.data_ptr(), self.accepted.data_ptr(),
                self.status.data_ptr(), self.world_tol.data_ptr(), P.joint_q.data_ptr(), P.joint_qd.data_ptr(),
                P.joint_q64.data_ptr(), P.joint_qd64.data_ptr(), P.body_q.data_ptr(), P.body_qd.data_ptr(),
                self.dt_outer, budget, int(max_iterations), flag, _lib.stream_handle()), "na_serve_articulated")
            if free and unfinished is not None:
                unfinished.bitwise_or_(self._flag_free)

    def _timed_event(self) -> torch.cuda.Event:
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self._stream)
        return ev

    def _account_launches(self) -> None:
        """Add the device time of every launch whose end event has completed to ``stats``."""
        keep = []
        for a, b in self._launch_events:
            if b.query():
                self.stats.device_s += a.elapsed_time(b) * 1e-3
            else:
                keep.append((a, b))
        self._launch_events = keep

    def _snapshot(self) -> _Snapshot:
        """Enqueue copies of the per-world status, accepted and frame counters of the launch just
        enqueued into pinned host buffers (two sets, used alternately) and an event after them."""
        bufs = self._snap_bufs[self._snap_flip]
        work = self._snap_work[self._snap_flip] if self._snap_work is not None else None
        self._snap_flip ^= 1
        for host, dev in zip(bufs, (self.status, self.accepted, self.frames_done)):
            host.copy_(dev, non_blocking=True)
        if work is not None:
            work.copy_(self.solver._work, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(self._stream)
        return _Snapshot(ev, bufs, list(self._last_live), work)

    def _retire_snapshot(self, snap: _Snapshot, now: float) -> list:
        """Retire the requests of ``snap``'s launch that had finished by its end.  Their final
        states are gathered on the stream behind the launches already enqueued (which skip
        finished worlds, so the states are final) into pinned buffers; the slots and arena ranges
        are freed at once, since any later upload into them is enqueued behind the gathers."""
        status, acc, frames = (b.numpy() for b in snap.bufs)
        if snap.work is not None:
            self._estimate(snap, status, acc, frames)
        done = [r for r in snap.live if status[r.slot] != 0 and self.resident.get(r.slot) is r]
        if not done:
            return []
        for r in done:
            del self.resident[r.slot]
        d = self.pool.device
        bidx = _dev(np.concatenate([np.arange(r.b0, r.b0 + r.body_count) for r in done]), torch.int64, d)

        def gather(src, idx):
            host = torch.empty((idx.numel(),) + tuple(src.shape[1:]), dtype=src.dtype, pin_memory=True)
            host.copy_(src.index_select(0, idx), non_blocking=True)
            return host

        q, qd = gather(self.pool.body_q, bidx), gather(self.pool.body_qd, bidx)
        jq = jqd = None
        art = [r for r in done if r.articulated]
        if art:
            ci = _dev(np.concatenate([np.arange(r.c0, r.c0 + r.coord_count) for r in art]), torch.int64, d)
            di = _dev(np.concatenate([np.arange(r.d0, r.d0 + r.dof_count) for r in art]), torch.int64, d)
            jq, jqd = gather(self.pool.joint_q, ci), gather(self.pool.joint_qd, di)
        ev = torch.cuda.Event()
        ev.record(self._stream)
        self._copyouts.append((ev, done, q, qd, jq, jqd))
        for r in done:
            if r.articulated:
                self._coords.release(r.c0, r.coord_count)
                self._dofs.release(r.d0, r.dof_count)
            r.status, r.accepted, r.frames = int(status[r.slot]), int(acc[r.slot]), int(frames[r.slot])
            r.finished_at = now
            self.resident_bytes -= self.request_bytes(r)
            self._bodies.release(r.b0, r.body_count)
            self._shapes.release(r.s0, r.shape_count)
            self._free_slots.append(r.slot)
            self.stats.accepted_steps += r.accepted
            if r.status == 1:
                self.stats.completed += 1
                self.stats.latencies.append(r.latency)
            else:
                self.stats.failed += 1
            self.finished.append(r)
        return done

    def _estimate(self, snap: _Snapshot, status, acc, frames) -> None:
        """Remaining work of each running request of ``snap``'s launch, from its measured cost so far
        (PCG iterations + 5 per Newton iteration + 10 per attempt, the relative costs of those phases)
        per frame (or per accepted step, for a step horizon) times what is left of its horizon."""
        w = snap.work.numpy()
        for r in snap.live:
            if status[r.slot] != 0 or self.resident.get(r.slot) is not r:
                continue
            c = float(w[r.slot, 2] + 5 * w[r.slot, 1] + 10 * w[r.slot, 0])
            if r.horizon_frames is not None:
                done, left = int(frames[r.slot]), r.horizon_frames - int(frames[r.slot])
            else:
                done, left = int(acc[r.slot]), r.horizon_accepted - int(acc[r.slot])
            r.cost_estimate = c / max(done, 1) * max(left, 0)

    def _collect_copyouts(self, block: bool) -> None:
        """Hand the gathered final states to their requests once the gathers have run."""
        keep = []
        for ev, done, q, qd, jq, jqd in self._copyouts:
            if block:
                ev.synchronize()
            elif not ev.query():
                keep.append((ev, done, q, qd, jq, jqd))
                continue
            q, qd = q.numpy(), qd.numpy()
            jq = jq.numpy() if jq is not None else None
            jqd = jqd.numpy() if jqd is not None else None
            o = oc = od = 0
            for r in done:
                n = r.body_count
                r.body_q, r.body_qd = q[o:o + n].copy(), qd[o:o + n].copy()
                o += n
                if r.articulated:
                    r.joint_q = jq[oc:oc + r.coord_count].copy()
                    r.joint_qd = jqd[od:od + r.dof_count].copy()
                    oc += r.coord_count
                    od += r.dof_count
        self._copyouts = keep

    def retire(self, now: float) -> list:
        """Wait for the enqueued work, then retire every finished resident request: final state
        copied out, latency recorded, slot freed."""
        self._stream.synchronize()
        if not self.resident:
            return []
        self._last_live = list(self.resident.values())
        snap = self._snapshot()
        snap.event.synchronize()
        done = self._retire_snapshot(snap, now)
        self._collect_copyouts(block=True)
        self._account_launches()
        return done

    def run(self, requests=None, *, realtime: bool = True, max_wall_s: float | None = None,
            pipelined: bool = True) -> dict:
        """Serve ``requests`` (plus anything already queued) to completion.

        With ``realtime`` the requests' ``arrival`` times are seconds after the start of this call
        and requests are admitted only once they have arrived; otherwise all are admitted at
        once.  Returns :meth:`ServingStats.summary`.

        Launches are pipelined: the next round's launch is enqueued before the host waits for the
        previous one, so the device never waits for the host's retirement and admission work.
        A launch lists every resident world the host has not yet seen finish; a world that
        finished in the launch before is skipped on the device.  With ``pipelined=False`` each
        round waits for the previous one, so its order and tiles use the latest cost estimates.
        """
        if requests is not None:
            self.submit(requests)
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0 if realtime else math.inf

        if self.schedule == "static":
            return self._run_static(t0, now, max_wall_s)

        self.admit(now())
        prev = None
        while self.pending or self.resident or prev is not None:
            snap = self._snapshot() if self.resident and self.launch() else None
            if not pipelined and snap is not None:
                prev, snap = snap, None
            if prev is not None:
                prev.event.synchronize()
                self._retire_snapshot(prev, time.perf_counter() - t0)
            self._collect_copyouts(block=False)
            self._account_launches()
            if snap is None and prev is None and not self.resident and self.pending:
                wait = self.pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
            self.admit(now())
            prev = snap
            if max_wall_s is not None and time.perf_counter() - t0 > max_wall_s:
                break
        self._stream.synchronize()
        self._collect_copyouts(block=True)
        self._account_launches()
        self.stats.wall_s = time.perf_counter() - t0
        return self.stats.summary()

    def _run_static(self, t0, now, max_wall_s) -> dict:
        iterations = 0
        while self.pending or self.resident:
            if not self.admit(now()):
                wait = self.pending[0].arrival - (time.perf_counter() - t0)
                time.sleep(min(max(wait, 0.0), 0.05))
                continue
            while True:  # one fixed batch, lock step
                self.launch(max_iterations=1, unfinished=self._unfinished)
                self._unfinished_host.copy_(self._unfinished, non_blocking=True)
                self._stream.synchronize()
                iterations += 1
                if int(self._unfinished_host[0]) == 0:
                    break
                if max_wall_s is not None and time.perf_counter() - t0 > max_wall_s:
                    break
            self.retire(time.perf_counter() - t0)
            if max_wall_s is not None and time.perf_counter() - t0 > max_wall_s:
                break
        self.stats.wall_s = time.perf_counter() - t0
        out = self.stats.summary()
        out["lockstep_iterations"] = iterations
        return out
