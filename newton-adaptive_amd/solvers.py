"""Per-world step-doubling solvers with the reference's Python surface.

``SolverSAPAdaptive`` mirrors ``newton/_src/solvers/sap/solver_sap_adaptive.py:437-1001``
and ``SolverMuJoCoAdaptive`` mirrors ``newton/_src/solvers/mujoco/solver_mujoco_adaptive.py:418-1042``
(constructor kwargs, ``step`` / ``step_dt`` / ``reset``, telemetry properties, ``ValueError``
on invalid settings, ``RuntimeError`` when the inner convex solve does not converge).  Both
run the same gfx950 path in ``libnadapt.so``: contact generation, the convex compliant
contact solve and step doubling per world inside one wavefront; they differ only in the
controller (``_adapt_dt`` with fixed/adaptive modes vs ``_calc_adjusted_step`` with the
separate commit mask and floor-divergence latch).

The controller runs in FP64 on the device; ``dt`` / ``sim_time`` / errors are float64
tensors.  Body poses accumulate in an FP64 shadow (``body_q64``) that is kept across
frames while the caller's float32 ``body_q`` still equals its rounding.
"""

from __future__ import annotations

import ctypes as C
import math
import weakref

import torch

from . import _lib
from .contacts import BROAD_PHASES, CollisionPipeline
from .native import model_handle

_MODE_CODES = {"fixed": 0, "adaptive": 1}
_LINE_SEARCHES = {"exact_root": 0, "armijo_decay": 1}
# Newton seeds of an attempt's three solves: "reference" as solver_sap_adaptive.py:756-771 (v_t,
# (v_t + v_full) / 2, v_full); "extrapolate": v_t + dt/2 a (a: the body's mean acceleration over
# its last committed step), (v_t + v_full) / 2, 2 v_mid - v_t.  Same converged solves, fewer
# Newton iterations on settled piles.
WARM_STARTS = {"reference": 0, "extrapolate": 1, "concurrent": 2}
_CONTROL_FIELDS = ("joint_f", "joint_target_q", "joint_target_qd", "joint_act")
_WORK_FIELDS = ("attempts", "newton_iterations", "cg_iterations", "contacts", "pairs", "solves",
                "cycles_collide", "cycles_setup", "cycles_residual", "cycles_preconditioner", "cycles_pcg",
                "cycles_line_search")


class _SolverBase:
    _controller = 0

    def __init__(self, model, *, mode="adaptive", tol=1e-3, dt_inner_init=0.01, dt_inner_min=1e-6,
                 dt_inner_max=None, max_substeps=16, max_rigid_contact=128, max_iterations=100,
                 contact_preset_variant="drake", line_search_variant="exact_root", contact_tau_d=0.01,
                 broad_phase="sap", optimality_rel_tol=None, schedule="served", cg_rel_tol=1.0e-4,
                 cg_forcing=0.1, world_order="cost", ls_rel_tol=1.0e-4, warm_start="extrapolate", _contact_law=0,
                 _impratio=1.0, **_kw):
        if mode not in _MODE_CODES:
            raise ValueError(f"mode must be one of {tuple(_MODE_CODES)}, got {mode!r}.")
        if float(tol) <= 0.0:
            raise ValueError(f"tol must be > 0, got {tol!r}.")
        if float(dt_inner_init) <= 0.0:
            raise ValueError(f"dt_inner_init must be > 0, got {dt_inner_init!r}.")
        if float(dt_inner_min) <= 0.0:
            raise ValueError(f"dt_inner_min must be > 0, got {dt_inner_min!r}.")
        if dt_inner_max is not None and float(dt_inner_max) <= 0.0:
            raise ValueError(f"dt_inner_max must be > 0 when provided, got {dt_inner_max!r}.")
        if int(max_substeps) < 1:
            raise ValueError(f"max_substeps must be >= 1, got {max_substeps!r}.")
        if not 1 <= int(max_rigid_contact) <= 128:
            raise ValueError(f"max_rigid_contact must be in [1, 128] per world, got {max_rigid_contact!r}.")
        if contact_preset_variant != "drake":
            raise ValueError(f"contact_preset_variant must be 'drake', got {contact_preset_variant!r}.")
        if line_search_variant not in _LINE_SEARCHES:
            raise ValueError(f"line_search_variant must be one of {tuple(_LINE_SEARCHES)}, "
                             f"got {line_search_variant!r}.")
        if broad_phase not in BROAD_PHASES:
            raise ValueError(f"broad_phase must be one of {tuple(BROAD_PHASES)}, got {broad_phase!r}.")
        if not 0.0 < float(cg_rel_tol) < 1.0:
            raise ValueError(f"cg_rel_tol must be in (0, 1), got {cg_rel_tol!r}.")
        if not 0.0 <= float(cg_forcing) < 1.0:
            raise ValueError(f"cg_forcing must be in [0, 1), got {cg_forcing!r}.")
        if not 0.0 < float(ls_rel_tol) < 1.0:
            raise ValueError(f"ls_rel_tol must be in (0, 1), got {ls_rel_tol!r}.")
        if world_order not in ("cost", "index"):
            raise ValueError(f"world_order must be 'cost' or 'index', got {world_order!r}.")
        if warm_start not in WARM_STARTS:
            raise ValueError(f"warm_start must be one of {tuple(WARM_STARTS)}, got {warm_start!r}.")
        if schedule not in ("served", "plain", "plain_batch"):
            raise ValueError(f"schedule must be 'served', 'plain' or 'plain_batch', got {schedule!r}.")
        plain_batch = schedule == "plain_batch"
        if plain_batch:
            schedule = "plain"
        self.model = model
        self._mode = mode
        self._tol = float(tol)
        self._dt_inner_init = float(dt_inner_init)
        self._dt_min = float(dt_inner_min)
        self._dt_max = float(dt_inner_max) if dt_inner_max is not None else math.inf
        self._max_substeps = int(max_substeps)
        self._max_rigid_contact = int(max_rigid_contact)
        self._schedule = schedule
        self._world_order = world_order
        self._world_count = int(model.world_count)
        W, B, dev = self._world_count, model.body_count, model.device

        cfg = _lib.NaConfig()
        cfg.mode = _MODE_CODES[mode]
        cfg.controller = self._controller
        cfg.tol = self._tol
        cfg.dt_inner_init = self._dt_inner_init
        cfg.dt_inner_min = self._dt_min
        cfg.dt_inner_max = self._dt_max if math.isfinite(self._dt_max) else 1.0e30
        cfg.divergence_threshold = 1.0e9
        cfg.max_substeps = self._max_substeps
        cfg.max_iterations = int(max_iterations)
        cfg.max_rigid_contact = self._max_rigid_contact
        cfg.broad_phase = BROAD_PHASES[broad_phase]
        # CENIC kappa*tol with kappa = 1e-3, floored at 1e-8 (solver_sap_adaptive.py:485-501)
        self.optimality_rel_tol = float(optimality_rel_tol) if optimality_rel_tol is not None \
            else max(1.0e-3 * self._tol, 1.0e-8)
        cfg.optimality_rel_tol = self.optimality_rel_tol
        cfg.optimality_abs_tol = 1.0e-12
        cfg.contact_tau_d = float(contact_tau_d)
        cfg.sigma = 1.0e-3
        cfg.beta = 1.0
        cfg.ls_rel_tol = float(ls_rel_tol)
        cfg.cg_rel_tol = float(cg_rel_tol)
        cfg.max_cg = 200
        cfg.cg_forcing = float(cg_forcing)
        cfg.line_search = _LINE_SEARCHES[line_search_variant]
        cfg.contact_law = int(_contact_law)
        cfg.impratio = float(_impratio)
        cfg.warm_start = WARM_STARTS[warm_start]
        self.warm_start = warm_start
        self.tiles = 0
        self.line_search_variant = line_search_variant
        self._cfg = cfg

        f64 = dict(dtype=torch.float64, device=dev)
        self._dt = torch.full((W,), self._dt_inner_init, **f64)
        self._dt_half = torch.full((W,), 0.5 * self._dt_inner_init, **f64)
        self._ideal_dt = torch.full((W,), self._dt_inner_init, **f64)
        self._sim_time = torch.zeros(W, **f64)
        self._next_time = torch.zeros(W, **f64)
        self._last_error = torch.zeros(W, **f64)
        self._accepted_error = torch.zeros(W, **f64)
        self._accepted = torch.zeros(W, dtype=torch.bool, device=dev)
        self._commit = torch.zeros(W, dtype=torch.bool, device=dev)
        self._diverged = torch.zeros(W, dtype=torch.bool, device=dev)
        self._substeps_frame = torch.zeros(W, dtype=torch.int32, device=dev)
        self._iterations = torch.zeros(W, dtype=torch.int32, device=dev)
        self._solve_fail = torch.zeros(W, dtype=torch.int32, device=dev)
        self._overflow = torch.zeros(W, dtype=torch.int32, device=dev)
        self._cum_accepted = torch.zeros(1, dtype=torch.int64, device=dev)
        self._work = torch.zeros((W, len(_WORK_FIELDS)), dtype=torch.int64, device=dev)
        self._q64 = torch.zeros((max(B, 1), 7), **f64)
        # each body's mean acceleration over its last committed step: the full step's Newton seed
        # under warm_start "extrapolate" (zero: no history yet)
        self._trend = torch.zeros((max(B, 1), 6), dtype=torch.float32, device=dev)
        self._last_state = None
        self._full_world_mask = torch.ones(W, dtype=torch.bool, device=dev)

        t = _lib.NaTelemetry()
        t.dt, t.dt_half, t.ideal_dt = self._dt.data_ptr(), self._dt_half.data_ptr(), self._ideal_dt.data_ptr()
        t.sim_time, t.next_time = self._sim_time.data_ptr(), self._next_time.data_ptr()
        t.last_raw_error, t.last_error = self._last_error.data_ptr(), self._accepted_error.data_ptr()
        t.accepted, t.commit, t.diverged = (self._accepted.data_ptr(), self._commit.data_ptr(),
                                            self._diverged.data_ptr())
        t.substeps, t.iterations = self._substeps_frame.data_ptr(), self._iterations.data_ptr()
        t.solve_fail, t.overflow = self._solve_fail.data_ptr(), self._overflow.data_ptr()
        t.cum_accepted, t.work, t.body_q64 = (self._cum_accepted.data_ptr(), self._work.data_ptr(),
                                              self._q64.data_ptr())
        t.body_trend = self._trend.data_ptr()
        self._handle = model_handle(model, with_pairs=broad_phase == "explicit")
        h = C.c_void_p()
        _lib.check(_lib.lib().na_solver_create(self._handle.handle, C.byref(cfg), C.byref(t), C.byref(h)),
                   "na_solver_create")
        self._h = h
        if plain_batch:
            _lib.check(_lib.lib().na_solver_set_plain_batch(self._h, 1), "na_solver_set_plain_batch")
        self.plain_batch = plain_batch
        self._order = None
        self._work_prev = None
        self._pipeline = None
        self._contacts = None
        self.last_plain_iterations = 0

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                _lib.lib().na_solver_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ------------------------------------------------------------------ schedule
    def set_schedule(self, grid_waves: int = 0, order: torch.Tensor | None = None, _checked: bool = False) -> None:
        """Persistent grid size (0: one wave per resident slot) and world visiting order.  ``order``
        must be a permutation of ``range(world_count)``: a repeated world would run on two waves at
        once and another world would never step."""
        if order is not None:
            order = order.to(device=self.model.device, dtype=torch.int32).contiguous()
            if order.numel() != self._world_count:
                raise ValueError("order must be a permutation of the worlds")
            if not _checked:
                ref = torch.arange(self._world_count, dtype=torch.int32, device=order.device)
                if not torch.equal(torch.sort(order).values, ref):
                    raise ValueError("order must be a permutation of the worlds (each index once)")
        self._order = order
        _lib.check(_lib.lib().na_solver_set_schedule(self._h, int(grid_waves),
                                                     order.data_ptr() if order is not None else None),
                   "na_solver_set_schedule")

    def set_tiles(self, tiles: int) -> None:
        """Run the first ``tiles`` worlds of each boundary call's visiting order (under
        ``world_order="cost"``, the previous call's costliest worlds) and of each served launch's
        list (``na_serve``, the serving engine) as two-wave tiles: two wavefronts per world, one
        solving each attempt's full step while the other solves its first half step.  Needs
        ``warm_start="concurrent"`` (the first half step's Newton seed must not depend on the full
        step's result).  Tiles apply to free-body worlds (a serving pool's articulated worlds and
        articulated boundary calls run one world per wave).  A world's results are the same with
        and without tiles, bit for bit."""
        tiles = int(tiles)
        if tiles < 0:
            raise ValueError("tiles must be >= 0")
        if tiles and self.warm_start != "concurrent":
            raise ValueError("two-wave tiles need warm_start='concurrent'")
        _lib.check(_lib.lib().na_solver_set_tiles(self._h, tiles), "na_solver_set_tiles")
        self.tiles = tiles

    # ------------------------------------------------------------------ boundary call
    def _control_desc(self, control):
        """``na_control_desc`` of a :class:`Control` (None: no control).  Every array must hold
        one float per dof; the tensors are kept alive until the next call."""
        if control is None:
            self._control_keep = None
            return None
        m = self.model
        D = int(m.joint_dof_count)
        desc = _lib.NaControlDesc()
        keep = []
        for name in _CONTROL_FIELDS:
            t = getattr(control, name, None)
            if t is None:
                continue
            t = torch.as_tensor(t).to(device=m.device, dtype=torch.float32).contiguous().view(-1)
            if t.numel() != D:
                raise ValueError(f"control.{name} must hold one value per dof ({D}), got {t.numel()}")
            keep.append(t)
            setattr(desc, name, t.data_ptr())
        self._control_keep = (desc, keep)
        return C.byref(desc)

    def _body_f(self, state):
        bf = getattr(state, "body_f", None)
        if bf is None:
            return None
        if bf.dtype != torch.float32 or not bf.is_contiguous() or bf.numel() != 6 * self.model.body_count:
            raise ValueError("state.body_f must be a contiguous float32 [body_count, 6] tensor")
        return bf

    def integrate(self, state, control, dt_outer: float, stream=None):
        s = _lib.stream_handle(stream)
        bf = self._body_f(state)
        ctl = self._control_desc(control)
        if self.model.is_articulated:
            # generalized coordinates: joint_q / joint_qd advance, body_q / body_qd follow (eval_fk)
            _lib.check(_lib.lib().na_step_articulated(self._h, state.joint_q.data_ptr(), state.joint_qd.data_ptr(),
                                                      state.body_q.data_ptr(), state.body_qd.data_ptr(),
                                                      bf.data_ptr() if bf is not None else None, ctl,
                                                      float(dt_outer), s), "na_step_articulated")
            if self._world_order == "cost":
                self._order_by_cost()
            if int(self._solve_fail.max().item()) != 0:
                self._raise_solve_failure()
            return state
        if self._schedule == "served":
            st = _lib.check(_lib.lib().na_step(self._h, state.body_q.data_ptr(), state.body_qd.data_ptr(),
                                               bf.data_ptr() if bf is not None else None, ctl, float(dt_outer), s),
                            "na_step")
            if self._world_order == "cost":
                self._order_by_cost()
            # the served launch ends with per-world solve status on the device
            if int(self._solve_fail.max().item()) != 0:
                self._raise_solve_failure()
        else:
            it = C.c_int32(0)
            st = _lib.check(_lib.lib().na_step_plain(self._h, state.body_q.data_ptr(), state.body_qd.data_ptr(),
                                                     bf.data_ptr() if bf is not None else None, ctl,
                                                     float(dt_outer), C.byref(it), s), "na_step_plain")
            self.last_plain_iterations = it.value
            if st == 2:
                self._raise_solve_failure()
        return state

    def _order_by_cost(self):
        """Visit the most expensive worlds of the last boundary call first in the next one
        (longest-first list scheduling over the persistent grid's atomic world queue).  The cost
        of a world is its PCG iterations plus 5 per Newton iteration and 10 per attempt, the
        measured relative costs of those phases.  The order only changes which wave runs a
        world, never its arithmetic."""
        w = self._work
        if self._work_prev is None or self._work_prev.shape != w.shape:
            self._work_prev = w.clone()
            return
        d = w - self._work_prev
        self._work_prev.copy_(w)
        cost = d[:, 2] + 5 * d[:, 1] + 10 * d[:, 0]
        order = torch.argsort(cost, descending=True, stable=True).to(torch.int32)
        self.set_schedule(0, order, _checked=True)  # an argsort is a permutation

    def _raise_solve_failure(self):
        raise RuntimeError(
            f"{type(self).__name__} inner convex solve failed to converge to "
            f"optimality_rel_tol={self.optimality_rel_tol:.3e}.")

    def substep(self, state_in, state_out, control, contacts, dt, guess=None):
        """ONE inner physics step of every world at its own ``dt`` (``solver_sap_adaptive.py:714-729``):
        contacts at ``state_in``'s pose, one convex solve, FP64 integration into ``state_out``; no
        step doubling, no controller.  ``dt``: per-world tensor or scalar [s].  ``guess``: an explicit
        per-body velocity seed [B, 6] (``None``: start from ``state_in``'s velocity); a seed that
        already meets the optimality test costs no Newton iteration.  The converged FP64 velocity is
        kept in ``last_substep_velocity``.  Returns the per-world stats [W, 6] (Newton iterations,
        CG iterations, converged, contacts, exit reason, residual bits) for free bodies, [W, 4]
        (Newton iterations, factorizations, converged, contacts) for articulated models."""
        m = self.model
        W, dev = m.world_count, m.device
        h = dt.to(device=dev, dtype=torch.float64).contiguous() if torch.is_tensor(dt) \
            else torch.full((W,), float(dt), dtype=torch.float64, device=dev)
        if h.numel() != W:
            raise ValueError(f"dt must hold one step per world ({W}), got {h.numel()}")
        s = _lib.stream_handle()
        if m.is_articulated:
            if guess is not None:
                raise ValueError("an explicit velocity guess is not supported for articulated models")
            q = state_in.joint_q.to(torch.float64).contiguous()
            v = state_in.joint_qd.to(torch.float64).contiguous()
            qo, vo = torch.empty_like(q), torch.empty_like(v)
            stats = torch.zeros((W, 4), dtype=torch.int32, device=dev)
            bf = self._body_f(state_in)
            _lib.check(_lib.lib().na_art_substep(self._h, q.data_ptr(), v.data_ptr(), h.data_ptr(),
                                                 bf.data_ptr() if bf is not None else None,
                                                 self._control_desc(control), qo.data_ptr(), vo.data_ptr(),
                                                 stats.data_ptr(), s), "na_art_substep")
            state_out.joint_q.copy_(qo)
            state_out.joint_qd.copy_(vo)
            self.last_substep_velocity = vo
            return stats
        B = m.body_count
        q64 = state_in.body_q.to(torch.float64).contiguous()
        qd = state_in.body_qd.contiguous()
        bf = getattr(state_in, "body_f", None)
        f = bf.contiguous() if bf is not None else torch.zeros_like(qd)
        g = None
        if guess is not None:
            g = torch.as_tensor(guess, dtype=torch.float64, device=dev).reshape(B, 6).contiguous()
        qo = torch.empty((B, 7), dtype=torch.float64, device=dev)
        vo = torch.empty((B, 6), dtype=torch.float32, device=dev)
        v64 = torch.empty((B, 6), dtype=torch.float64, device=dev)
        stats = torch.zeros((W, 6), dtype=torch.int32, device=dev)
        _lib.check(_lib.lib().na_substep(self._h, q64.data_ptr(), qd.data_ptr(), f.data_ptr(), h.data_ptr(),
                                         g.data_ptr() if g is not None else None, qo.data_ptr(), vo.data_ptr(),
                                         v64.data_ptr(), stats.data_ptr(), s), "na_substep")
        state_out.body_q.copy_(qo)
        state_out.body_qd.copy_(vo)
        self.last_substep_velocity = v64
        return stats

    def step(self, state_in, state_out, control, contacts=None, dt=None, apply_forces=None):
        """Advance every world by ``dt`` (= ``dt_outer``) in place; returns ``(state_in, state_out)``
        (``solver_sap_adaptive.py:962-975``)."""
        if dt is None:
            raise ValueError(f"{type(self).__name__}.step requires dt (the outer boundary period).")
        if apply_forces is not None:
            apply_forces(state_in)
        last = self._last_state() if self._last_state is not None else None
        if last is not state_in:
            # a state this solver did not advance last: its step history does not apply
            self._trend.zero_()
        self.integrate(state_in, control, float(dt))
        self._last_state = weakref.ref(state_in)
        return state_in, state_out

    def step_dt(self, dt_outer, state_0, state_1, control, apply_forces=None):
        """Legacy ``(dt, s0, s1, control)`` order (``solver_sap_adaptive.py:977-979``)."""
        return self.step(state_0, state_1, control, None, dt_outer, apply_forces=apply_forces)

    def reset(self, state=None, world_mask: torch.Tensor | None = None, flags=0) -> None:
        """Restore per-world controller state for masked worlds (``_reset_worlds``)."""
        mask = None
        if world_mask is not None:
            mask = world_mask.to(device=self.model.device, dtype=torch.uint8).contiguous()
            if mask.numel() != self._world_count:
                raise ValueError("world_mask must have one entry per world")
        _lib.check(_lib.lib().na_reset(self._h, mask.data_ptr() if mask is not None else None,
                                       _lib.stream_handle()), "na_reset")  # also clears their trend

    # ------------------------------------------------------------------ telemetry
    @property
    def diverged(self):
        return self._diverged

    @property
    def dt(self):
        return self._dt

    @property
    def sim_time(self):
        return self._sim_time

    @property
    def last_error(self):
        return self._accepted_error

    @property
    def last_raw_error(self):
        return self._last_error

    @property
    def accepted(self):
        return self._accepted

    @property
    def substeps(self):
        return self._substeps_frame

    @property
    def iteration_count(self):
        return self._iterations.max().view(1)

    @property
    def mode(self) -> str:
        return self._mode

    @property
    def tiling(self) -> str:
        return "adaptive"

    @property
    def body_q64(self):
        return self._q64

    @property
    def contacts(self):
        """Contacts at the working state (generated on request; the step keeps its own in LDS)."""
        return self._contacts

    def collide(self, state):
        if self._pipeline is None:
            self._pipeline = CollisionPipeline(self.model, rigid_contact_max_per_world=self._max_rigid_contact)
            self._contacts = self._pipeline.contacts()
        return self._pipeline.collide(state, self._contacts)

    def work_counters(self) -> dict:
        w = self._work.sum(dim=0).tolist()
        return dict(zip(_WORK_FIELDS, w))

    # ------------------------------------------------------------------ autotuner
    _TUNE_STATE = ("_dt", "_dt_half", "_ideal_dt", "_sim_time", "_next_time", "_last_error", "_accepted_error",
                   "_accepted", "_commit", "_diverged", "_substeps_frame", "_iterations", "_solve_fail", "_overflow",
                   "_cum_accepted", "_work", "_q64", "_trend")

    def autotune(self, state, dt_outer: float, frames: int = 2, waves_per_cu=(2, 3, 4),
                 orders=("cost", "size", "index"), control=None, tiles=None) -> dict:
        """Pick the persistent launch schedule of the boundary call by timing it on the device.

        Each candidate (waves per CU of the persistent grid, world visiting order) runs ``frames``
        boundary calls from the same snapshot of ``state`` and of the controller; the device time
        of those calls decides.  Orders: "cost" (the worlds' work in an untimed probe run, most
        expensive first), "size" (bodies + shapes, largest first), "index".  A schedule decides
        only which wave runs a world, never its arithmetic, so every candidate must leave the same
        state bit for bit; that is checked, and a mismatch raises.  Under ``warm_start="concurrent"``
        the candidates also include two-wave tiles for the first k worlds of the cost order, k in
        ``tiles`` (default 64, 128, 256; see :meth:`set_tiles`).  The state and controller are
        restored afterwards, the chosen schedule is kept for the following calls (``world_order``
        becomes "index" so that the order is not replaced after the next call) and the table is
        returned and kept in ``schedule_choice``."""
        if self._schedule != "served":
            raise ValueError("autotune applies to the served (persistent launch) schedule")
        m = self.model
        if m.is_articulated:
            raise ValueError("autotune is built for free-body models (their whole step state is restorable)")
        dev = m.device
        snap = {k: getattr(self, k).clone() for k in self._TUNE_STATE}
        q0 = state.body_q.clone()
        qd0 = state.body_qd.clone()

        def restore():
            for k, v in snap.items():
                getattr(self, k).copy_(v)
            state.body_q.copy_(q0)
            state.body_qd.copy_(qd0)

        saved_order = self._world_order
        self._world_order = "index"
        # untimed probe: per-world work for the cost order
        self.set_schedule(0, None)
        w_before = self._work.clone()
        for _ in range(frames):
            self.step(state, state, control, None, dt_outer)
        cost_w = self._work - w_before
        restore()
        cost = (cost_w[:, 2] + 5 * cost_w[:, 1] + 10 * cost_w[:, 0]).to(torch.float64)
        wb = m.world_body_start.to(torch.int64)
        ws = m.world_shape_start.to(torch.int64)
        size = (wb[1:] - wb[:-1] + ws[1:] - ws[:-1]).to(torch.float64)
        order_of = {"cost": torch.argsort(cost, descending=True, stable=True),
                    "size": torch.argsort(size, descending=True, stable=True), "index": None}
        cus = torch.cuda.get_device_properties(dev).multi_processor_count
        stream = torch.cuda.current_stream(dev)
        if tiles is None:
            tiles = (64, 128, 256) if self.warm_start == "concurrent" else ()
        cands = [(int(w), name, 0) for w in waves_per_cu for name in orders]
        cands += [(4, "cost", int(k)) for k in tiles if k > 0]
        rows, ref = [], None
        for wpc, name, k in cands:
            order = order_of[name]
            self.set_schedule(int(wpc) * cus, order.to(torch.int32) if order is not None else None,
                              _checked=True)
            self.set_tiles(k)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(frames):
                self.step(state, state, control, None, dt_outer)
            e1.record(stream)
            torch.cuda.synchronize(dev)
            out = (state.body_q.clone(), state.body_qd.clone(), self._q64.clone())
            if ref is None:
                ref = out
            elif not all(torch.equal(a, b) for a, b in zip(out, ref)):
                restore()
                raise RuntimeError("autotune: schedules gave different states (a schedule changed arithmetic)")
            rows.append({"waves_per_cu": int(wpc), "order": name, "tiles": k, "ms": e0.elapsed_time(e1)})
            restore()
        best = min(rows, key=lambda r: r["ms"])
        order = order_of[best["order"]]
        self.set_schedule(best["waves_per_cu"] * cus, order.to(torch.int32) if order is not None else None,
                          _checked=True)
        self.set_tiles(best["tiles"])
        if best["order"] != "index" or saved_order == "index":
            self._world_order = "index"
        else:
            self._world_order = saved_order
        self.schedule_choice = {"best": best, "candidates": rows, "frames": int(frames)}
        return self.schedule_choice

    def reset_work_counters(self) -> None:
        self._work.zero_()
        if self._work_prev is not None:
            self._work_prev.zero_()

    def cumulative_substeps(self) -> int:
        """Accepted inner steps times the 3 evaluations of one step-doubling attempt."""
        return int(self._cum_accepted.item()) * 3

    def accepted_steps(self) -> int:
        return int(self._cum_accepted.item())

    def reset_compute_counter(self) -> None:
        self._cum_accepted.zero_()

    def get_max_contact_count(self) -> int:
        return self._max_rigid_contact * self._world_count

    def get_status_summary(self) -> dict:
        """Six-scalar summary (``solver_mujoco_adaptive.py:1021-1042``)."""
        return {
            "sim_time_min": float(self._sim_time.min()),
            "sim_time_max": float(self._sim_time.max()),
            "error_max": float(self._accepted_error.max()),
            "accept_count": int(self._accepted.sum()),
            "dt_min": float(self._dt.min()),
            "dt_max": float(self._dt.max()),
        }

    def update_contacts(self, contacts, state=None) -> None:
        return None

    def notify_model_changed(self, flags: int) -> None:
        return None


class SolverSAPAdaptive(_SolverBase):
    """Per-world adaptive (step-doubling) convex-contact integrator (``solver_sap_adaptive.py:437``)."""

    _controller = 0


_MJ_SOLVERS = {"newton": 2, "cg": 1}       # mjtSolver: both minimise the same convex problem
_MJ_INTEGRATORS = {"euler": 0, "implicitfast": 3}  # mjtIntegrator
_MJ_CONES = {"pyramidal": 0}              # mjtCone


def _mj_option(value, table, name, default):
    """A MuJoCo enum option given by name or by its MuJoCo integer constant."""
    if value is None:
        return default
    if isinstance(value, str):
        key = value.lower()
        if key in table:
            return key
    else:
        inv = {v: k for k, v in table.items()}
        if int(value) in inv:
            return inv[int(value)]
    raise ValueError(f"{name} must be one of {tuple(table)} (or their MuJoCo constants), got {value!r}.")


class SolverMuJoCoAdaptive(_SolverBase):
    """Adaptive solver with the MuJoCo-variant controller (``solver_mujoco_adaptive.py:418``):
    ``accepted`` advances time, ``commit`` writes the doubled state, a floor-diverged world holds
    its last good state and is latched in ``diverged``.

    Each inner step is MuJoCo's soft-constraint step (``contact_model="mujoco"``, the default on
    free-body models): contacts are rows with MuJoCo's reference acceleration (solref), impedance
    (solimp) and regulariser, mixed per pair from ``model.mujoco`` (``kernels.py:127-190``), on the
    pyramidal cone, minimised by Newton's method, then Euler (``solver_mujoco_adaptive.py:585-605``:
    ``mujoco_warp.step`` with the reference's defaults solver "newton", integrator "implicitfast",
    cone "pyramidal").  For free bodies without joint damping or actuators "euler" and
    "implicitfast" are the same update.  ``contact_model="sap"`` keeps the SAP contact solve under
    this controller; articulated models need it (their MuJoCo inner step is not built).

    The MuJoCo options of ``SolverMuJoCo`` that bear on this step are checked, not dropped:
    ``solver`` ("newton" / "cg": the same minimiser, solved here by Newton), ``integrator``
    ("euler" / "implicitfast"), ``cone`` ("pyramidal"), ``impratio`` (> 0), ``iterations`` (the
    Newton cap), ``ls_iterations``, ``tolerance``, ``ls_tolerance`` (> 0; the stopping rule here is
    this solver's optimality test), ``nconmax`` (at most 128 contacts per world), ``njmax``,
    ``disable_contacts`` (False only), ``use_mujoco_cpu`` (False only).  Other keywords raise
    ``TypeError``."""

    _controller = 1

    def __init__(self, model, *, tol=1e-3, dt_inner_init=0.01, dt_inner_min=1e-6, dt_inner_max=None,
                 dt_mode="per_world", tiling="ragged", max_substeps=256, contact_model=None, solver=None,
                 integrator=None, cone=None, impratio=None, iterations=None, ls_iterations=None, tolerance=None,
                 ls_tolerance=None, njmax=None, nconmax=None, disable_contacts=False, use_mujoco_cpu=False,
                 separate_worlds=None, **kwargs):
        if dt_mode != "per_world":
            raise ValueError(f"dt_mode must be 'per_world' (the only supported mode), got {dt_mode!r}")
        if tiling != "ragged":
            raise ValueError(f"tiling must be 'ragged' (the only supported mode), got {tiling!r}")
        articulated = bool(getattr(model, "is_articulated", False))
        if contact_model is None:
            contact_model = "sap" if articulated else "mujoco"
        if contact_model not in ("mujoco", "sap"):
            raise ValueError(f"contact_model must be 'mujoco' or 'sap', got {contact_model!r}")
        if contact_model == "mujoco" and articulated:
            raise ValueError("contact_model='mujoco' is built for free-body models; pass contact_model='sap' "
                             "for articulated models")
        self.mj_solver = _mj_option(solver, _MJ_SOLVERS, "solver", "newton")
        self.mj_integrator = _mj_option(integrator, _MJ_INTEGRATORS, "integrator", "implicitfast")
        self.mj_cone = _mj_option(cone, _MJ_CONES, "cone", "pyramidal")
        impratio = 1.0 if impratio is None else float(impratio)
        if not impratio > 0.0:
            raise ValueError(f"impratio must be > 0, got {impratio!r}")
        for name, val in (("tolerance", tolerance), ("ls_tolerance", ls_tolerance)):
            if val is not None and not float(val) > 0.0:
                raise ValueError(f"{name} must be > 0, got {val!r}")
        for name, val in (("iterations", iterations), ("ls_iterations", ls_iterations), ("njmax", njmax),
                          ("nconmax", nconmax)):
            if val is not None and int(val) < 1:
                raise ValueError(f"{name} must be >= 1, got {val!r}")
        if nconmax is not None and int(nconmax) > 128:
            raise ValueError(f"nconmax must be at most 128 contacts per world, got {nconmax!r}")
        if disable_contacts:
            raise ValueError("disable_contacts=True is not supported by this solver")
        if use_mujoco_cpu:
            raise ValueError("use_mujoco_cpu=True is not supported: the step runs on the device")
        allowed = {"max_rigid_contact", "max_iterations", "broad_phase", "optimality_rel_tol", "schedule",
                   "cg_rel_tol", "cg_forcing", "world_order", "ls_rel_tol", "line_search_variant",
                   "contact_tau_d", "contact_preset_variant", "warm_start"}
        unknown = set(kwargs) - allowed
        if unknown:
            raise TypeError(f"SolverMuJoCoAdaptive got unexpected keyword arguments {sorted(unknown)}")
        if iterations is not None:
            kwargs.setdefault("max_iterations", int(iterations))
        if nconmax is not None:
            kwargs.setdefault("max_rigid_contact", int(nconmax))
        self.contact_model = contact_model
        self.impratio = impratio
        super().__init__(model, mode="adaptive", tol=tol, dt_inner_init=dt_inner_init, dt_inner_min=dt_inner_min,
                         dt_inner_max=dt_inner_max, max_substeps=max_substeps,
                         _contact_law=1 if contact_model == "mujoco" else 0, _impratio=impratio, **kwargs)
        self._tiling = tiling

    @property
    def tiling(self) -> str:
        return "ragged"

    @property
    def commit(self):
        return self._commit

    @property
    def cumulative_iterations(self):
        return self._iterations.sum().view(1)
