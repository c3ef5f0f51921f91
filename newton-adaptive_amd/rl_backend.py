"""Physics-backend seam of the RL study: one ``advance()`` over the adaptive boundary call and a
fixed-step substep loop, on the same model, control and gains.

Mirrors ``scripts/rl/anymal_study/backends.py:17-130`` (``BackendSpec``, ``REF_SPECS``,
``Backend``, ``make_backend``): the adaptive kinds drive ``SolverMuJoCoAdaptive.step_dt``; the
fixed kinds run ``round(control_dt / fixed_dt)`` substeps of a fixed-step solver (this build's
solver in ``mode="fixed"``: one convex step of ``fixed_dt`` per substep, no step doubling), clearing
forces and applying ``apply_forces`` before every substep.  ``advance`` returns the state pair with
the live (post-step) state first for both kinds.  The MuJoCo constraint-solver settings of the spec
(``solver``, ``ls_iterations``, ``njmax``, ``nconmax``) are accepted for interface parity; the
inner solve here is the convex contact solve of this build.
"""

from __future__ import annotations

from dataclasses import dataclass

from .solvers import SolverMuJoCoAdaptive, SolverSAPAdaptive


@dataclass
class BackendSpec:
    """Solver configuration of one physics backend (``backends.py:17-58``).

    ``kind``: ``"adaptive"``, ``"fixed"``, ``"ref_tol"`` (tight-tolerance adaptive reference) or
    ``"ref_dt"`` (small fixed-dt reference).  ``tol`` [m or rad], ``dt_inner_*`` [s] and
    ``dt_mode`` configure the adaptive kinds; ``fixed_dt`` [s] the fixed ones.
    """

    kind: str = "adaptive"
    tol: float = 1e-3
    dt_inner_init: float = 5e-3
    dt_inner_min: float = 2e-7
    dt_inner_max: float = 0.02
    dt_mode: str = "per_world"
    fixed_dt: float = 5e-3
    solver: str = "newton"
    ls_iterations: int = 50
    njmax: int = 50
    nconmax: int = 100

    @property
    def is_adaptive(self) -> bool:
        return self.kind in ("adaptive", "ref_tol")


# High-fidelity reference backends of the transfer matrix (backends.py:61-64).
REF_SPECS: dict = {
    "ref_tol": BackendSpec(kind="ref_tol", tol=1e-5, dt_inner_min=5e-8, dt_inner_max=2e-3),
    "ref_dt": BackendSpec(kind="ref_dt", fixed_dt=5e-4),
}

_KINDS = ("adaptive", "fixed", "ref_tol", "ref_dt")


class Backend:
    """Uniform ``advance()`` over the adaptive ``step_dt`` and a fixed-step substep loop
    (``backends.py:67-125``)."""

    def __init__(self, model, spec: BackendSpec):
        if spec.kind not in _KINDS:
            raise ValueError(f"BackendSpec.kind must be one of {_KINDS}, got {spec.kind!r}")
        self.spec = spec
        self.model = model
        if spec.is_adaptive:
            self.solver = SolverMuJoCoAdaptive(model, tol=spec.tol, dt_inner_init=spec.dt_inner_init,
                                               dt_inner_min=spec.dt_inner_min, dt_inner_max=spec.dt_inner_max,
                                               dt_mode=spec.dt_mode, njmax=spec.njmax, nconmax=spec.nconmax,
                                               solver=spec.solver, ls_iterations=spec.ls_iterations)
        else:
            if not spec.fixed_dt > 0.0:
                raise ValueError(f"fixed_dt must be > 0, got {spec.fixed_dt!r}")
            self.solver = SolverSAPAdaptive(model, mode="fixed", dt_inner_init=spec.fixed_dt,
                                            dt_inner_max=spec.fixed_dt, max_substeps=1)

    @property
    def is_adaptive(self) -> bool:
        return self.spec.is_adaptive

    def advance(self, control_dt: float, s0, s1, control, apply_forces=None):
        """Advance every world by ``control_dt`` seconds; returns ``(state_0, state_1)`` with the
        live state first."""
        if self.is_adaptive:
            return self.solver.step_dt(control_dt, s0, s1, control, apply_forces=apply_forces)
        n = max(1, round(control_dt / self.spec.fixed_dt))
        for _ in range(n):
            s0.clear_forces()
            if apply_forces is not None:
                apply_forces(s0)
            # the fixed-step solver advances its input state in place: the live state stays s0
            self.solver.step(s0, s1, control, None, self.spec.fixed_dt)
        return s0, s1

    def diagnostics(self) -> dict | None:
        """Adaptive-only stepping diagnostics (reads device arrays: keep off the hot path)."""
        if not self.is_adaptive:
            return None
        return self.solver.get_status_summary()


def make_backend(model, spec: BackendSpec) -> Backend:
    return Backend(model, spec)
