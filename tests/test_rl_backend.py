"""The RL physics-backend seam (``scripts/rl/anymal_study/backends.py``): one ``advance()`` over
the adaptive boundary call and a fixed-step substep loop.

CPU: the spec defaults, the reference backends and argument checks.
GPU: an adaptive ``advance`` equals ``SolverMuJoCoAdaptive.step_dt`` on the same state bit for bit
and lands every world on the boundary; a fixed ``advance`` equals ``round(control_dt / fixed_dt)``
fixed steps, clearing forces and calling ``apply_forces`` before each; articulated models advance
through the same seam.
"""

import importlib

import numpy as np
import pytest
import torch

na = importlib.import_module("newton-adaptive_amd")
rl = importlib.import_module("newton-adaptive_amd.rl_backend")
DEV = "cuda:0"


def test_spec_defaults_and_reference_backends():
    s = rl.BackendSpec()
    assert s.kind == "adaptive" and s.is_adaptive and s.dt_mode == "per_world"
    assert rl.REF_SPECS["ref_tol"].is_adaptive and rl.REF_SPECS["ref_tol"].tol == 1e-5
    assert not rl.REF_SPECS["ref_dt"].is_adaptive and rl.REF_SPECS["ref_dt"].fixed_dt == 5e-4
    m = na.scenes.box_drop(2, device="cpu")
    with pytest.raises(ValueError):
        rl.Backend(m, rl.BackendSpec(kind="bogus"))


@pytest.mark.gpu
def test_adaptive_advance_equals_step_dt():
    model = na.scenes.box_drop(16, seed=2, device=DEV)
    spec = rl.BackendSpec(kind="adaptive", tol=1e-3, dt_inner_init=1e-3, dt_inner_min=1e-6, dt_inner_max=0.01)
    be = rl.make_backend(model, spec)
    ref = na.SolverMuJoCoAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_min=1e-6, dt_inner_max=0.01)
    a0, a1 = model.state(), model.state()
    b0, b1 = model.state(), model.state()
    for _ in range(4):
        a0, a1 = be.advance(0.02, a0, a1, model.control())
        b0, b1 = ref.step_dt(0.02, b0, b1, model.control())
    torch.cuda.synchronize()
    assert np.array_equal(a0.body_q.cpu().numpy(), b0.body_q.cpu().numpy())
    assert np.allclose(be.solver.sim_time.cpu().numpy(), 0.02, rtol=0, atol=1e-12)
    d = be.diagnostics()
    assert set(d) == {"sim_time_min", "sim_time_max", "error_max", "accept_count", "dt_min", "dt_max"}


@pytest.mark.gpu
def test_fixed_advance_is_substep_loop_with_force_callback():
    model = na.scenes.box_drop(8, seed=3, device=DEV)
    be = rl.Backend(model, rl.BackendSpec(kind="fixed", fixed_dt=5e-3))
    assert be.diagnostics() is None
    calls = []

    def push(state):
        calls.append(float(state.body_f.abs().sum()))  # cleared before every substep
        state.body_f[:, 2] += 1.0

    s0, s1 = model.state(), model.state()
    s0, s1 = be.advance(0.02, s0, s1, model.control(), apply_forces=push)
    torch.cuda.synchronize()
    assert len(calls) == 4 and all(c == 0.0 for c in calls)
    ref = na.SolverSAPAdaptive(model, mode="fixed", dt_inner_init=5e-3, dt_inner_max=5e-3, max_substeps=1)
    r = model.state()
    for _ in range(4):
        r.clear_forces()
        r.body_f[:, 2] += 1.0
        ref.step(r, r, model.control(), None, 5e-3)
    torch.cuda.synchronize()
    assert np.array_equal(s0.body_q.cpu().numpy(), r.body_q.cpu().numpy())
    assert ref.accepted_steps() == 4 * model.world_count


@pytest.mark.gpu
def test_articulated_model_advances_through_the_seam():
    model = na.scenes.articulated_mix(8, seed=4, device=DEV)
    be = rl.Backend(model, rl.BackendSpec(kind="adaptive", tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01))
    s0, s1 = model.state(), model.state()
    q_start = s0.joint_q.clone()
    for _ in range(3):
        s0, s1 = be.advance(0.02, s0, s1, model.control())
    torch.cuda.synchronize()
    assert torch.isfinite(s0.joint_q).all() and not torch.equal(s0.joint_q, q_start)
    assert np.allclose(be.solver.sim_time.cpu().numpy(), 0.02, rtol=0, atol=1e-12)
