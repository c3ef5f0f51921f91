"""Control inputs and applied forces through the boundary (``Control`` of ``control.py:31-180``,
``State.body_f``), and the ``armijo_decay`` line search, on the GPU against the FP64 oracle.

* articulated inner step with joint forces (revolute torques and a FREE root's COM wrench), drive
  position / velocity targets, feedforward and applied body wrenches: coordinates within 1e-6,
  velocities within the capped sensitivity bar of ``test_gpu_articulation`` (at most 1e-4
  relative); every world's result moves away from the uncontrolled one by more than the bar;
* per-frame targets written into ``Control`` between boundary calls move the arms and quadrupeds
  exactly as the oracle's frame driver does with the same targets (equal accepted-step counts,
  coordinates within 1e-5), and ``apply_forces`` changes an articulated world;
* a free body's FREE-joint force equals the same wrench given as ``body_f`` (bit-identical);
* the ``armijo_decay`` line search: an inner step of piles and of articulated worlds matches the
  oracle's backtracking restatement (pose within 1e-5 / 1e-6) and converges.
"""

import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
na = importlib.import_module("newton-adaptive_amd")
_T = importlib.import_module("tests.test_gpu_articulation")


def _control(model, rng, scale=1.0):
    """Random control of an articulated model: targets near the model's, velocity targets,
    feedforward, joint forces (revolute torques, FREE wrenches), and body wrenches."""
    D, B = model.joint_dof_count, model.body_count
    c = model.control()
    tq = model.host("joint_target_pos").astype(np.float64) + rng.normal(size=D) * 0.3 * scale
    c.joint_target_q = torch.as_tensor(tq, dtype=torch.float32, device=DEV)
    c.joint_target_qd = torch.as_tensor(rng.normal(size=D) * 0.5 * scale, dtype=torch.float32, device=DEV)
    c.joint_act = torch.as_tensor(rng.normal(size=D) * 0.5 * scale, dtype=torch.float32, device=DEV)
    c.joint_f = torch.as_tensor(rng.normal(size=D) * 2.0 * scale, dtype=torch.float32, device=DEV)
    bf = (rng.normal(size=(B, 6)) * np.array([5.0, 5.0, 5.0, 0.5, 0.5, 0.5]) * scale).astype(np.float32)
    return c, bf


def _oracle_control(oa, c, bf):
    oa.set_control(joint_f=c.joint_f.cpu().numpy(), joint_target_q=c.joint_target_q.cpu().numpy(),
                   joint_target_qd=c.joint_target_qd.cpu().numpy(), joint_act=c.joint_act.cpu().numpy(),
                   body_f=bf)


def _vel_bar(sens):
    """Velocity bar: 100x the oracle's own response to a float32-size coordinate perturbation,
    at least 1e-5 and at most 1e-4 (relative to max(|v|, 1))."""
    return min(max(1e-5, 100.0 * sens), 1e-4)


def test_art_control_substep_matches_oracle(O, nalib):
    model = na.scenes.articulated_mix(16, seed=2, device=DEV)
    q, v = _T._rollout(O, model, 40)
    rng = np.random.default_rng(11)
    ctl, bf = _control(model, rng)
    print(f"articulated substep with control: worlds={model.world_count}", flush=True)
    solver = na.SolverSAPAdaptive(model, tol=1e-3, optimality_rel_tol=1e-9, max_iterations=200)
    qo, vo, st = _T._art_substep(nalib, solver, q, v, 2e-3, body_f=bf, control=ctl)
    q0o, _, _ = _T._art_substep(nalib, solver, q, v, 2e-3)
    assert st[:, 2].min() == 1, f"GPU solves not converged: {st.tolist()}"
    om, oa = O.OracleModel(model), O.OracleArt(model)
    _oracle_control(oa, ctl, bf)
    cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-12, max_iterations=200)
    rng = np.random.default_rng(0)
    worst = 0.0
    for w in range(model.world_count):
        q0, q1 = oa.world_q_start[w], oa.world_q_start[w + 1]
        d0, d1 = oa.world_qd_start[w], oa.world_qd_start[w + 1]
        rq, rv, rst, ok = O.art_substep(om, oa, cfg, w, q[q0:q1], v[d0:d1], 2e-3)
        assert ok and rst[3] == st[w, 3], f"world {w}: contacts {st[w, 3]} vs {rst[3]}"
        dq = np.abs(qo[q0:q1] - rq).max()
        assert dq <= 1e-6, f"world {w}: dq={dq:.2e}"
        moved = np.abs(qo[q0:q1] - q0o[q0:q1]).max()
        assert moved > 1e-5, f"world {w}: the control changed the step by only {moved:.2e}"
        vscale = max(np.abs(rv).max(), 1.0)
        dv = np.abs(vo[d0:d1] - rv).max() / vscale
        sens = 0.0
        for _ in range(2):
            qp = q[q0:q1] * (1.0 + rng.normal(size=q1 - q0) * 6e-8)
            _, pv, pst, _ = O.art_substep(om, oa, cfg, w, qp, v[d0:d1], 2e-3)
            if pst[3] == rst[3]:
                sens = max(sens, np.abs(pv - rv).max() / vscale)
        worst = max(worst, sens)
        assert dv <= _vel_bar(sens), f"world {w}: dv={dv:.2e} oracle sensitivity={sens:.2e}"
    print(f"max oracle sensitivity s_w = {worst:.2e}", flush=True)


def test_per_frame_targets_and_forces_match_oracle_driver(O):
    """An RL-style loop: new drive targets, joint forces and body wrenches in every frame."""
    model = na.scenes.articulated_mix(16, seed=7, device=DEV)
    q, v = _T._rollout(O, model, 30)
    state = model.state()
    state.joint_q.copy_(torch.as_tensor(q, dtype=torch.float32, device=DEV))
    state.joint_qd.copy_(torch.as_tensor(v, dtype=torch.float32, device=DEV))
    solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=2e-3, max_substeps=64)
    om, oa = O.OracleModel(model), O.OracleArt(model)
    cfg = O.make_config(tol=1e-3, dt_init=2e-3, max_substeps=64, opt_rel_tol=1e-6)
    ref = O.OracleSolver(om, cfg)
    rq, rv = q.copy(), v.copy()
    rng = np.random.default_rng(5)
    free_q = model.state()
    free_q.joint_q.copy_(state.joint_q)
    free_q.joint_qd.copy_(state.joint_qd)
    free_solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=2e-3, max_substeps=64)
    for frame in range(4):
        ctl, bf = _control(model, rng, scale=0.5)
        print(f"frame {frame}: new targets and forces", flush=True)

        def apply_forces(s, bf=bf):
            s.body_f.copy_(torch.as_tensor(bf, device=DEV))

        solver.step(state, state, ctl, None, 0.01, apply_forces=apply_forces)
        free_solver.step(free_q, free_q, model.control(), None, 0.01)
        _oracle_control(oa, ctl, bf)
        assert ref.step_articulated(oa, rq, rv, 0.01) == 0
        rv[:] = rv.astype(np.float32).astype(np.float64)  # float32 velocity across frames, as on the device
        torch.cuda.synchronize()
        sub = solver.substeps.cpu().numpy()
        assert np.array_equal(sub, ref.a["substeps"]), f"frame {frame}: {sub} vs {ref.a['substeps']}"
        dq = np.abs(state.joint_q.cpu().numpy().astype(np.float64) - rq).max()
        assert dq <= 1e-5, f"frame {frame}: max |dq| = {dq:.2e}"
    # the controlled worlds are not where the uncontrolled ones went
    qs = model.host("world_q_start")
    diff = np.abs(state.joint_q.cpu().numpy() - free_q.joint_q.cpu().numpy())
    for w in range(model.world_count):
        e = int(qs[w + 1]) if w + 1 < len(qs) else len(diff)
        assert diff[int(qs[w]):e].max() > 1e-4, f"world {w}: control had no effect"


def test_free_body_joint_f_equals_body_f():
    """A free body's FREE-joint force is a COM wrench in the world frame, like State.body_f."""
    results = []
    for via in ("joint_f", "body_f"):
        model = na.scenes.box_drop(32, seed=3, device=DEV)
        state = model.state()
        ctl = model.control()
        rng = np.random.default_rng(2)
        wrench = torch.as_tensor(rng.normal(size=(model.body_count, 6)) * np.array([3, 3, 3, 0.05, 0.05, 0.05]),
                                 dtype=torch.float32, device=DEV)
        if via == "joint_f":
            ctl.joint_f.copy_(wrench.view(-1))
        else:
            state.body_f.copy_(wrench)
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128)
        for _ in range(5):
            solver.step(state, state, ctl, None, 0.01)
        torch.cuda.synchronize()
        results.append((state.body_q.cpu().numpy(), state.body_qd.cpu().numpy()))
    assert np.array_equal(results[0][0], results[1][0]) and np.array_equal(results[0][1], results[1][1])
    still = na.scenes.box_drop(32, seed=3, device=DEV)
    s0 = still.state()
    sol0 = na.SolverSAPAdaptive(still, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128)
    for _ in range(5):
        sol0.step(s0, s0, still.control(), None, 0.01)
    torch.cuda.synchronize()
    assert not np.array_equal(s0.body_q.cpu().numpy(), results[0][0])


@pytest.mark.parametrize("frames", [3, 30])
def test_armijo_decay_substep_matches_oracle(O, nalib, frames):
    model = na.scenes.uneven_pile(32, seed=6, device=DEV)
    state = model.state()
    settle = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128)
    for _ in range(frames):
        settle.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    print(f"armijo_decay substep: pile frames={frames}", flush=True)
    solver = na.SolverSAPAdaptive(model, tol=1e-3, optimality_rel_tol=1e-9, max_iterations=200,
                                  line_search_variant="armijo_decay")
    W, B = model.world_count, model.body_count
    q64 = state.body_q.double().contiguous()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    qo = torch.zeros((B, 7), dtype=torch.float64, device=DEV)
    vo = torch.zeros((B, 6), dtype=torch.float32, device=DEV)
    st = torch.zeros((W, 6), dtype=torch.int32, device=DEV)
    hh = torch.full((W,), 2e-3, dtype=torch.float64, device=DEV)
    nalib.check(nalib.lib().na_substep(solver._h, q64.data_ptr(), qd.data_ptr(), f.data_ptr(), hh.data_ptr(), None,
                                       qo.data_ptr(), vo.data_ptr(), None, st.data_ptr(), nalib.stream_handle()),
                "substep")
    torch.cuda.synchronize()
    st = st.cpu().numpy()
    assert st[:, 2].min() == 1, f"exit reasons {np.bincount(st[:, 4]).tolist()}"
    om = O.OracleModel(model)
    cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-12, max_iterations=200, line_search="armijo_decay")
    qn, qdn, qo = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), qo.cpu().numpy()
    zero = np.zeros_like(qdn)
    for w in range(W):
        b0, b1 = int(om.world_body_start[w]), int(om.world_body_start[w + 1])
        rq, rv, rst, ok = O.substep_world(om, cfg, w, qn, qdn, zero, 2e-3)
        assert ok and rst[3] == st[w, 3], f"world {w}: contacts {st[w, 3]} vs {rst[3]}"
        dq = np.abs(qo[b0:b1] - rq).max() / max(np.abs(rq).max(), 1.0)
        assert dq <= 1e-5, f"world {w}: pose dq={dq:.2e}"


def test_armijo_decay_articulated_attempt_matches_oracle(O, nalib):
    model = na.scenes.articulated_mix(16, seed=2, device=DEV)
    q, v = _T._rollout(O, model, 40)
    print("armijo_decay articulated attempt", flush=True)
    solver = na.SolverSAPAdaptive(model, tol=1e-3, line_search_variant="armijo_decay")
    out = _T._art_attempt(nalib, solver, q, v, 2e-3)
    om, oa = O.OracleModel(model), O.OracleArt(model)
    ref = O.art_attempt_all(om, oa, O.make_config(tol=1e-3, opt_rel_tol=1e-6, line_search="armijo_decay"), q, v,
                            np.full(model.world_count, 2e-3))
    assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
    assert np.array_equal(out["n_contacts"], ref["n_contacts"])
    de = np.abs(out["err"] - ref["err"])
    assert np.all(de <= 1e-7 + 1e-3 * ref["err"]), f"err gpu {out['err']} oracle {ref['err']}"
    assert np.abs(out["q_double"] - ref["q_double"]).max() <= 1e-6
