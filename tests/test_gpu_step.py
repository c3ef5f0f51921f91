"""GPU step-doubling path against the FP64 oracle (tolerances are stated per check).

* one inner step per world (``na_substep``, both solves converged to a relative optimality of
  1e-9): pose within 1e-5 relative to max(|q|, 1); velocity within min(max(1e-5, 100 s_w), 1e-4)
  relative to max(|v|, 1 m/s), where s_w is how far the FP64 oracle's own velocity moves when its
  input pose is perturbed at float32 rounding level (6e-8 relative) -- the contact problem of a
  settled pile amplifies input rounding by up to ~1e3, so a fixed 1e-5 velocity bar would test the
  conditioning of the problem rather than the kernel; the relaxation is capped at 1e-4, and the
  largest s_w and every world above 1e-5 are printed;
* one step-doubling attempt (``na_attempt``): accepted-step error norms within 1e-6 absolute,
  doubled pose within 1e-5, for every world: on the contact sets the device solved with, and end
  to end wherever the oracle's own collision found the same contact counts (``tests/parity.py``);
* boundary calls: every world lands on dt_outer; repeated runs are bit-identical; the lock-step
  form (``na_step_plain``) and the persistent served form give bit-identical states (same
  per-world arithmetic); fixed mode does constant work; tighter tolerance does more work.
"""

import importlib
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GE = importlib.import_module("__graft_entry__")
P = importlib.import_module("parity")


def _settle(na, model, frames, **kw):
    state = model.state()
    solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128, **kw)
    for _ in range(frames):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    return state, solver


def _substep(nalib, solver, q64, qd, f, h):
    W, B = solver.model.world_count, solver.model.body_count
    qo = torch.zeros((B, 7), dtype=torch.float64, device=DEV)
    vo = torch.zeros((B, 6), dtype=torch.float32, device=DEV)
    st = torch.zeros((W, 6), dtype=torch.int32, device=DEV)
    hh = torch.full((W,), h, dtype=torch.float64, device=DEV)
    nalib.check(nalib.lib().na_substep(solver._h, q64.data_ptr(), qd.data_ptr(), f.data_ptr(), hh.data_ptr(), None,
                                       qo.data_ptr(), vo.data_ptr(), None, st.data_ptr(), nalib.stream_handle()),
                "substep")
    torch.cuda.synchronize()
    return qo.cpu().numpy(), vo.cpu().numpy(), st.cpu().numpy()


SCENES = [("drop", 25), ("drop", 40), ("stack", 3), ("stack", 30), ("pile", 20), ("pile", 60)]


def _model(na, name):
    if name == "drop":
        return na.scenes.box_drop(64, seed=4, device=DEV)
    if name == "stack":
        return na.scenes.box_stack(64, seed=5, device=DEV)
    return na.scenes.uneven_pile(64, seed=6, device=DEV)


@pytest.mark.parametrize("name,frames", SCENES)
def test_substep_matches_oracle(na, O, nalib, name, frames):
    model = _model(na, name)
    state, _ = _settle(na, model, frames)
    solver = na.SolverSAPAdaptive(model, tol=1e-3, optimality_rel_tol=1e-9, max_iterations=200)
    q64 = _settle_q64(na, model, state)
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    qo, vo, st = _substep(nalib, solver, q64, qd, f, 2e-3)
    assert st[:, 2].min() == 1, f"GPU solve did not converge: exit reasons {np.bincount(st[:, 4]).tolist()}"
    om = O.OracleModel(model)
    cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-12, max_iterations=200)
    qn, qdn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(0)
    zero = np.zeros_like(qdn)
    worst, loose = 0.0, []
    for w in range(model.world_count):
        b0, b1 = int(om.world_body_start[w]), int(om.world_body_start[w + 1])
        rq, rv, rst, ok = O.substep_world(om, cfg, w, qn, qdn, zero, 2e-3)
        assert ok and rst[3] == st[w, 3], f"world {w}: contacts {st[w, 3]} vs {rst[3]}"
        dq = np.abs(qo[b0:b1] - rq).max() / max(np.abs(rq).max(), 1.0)
        assert dq <= 1e-5, f"world {w}: pose dq={dq:.2e}"
        vscale = max(np.abs(rv).max(), 1.0)
        dv = np.abs(vo[b0:b1] - rv).max() / vscale
        sens = 0.0
        for _ in range(2):
            qp = qn * (1.0 + rng.normal(size=qn.shape) * 6e-8)
            _, pv, pst, _ = O.substep_world(om, cfg, w, qp, qdn, zero, 2e-3)
            if pst[3] == rst[3]:
                sens = max(sens, np.abs(pv - rv).max() / vscale)
        worst = max(worst, sens)
        if dv > 1e-5:
            loose.append((w, float(dv), float(sens)))
        assert dv <= min(max(1e-5, 100.0 * sens), 1e-4), f"world {w}: dv={dv:.2e} oracle sensitivity={sens:.2e}"
    print(f"{name} after {frames} frames: max oracle sensitivity s_w = {worst:.2e}; worlds with dv > 1e-5 "
          f"(world, dv, s_w): {[(w, f'{a:.1e}', f'{b:.1e}') for w, a, b in loose]}", flush=True)


def _settle_q64(na, model, state):
    """FP64 pose equal to the settled float32 pose (the shadow rounds to it)."""
    return state.body_q.to(torch.float64).contiguous()


@pytest.mark.parametrize("name,frames", SCENES)
def test_attempt_error_norm_matches_oracle(na, O, name, frames):
    model = _model(na, name)
    state, solver = _settle(na, model, frames)
    q64 = solver.body_q64.clone()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    om = O.OracleModel(model)
    for dt in (1e-3, 4e-3):
        dtt = torch.full((model.world_count,), dt, dtype=torch.float64, device=DEV)
        out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True)
        cfg = O.make_config(tol=1e-3)
        qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
        ref = O.attempt_all(om, cfg, qn, qdn, fn, dtn)
        assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
        print(f"{name} after {frames} frames, dt {dt}", flush=True)
        P.check_free_attempts(O, om, cfg, out, ref, qn, qdn, fn, dtn, model.host)


@pytest.mark.parametrize("frames", [3, 30])
def test_stiff_stack_attempt_at_tol_1e4_matches_oracle(na, O, frames):
    """Stiff contact stack (4 boxes, ke 1e6, mu ~ U[0.3, 1]) at tol 1e-4: error norms within 1e-6
    and the doubled pose within 1e-5 of the FP64 oracle."""
    model = na.scenes.box_stack(64, seed=5, device=DEV)
    state = model.state()
    solver = na.SolverSAPAdaptive(model, tol=1e-4, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)
    for _ in range(frames):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    q64 = solver.body_q64.clone()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    om = O.OracleModel(model)
    for dt in (2.5e-4, 1e-3):
        dtt = torch.full((model.world_count,), dt, dtype=torch.float64, device=DEV)
        out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True)
        cfg = O.make_config(tol=1e-4)
        qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
        ref = O.attempt_all(om, cfg, qn, qdn, fn, dtn)
        assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
        P.check_free_attempts(O, om, cfg, out, ref, qn, qdn, fn, dtn, model.host)


def test_boundary_determinism_and_plain_equals_served(na):
    runs = {}
    for schedule in ("served", "plain", "served"):
        model = na.scenes.uneven_pile(128, seed=9, device=DEV, max_prims=12)
        state = model.state()
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256,
                                      schedule=schedule)
        for _ in range(6):
            solver.step(state, state, model.control(), None, 1.0 / 60.0)
            torch.cuda.synchronize()
            st = solver.sim_time.cpu().numpy()
            assert np.allclose(st, 1.0 / 60.0, rtol=1e-12, atol=1e-15)
        q = state.body_q.cpu().numpy()
        assert np.all(np.isfinite(q))
        # FP64 shadow rounds to the caller's float32 pose
        assert np.array_equal(solver.body_q64.float().cpu().numpy(), q)
        key = schedule if schedule not in runs else schedule + "2"
        runs[key] = (q, state.body_qd.cpu().numpy(), solver.accepted_steps())
    assert np.array_equal(runs["served"][0], runs["served2"][0])
    assert np.array_equal(runs["served"][0], runs["plain"][0])
    assert np.array_equal(runs["served"][1], runs["plain"][1])
    assert runs["served"][2] == runs["plain"][2]


def test_fixed_mode_constant_work_and_tol_ordering(na):
    model = na.scenes.box_drop(32, seed=1, device=DEV)
    state = model.state()
    fixed = na.SolverSAPAdaptive(model, mode="fixed", dt_inner_init=0.0025, max_substeps=64)
    work = []
    for _ in range(10):
        fixed.step(state, state, model.control(), None, 0.01)
        work.append(int(fixed.substeps.sum().item()))
    assert len(set(work)) == 1 and work[0] == 32 * 4
    acc = {}
    for tol in (1e-2, 1e-5):
        m = na.scenes.box_drop(32, seed=1, device=DEV)
        s = m.state()
        sol = na.SolverSAPAdaptive(m, tol=tol, dt_inner_init=0.01, max_substeps=256)
        for _ in range(40):
            sol.step(s, s, m.control(), None, 0.01)
        acc[tol] = sol.accepted_steps()
    assert acc[1e-5] >= acc[1e-2]


def test_mujoco_variant_controller_runs(na):
    model = na.scenes.box_drop(16, seed=2, device=DEV)
    state = model.state()
    solver = na.SolverMuJoCoAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01)
    for _ in range(60):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    q = state.body_q.cpu().numpy()
    assert np.all(np.isfinite(q)) and np.all(q[:, 2] > 0.03)
    assert np.allclose(solver.sim_time.cpu().numpy(), 0.01, rtol=1e-12)
    assert not bool(solver.diverged.any())


def test_reset_masked_and_invalid_args(na):
    model = na.scenes.box_drop(8, seed=2, device=DEV)
    state = model.state()
    solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3)
    solver.step(state, state, model.control(), None, 0.01)
    mask = torch.tensor([1, 0] * 4, dtype=torch.bool, device=DEV)
    solver.reset(state, world_mask=mask)
    torch.cuda.synchronize()
    st = solver.sim_time.cpu().numpy()
    assert np.all(st[::2] == 0.0) and np.all(st[1::2] > 0.0)
    assert np.all(solver.dt.cpu().numpy()[::2] == 1e-3)
    with pytest.raises(ValueError):
        solver.step(state, state, model.control(), None, None)
    with pytest.raises(ValueError):
        na.SolverSAPAdaptive(model, mode="bogus")
    with pytest.raises(ValueError):
        na.SolverSAPAdaptive(model, tol=0.0)


def _resting_spheres(na, worlds=8, radius=0.1):
    """Four spheres per world pressed into the ground (test_adaptive_warmstart_gate.py:50-62)."""
    b = na.ModelBuilder()
    for w in range(worlds):
        b.begin_world()
        for i in range(4):
            body = b.add_body(xform=(0.3 * i, 0.0, radius - 0.01 - 0.002 * w, 0.0, 0.0, 0.0, 1.0))
            b.add_shape_sphere(body, radius=radius)
        b.end_world()
    b.add_ground_plane()
    return b.finalize(DEV)


def test_warm_start_gate(na):
    """substep(..., guess): a seed at the solve's own converged velocity meets the optimality test
    before any Newton iteration and is returned unchanged; a seed 0.2 m/s off converges back
    within 1e-6 (the reference's warm-start gate, test_adaptive_warmstart_gate.py:66-130)."""
    model = _resting_spheres(na)
    dt = 1.0 / 60.0
    solver = na.SolverSAPAdaptive(model, mode="fixed", dt_inner_init=dt, max_substeps=1, max_iterations=80)
    state, out = model.state(), model.state()
    state.body_qd[:, 2] = -1.0
    for _ in range(8):
        solver.step(state, state, model.control(), None, dt)
    st0 = solver.substep(state, out, model.control(), None, dt)
    sol = solver.last_substep_velocity.clone()
    torch.cuda.synchronize()
    assert st0[:, 2].min().item() == 1 and st0[:, 3].min().item() >= 4
    st1 = solver.substep(state, out, model.control(), None, dt, guess=sol)
    torch.cuda.synchronize()
    assert st1[:, 0].max().item() == 0 and st1[:, 2].min().item() == 1
    assert torch.equal(solver.last_substep_velocity, sol)
    st2 = solver.substep(state, out, model.control(), None, dt, guess=sol + 0.2)
    torch.cuda.synchronize()
    assert st2[:, 0].min().item() >= 1 and st2[:, 2].min().item() == 1
    assert (solver.last_substep_velocity - sol).abs().max().item() <= 1e-6


def test_box_drop_energy_bound_and_rest(na):
    """The box drop on the GPU (test_rigid_contact.py:766 box-drop invariant): no box ever moves
    faster than free fall from its start height allows, sqrt(2 g (z0 - h)) (1% margin for the
    contact compliance), and after 1000 accepted steps every box rests on a face at COM height =
    half extent."""
    m = na.scenes.box_drop(64, seed=0, device=DEV)
    state = m.state()
    z0 = state.body_q[:, 2].double().clone()
    vmax = torch.sqrt(2.0 * 9.81 * (z0 - 0.05).clamp(min=0.0)) * 1.01 + 1e-3
    solver = na.SolverSAPAdaptive(m, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)
    frames = 0
    per_world = torch.zeros(m.world_count, dtype=torch.int64, device=DEV)
    while int(per_world.min().item()) < 1000:
        solver.step(state, state, m.control(), None, 0.01)
        per_world += solver.substeps
        speed = state.body_qd[:, :3].double().norm(dim=1)
        assert bool((speed <= vmax).all()), f"frame {frames}: speed {speed.max().item():.3f}"
        frames += 1
        assert frames < 5000
    torch.cuda.synchronize()
    q = state.body_q.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(q[:, 2] - 0.05) < 2e-3), q[:, 2]
    assert state.body_qd.abs().max().item() < 1e-3
    x, y, z, w = q[:, 3], q[:, 4], q[:, 5], q[:, 6]
    rz = np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)
    assert np.all(np.abs(rz).max(axis=1) > 1 - 1e-4)


def _ragged(na, counts, seed=11):
    """Worlds of counts[w] boxes (0 = an empty world) dropped onto the ground."""
    b = na.ModelBuilder()
    cfg = na.ModelBuilder.ShapeConfig(density=1000.0, ke=1.0e4, mu=0.5)
    for w, n in enumerate(counts):
        rng = np.random.default_rng(seed + w)
        b.begin_world()
        for k in range(n):
            body = b.add_body(xform=(0.25 * k, 0.0, rng.uniform(0.1, 0.3), *na.scenes.random_unit_quaternion(rng)))
            b.add_shape_box(body, hx=0.05, hy=0.05, hz=0.05, cfg=cfg)
        b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=1.0e4, mu=0.5))
    return b.finalize(DEV)


def test_empty_and_ragged_worlds(na):
    """Empty worlds reach every boundary without work, and do not change the others: a model with
    empty worlds interleaved steps its boxes bit for bit like the model without them."""
    counts = [1, 0, 3, 0, 0, 2]
    ragged = _ragged(na, counts)
    # the non-empty worlds alone, with their per-world seeds
    dense = na.ModelBuilder()
    cfg = na.ModelBuilder.ShapeConfig(density=1000.0, ke=1.0e4, mu=0.5)
    for w, n in enumerate(counts):
        if not n:
            continue
        rng = np.random.default_rng(11 + w)
        dense.begin_world()
        for k in range(n):
            body = dense.add_body(xform=(0.25 * k, 0.0, rng.uniform(0.1, 0.3), *na.scenes.random_unit_quaternion(rng)))
            dense.add_shape_box(body, hx=0.05, hy=0.05, hz=0.05, cfg=cfg)
        dense.end_world()
    dense.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=1.0e4, mu=0.5))
    dense = dense.finalize(DEV)
    assert torch.equal(ragged.body_q, dense.body_q)
    out = {}
    for name, m in (("ragged", ragged), ("dense", dense)):
        s = m.state()
        solver = na.SolverSAPAdaptive(m, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128)
        for _ in range(15):
            solver.step(s, s, m.control(), None, 0.01)
            assert np.allclose(solver.sim_time.cpu().numpy(), 0.01, rtol=0, atol=1e-12)
        torch.cuda.synchronize()
        out[name] = (s.body_q.cpu().numpy(), s.body_qd.cpu().numpy(), solver.substeps.cpu().numpy())
    assert np.array_equal(out["ragged"][0], out["dense"][0]) and np.array_equal(out["ragged"][1], out["dense"][1])
    assert np.array_equal(out["ragged"][2][np.array(counts) > 0], out["dense"][2])


def test_contact_overflow_is_flagged_and_deterministic(na):
    """Worlds with more than 128 candidate contacts (64 primitives with the builder's 0.1 m
    contact gap) keep their 128 closest, flag the overflow and stay finite and run-to-run
    identical."""
    m = na.scenes.uneven_pile(4, seed=9, device=DEV, min_prims=64, max_prims=64, gap=0.1)
    res = []
    for _ in range(2):
        s = m.state()
        solver = na.SolverSAPAdaptive(m, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)
        for _ in range(30):
            solver.step(s, s, m.control(), None, 0.01)
        torch.cuda.synchronize()
        res.append((s.body_q.cpu().numpy(), solver._overflow.cpu().numpy().copy()))
    assert np.isfinite(res[0][0]).all()
    assert (res[0][1] > 0).any(), "no world overflowed its 128-contact set"
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_nan_world_sap_raises_mujoco_holds_and_latches(na):
    """A world whose velocity is NaN: the SAP solver raises the reference's RuntimeError
    (solver_sap_adaptive.py:901-905); the MuJoCo variant has no such error, shrinks the step to the
    floor, holds the world's state, latches `diverged` and moves its clock to the boundary
    (solver_mujoco_adaptive.py:120-231), while every other world steps exactly as without it."""
    model = na.scenes.box_drop(8, seed=2, device=DEV)
    kw = dict(tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=64)
    bad = model.state()
    bad.body_qd[3, 2] = float("nan")
    with pytest.raises(RuntimeError):
        na.SolverSAPAdaptive(model, **kw).step(bad, bad, model.control(), None, 0.01)
    bad = model.state()
    bad.body_qd[3, 2] = float("nan")
    good = model.state()
    q_start = bad.body_q.clone()
    sb, sg = na.SolverMuJoCoAdaptive(model, **kw), na.SolverMuJoCoAdaptive(model, **kw)
    for _ in range(3):
        sb.step(bad, bad, model.control(), None, 0.01)
        sg.step(good, good, model.control(), None, 0.01)
        torch.cuda.synchronize()
        assert bool(sb.diverged[3]) and int(sb.diverged.sum()) == 1
        assert np.allclose(sb.sim_time.cpu().numpy(), 0.01, rtol=0, atol=1e-12)
    torch.cuda.synchronize()
    assert torch.equal(bad.body_q[3], q_start[3])
    others = [w for w in range(8) if w != 3]
    assert torch.equal(bad.body_q[others], good.body_q[others])
    assert torch.equal(bad.body_qd[others], good.body_qd[others])


def test_attempt_with_trend_matches_oracle(na, O):
    """na_attempt with the solver's own trend (the seed accelerations its boundary calls left)
    against the oracle's attempt given the same trend: the error-norm and pose bars of the attempt
    parity above, on every world."""
    model = _model(na, "pile")
    state, solver = _settle(na, model, 20)
    assert solver.warm_start == "extrapolate"
    trend = solver._trend.clone()
    assert trend.abs().max().item() > 0.0
    q64 = solver.body_q64.clone()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    om = O.OracleModel(model)
    dtt = torch.full((model.world_count,), 2e-3, dtype=torch.float64, device=DEV)
    out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True, trend=trend)
    cfg = O.make_config(tol=1e-3, max_iterations=100)
    qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
    tn = trend.cpu().numpy().astype(np.float64)
    ref = O.attempt_all(om, cfg, qn, qdn, fn, dtn, trend=tn)
    assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
    P.check_free_attempts(O, om, cfg, out, ref, qn, qdn, fn, dtn, model.host, trend=tn)


def test_trend_of_a_free_fall_is_gravity_and_reset_clears_it(na):
    """After a frame of free fall the per-body trend is gravity (float32 velocities over
    millisecond steps: within 1e-3 relative); reference seeds keep none; reset() clears it."""
    for seeds in ("extrapolate", "reference"):
        model = na.scenes.box_drop(16, seed=2, device=DEV)
        state = model.state()
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=64,
                                      warm_start=seeds)
        solver.step(state, state, model.control(), None, 0.01)
        torch.cuda.synchronize()
        assert solver.substeps.min().item() > 1
        tr = solver._trend.cpu().numpy()
        if seeds == "reference":
            assert not tr.any()
            continue
        assert np.abs(tr[:, 2] + 9.81).max() < 9.81e-3, tr
        assert np.abs(tr[:, [0, 1, 3, 4, 5]]).max() < 1e-3, tr
        mask = torch.zeros(model.world_count, dtype=torch.bool, device=DEV)
        mask[::2] = True
        solver.reset(world_mask=mask)
        torch.cuda.synchronize()
        tr2 = solver._trend.cpu().numpy()
        assert not tr2[::2].any() and np.array_equal(tr2[1::2], tr[1::2])


@pytest.mark.parametrize("seeds", ["extrapolate", "concurrent", "reference"])
def test_served_and_plain_identical_for_both_seeds(na, seeds):
    """The served and the lock-step boundary forms give bit-identical states under every Newton
    seed rule (the trend lives in device memory between the lock-step launches), also when the
    lock-step form runs the reference's fixed-batch work pattern (finished worlds still collide)."""
    runs = {}
    for schedule in ("served", "plain", "plain_batch"):
        model = na.scenes.uneven_pile(96, seed=21, device=DEV, max_prims=16)
        state = model.state()
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256,
                                      schedule=schedule, warm_start=seeds)
        for _ in range(8):
            solver.step(state, state, model.control(), None, 0.01)
        torch.cuda.synchronize()
        runs[schedule] = (state.body_q.cpu().numpy(), state.body_qd.cpu().numpy(), solver._trend.cpu().numpy())
    for k in range(3):
        assert np.array_equal(runs["served"][k], runs["plain"][k])
        assert np.array_equal(runs["served"][k], runs["plain_batch"][k])


@pytest.mark.parametrize("tiles", [1, 12, 96])
def test_tiled_boundary_calls_equal_single_wave(na, tiles):
    """Boundary calls with two-wave tiles (the previous call's costliest worlds first, each solved
    by two wavefronts) leave the same states as single-wave boundary calls and as the lock-step
    form, bit for bit, under the concurrent seeds; the first call has no cost history yet."""
    runs = {}
    for name in ("untiled", "tiled", "plain"):
        model = na.scenes.uneven_pile(96, seed=23, device=DEV, min_prims=2, max_prims=24)
        state = model.state()
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256,
                                      schedule="plain" if name == "plain" else "served", warm_start="concurrent")
        if name == "tiled":
            solver.set_tiles(tiles)
        for _ in range(6):
            solver.step(state, state, model.control(), None, 0.01)
        torch.cuda.synchronize()
        runs[name] = (state.body_q.cpu().numpy(), state.body_qd.cpu().numpy(), solver._trend.cpu().numpy(),
                      solver.accepted_steps())
    for k in range(3):
        assert np.array_equal(runs["tiled"][k], runs["untiled"][k])
        assert np.array_equal(runs["plain"][k], runs["untiled"][k])
    assert runs["tiled"][3] == runs["untiled"][3] > 0


def test_extrapolated_seeds_cut_newton_iterations(na):
    """On settled piles the extrapolated seeds take fewer Newton iterations than the reference's
    (the oracle study: 5-25% per heavy world); both runs reach every boundary with finite states."""
    model = na.scenes.uneven_pile(256, seed=31, device=DEV)
    state, warm = _settle(na, model, 20, warm_start="reference")
    q0, qd0 = state.body_q.clone(), state.body_qd.clone()
    newton = {}
    for seeds in ("reference", "extrapolate"):
        st = na.State(q0.clone(), qd0.clone(), torch.zeros_like(qd0))
        solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128,
                                      warm_start=seeds)
        solver.step(st, st, model.control(), None, 0.01)  # builds the trend
        solver.reset_work_counters()
        for _ in range(6):
            solver.step(st, st, model.control(), None, 0.01)
        torch.cuda.synchronize()
        assert np.isfinite(st.body_q.cpu().numpy()).all()
        assert np.allclose(solver.sim_time.cpu().numpy(), 0.01, rtol=1e-12, atol=1e-15)
        w = solver.work_counters()
        newton[seeds] = (w["newton_iterations"], w["attempts"])
    ratio = newton["extrapolate"][0] / newton["reference"][0]
    print(f"Newton iterations over 6 frames: reference {newton['reference']}, extrapolate {newton['extrapolate']} "
          f"(ratio {ratio:.3f})", flush=True)
    assert ratio < 0.97
