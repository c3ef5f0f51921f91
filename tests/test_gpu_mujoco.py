"""GPU MuJoCo soft-constraint inner step (SolverMuJoCoAdaptive, contact law 1) against the FP64
oracle and the reference's recorded single-drop run (tolerances stated per check).

* the single-drop scene of the reference's results/v1_single_drop.json (tests/golden): device
  inner steps of 2e-5 s (the gold run) and the five fixed steps reproduce the recorded lowest
  sphere heights within 1e-6 m (FP64 pose through the steps; float32 velocities between them);
* one inner step per world (``na_substep``) on dropped boxes, stacks and piles: pose within 1e-5
  relative, velocity within min(max(2e-4, 10 s_w), 1e-3) relative to max(|v|, 1 m/s) (s_w: the
  FP64 oracle's own velocity change under float32-level perturbation of its input pose).  MuJoCo's
  rows carry the contact position into the reference velocity as h K imp pos, K = 1/(dmax^2 tc^2
  dr^2) ~ 1.1e6 on the stiff stacks (ke 1e6, kd 100): a float32 contact distance (~1e-7 m at these
  heights) moves that velocity by h K imp 1e-7 ~ 2e-4 m/s at h = 2e-3, hence the floor; a
  just-landed stack also moves its own FP64 velocity by up to ~3e-4 under 6e-8 input rounding;
* one step-doubling attempt (``na_attempt``) for every world: on the contact sets the device
  solved with, error norms within 1e-6 and the doubled pose within 1e-5; end to end where the
  contact sets agree (``tests/parity.py``), the error norm and the doubled pose within 1e-5 and
  1e-4: each side's own collision rounds the contact positions differently (float32 against
  FP64), a MuJoCo row carries its position into the reference velocity as h K imp pos with K up
  to ~1e6 on the stiff stacks, and the settled piles amplify input rounding by up to ~1e5 (the
  oracle's own s_w reaches 9e-3 on the pile after 20 frames);
* boundary calls: the served launch and the per-iteration host loop leave bit-identical states,
  and three frames of dropped boxes match the oracle's frame driver (MuJoCo controller and rows).
"""

import importlib
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
HERE = os.path.dirname(os.path.abspath(__file__))
GE = importlib.import_module("__graft_entry__")
P = importlib.import_module("parity")


def _drop_scene(na, g):
    s = g["scene"]
    b = na.ModelBuilder()
    cfg = na.ModelBuilder.ShapeConfig(ke=s["ke"], kd=s["kd"], kf=0.0, mu=s["mu"], margin=s["margin"])
    body = b.add_body(xform=(0.0, 0.0, s["z0"], 0.0, 0.0, 0.0, 1.0))
    b.add_shape_sphere(body, radius=s["radius"], cfg=cfg)
    b.add_ground_plane()
    return b.finalize(DEV)


def _device_drop(na, nalib, g, dt, per_frame):
    m = _drop_scene(na, g)
    solver = na.SolverMuJoCoAdaptive(m, tol=1e-3, dt_inner_init=dt, dt_inner_min=1e-7, max_iterations=100)
    q = m.body_q.to(torch.float64).clone()
    v = torch.zeros((1, 6), dtype=torch.float32, device=DEV)
    v[0, 2] = g["scene"]["vz"]
    f = torch.zeros_like(v)
    qo, vo = torch.empty_like(q), torch.empty_like(v)
    st = torch.zeros((1, 6), dtype=torch.int32, device=DEV)
    h = torch.full((1,), dt, dtype=torch.float64, device=DEV)
    L = nalib.lib()
    zs = []
    for frame in range(int(round(g["scene"]["duration"] / g["scene"]["dt_outer"]))):
        for _ in range(per_frame):
            nalib.check(L.na_substep(solver._h, q.data_ptr(), v.data_ptr(), f.data_ptr(), h.data_ptr(), None,
                                     qo.data_ptr(), vo.data_ptr(), None, st.data_ptr(), nalib.stream_handle()),
                        "na_substep")
            q.copy_(qo)
            v.copy_(vo)
        zs.append(q[0, 2].clone())
    torch.cuda.synchronize()
    assert int(st[0, 2]) == 1, "the device solve did not converge"
    return float(torch.stack(zs).min())


def test_single_drop_gold_on_device(na, nalib):
    g = json.load(open(os.path.join(HERE, "golden", "v1_single_drop.json")))
    z = _device_drop(na, nalib, g, g["gold_dt"], int(round(g["scene"]["dt_outer"] / g["gold_dt"])))
    print(f"single drop, dt 2e-5: z_min {z:.9f}, recorded {g['gold_zmin']:.9f}", flush=True)
    assert abs(z - g["gold_zmin"]) <= 1e-6


@pytest.mark.parametrize("row", range(5))
def test_single_drop_fixed_steps_on_device(na, nalib, row):
    g = json.load(open(os.path.join(HERE, "golden", "v1_single_drop.json")))
    r = g["fixed"][row]
    z = _device_drop(na, nalib, g, r["fixed_dt"], int(round(g["scene"]["dt_outer"] / r["fixed_dt"])))
    err_mm = abs(z - g["gold_zmin"]) * 1e3
    assert abs(err_mm - r["pen_err_mm"]) <= 1e-3, f"dt {r['fixed_dt']}: {err_mm:.6f} mm against {r['pen_err_mm']:.6f}"


def _model(na, name):
    if name == "drop":
        return na.scenes.box_drop(64, seed=4, device=DEV)
    if name == "stack":
        return na.scenes.box_stack(64, seed=5, device=DEV)
    return na.scenes.uneven_pile(64, seed=6, device=DEV)


def _settle(na, model, frames, **kw):
    state = model.state()
    solver = na.SolverMuJoCoAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=128, **kw)
    for _ in range(frames):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    assert not bool(solver.diverged.any())
    return state, solver


SCENES = [("drop", 25), ("stack", 5), ("stack", 30), ("pile", 20), ("pile", 40)]


@pytest.mark.parametrize("name,frames", SCENES)
def test_mujoco_substep_matches_oracle(na, O, nalib, name, frames):
    model = _model(na, name)
    state, _ = _settle(na, model, frames)
    solver = na.SolverMuJoCoAdaptive(model, tol=1e-3, optimality_rel_tol=1e-9, max_iterations=200)
    q64 = state.body_q.to(torch.float64).contiguous()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    W, B = model.world_count, model.body_count
    qo = torch.zeros((B, 7), dtype=torch.float64, device=DEV)
    vo = torch.zeros((B, 6), dtype=torch.float32, device=DEV)
    st = torch.zeros((W, 6), dtype=torch.int32, device=DEV)
    hh = torch.full((W,), 2e-3, dtype=torch.float64, device=DEV)
    nalib.check(nalib.lib().na_substep(solver._h, q64.data_ptr(), qd.data_ptr(), f.data_ptr(), hh.data_ptr(), None,
                                       qo.data_ptr(), vo.data_ptr(), None, st.data_ptr(), nalib.stream_handle()),
                "substep")
    torch.cuda.synchronize()
    qo, vo, st = qo.cpu().numpy(), vo.cpu().numpy(), st.cpu().numpy()
    assert st[:, 2].min() == 1
    om = O.OracleModel(model)
    cfg = O.make_config(tol=1e-3, controller="mujoco", contact_law="mujoco", opt_rel_tol=1e-12, max_iterations=200)
    qn, qdn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64)
    zero = np.zeros_like(qdn)
    rng = np.random.default_rng(0)
    worst, loose, differ = 0.0, [], []
    for w in range(W):
        b0, b1 = int(om.world_body_start[w]), int(om.world_body_start[w + 1])
        rq, rv, rst, ok = O.substep_world(om, cfg, w, qn, qdn, zero, 2e-3)
        assert ok
        if rst[3] != st[w, 3]:
            # a contact at the gap threshold or a manifold tie: the attempt test compares this world
            # on the device's own contacts (its full step is this substep)
            differ.append((w, int(st[w, 3]), int(rst[3])))
            continue
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
        assert dv <= min(max(2e-4, 10.0 * sens), 1e-3), f"world {w}: dv={dv:.2e} oracle sensitivity={sens:.2e}"
    print(f"{name} after {frames} frames: max s_w = {worst:.2e}; worlds with dv > 1e-5 (world, dv, s_w): "
          f"{[(w, f'{a:.1e}', f'{b:.1e}') for w, a, b in loose]}; contact counts differ (world, gpu, oracle): "
          f"{differ}", flush=True)
    assert len(differ) <= 2


@pytest.mark.parametrize("name,frames", SCENES)
def test_mujoco_attempt_matches_oracle(na, O, name, frames):
    model = _model(na, name)
    state, solver = _settle(na, model, frames)
    q64 = solver.body_q64.clone()
    qd = state.body_qd.clone()
    f = torch.zeros_like(qd)
    om = O.OracleModel(model)
    cfg = O.make_config(tol=1e-3, controller="mujoco", contact_law="mujoco")
    for dt in (1e-3, 4e-3):
        dtt = torch.full((model.world_count,), dt, dtype=torch.float64, device=DEV)
        out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True)
        qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
        ref = O.attempt_all(om, cfg, qn, qdn, fn, dtn)
        assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
        print(f"{name} after {frames} frames, dt {dt}", flush=True)
        P.check_free_attempts(O, om, cfg, out, ref, qn, qdn, fn, dtn, model.host, e2e_err_bar=1e-5,
                              e2e_q_bar=1e-4)


def test_mujoco_served_equals_plain(na):
    model = na.scenes.uneven_pile(64, seed=7, device=DEV)
    a, b = model.state(), model.state()
    sa = na.SolverMuJoCoAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, schedule="served")
    sb = na.SolverMuJoCoAdaptive(model, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, schedule="plain")
    for _ in range(15):
        sa.step(a, a, model.control(), None, 0.01)
        sb.step(b, b, model.control(), None, 0.01)
    torch.cuda.synchronize()
    assert torch.equal(a.body_q, b.body_q) and torch.equal(a.body_qd, b.body_qd)
    assert np.allclose(sa.sim_time.cpu().numpy(), 0.01, rtol=0, atol=1e-12)  # rebased at every call


def test_mujoco_frames_match_oracle_driver(na, O):
    model = na.scenes.box_drop(16, seed=8, device=DEV)
    state, solver = _settle(na, model, 10)
    om = O.OracleModel(model)
    cfg = O.make_config(mode="adaptive", controller="mujoco", contact_law="mujoco", tol=1e-3, dt_init=1e-3,
                        dt_max=0.01, max_substeps=128, max_iterations=100, opt_rel_tol=1e-6)
    S = O.OracleSolver(om, cfg)
    for k in ("dt", "dt_half", "ideal_dt", "sim_time", "next_time"):
        S.a[k][:] = getattr(solver, "_" + k).cpu().numpy()
    q = solver.body_q64.cpu().numpy().copy()
    qd = state.body_qd.cpu().numpy().astype(np.float64)
    for frame in range(3):
        before = solver.accepted_steps()
        solver.step(state, state, model.control(), None, 0.01)
        torch.cuda.synchronize()
        S.step(q, qd, np.zeros_like(qd), 0.01)
        qd = qd.astype(np.float32).astype(np.float64)  # the boundary rounds velocities to float32
        dq = np.abs(solver.body_q64.cpu().numpy() - q).max()
        print(f"frame {frame}: |dq| {dq:.2e}, accepted {solver.accepted_steps() - before}", flush=True)
        assert dq <= 1e-5
