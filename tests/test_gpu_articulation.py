"""Articulated worlds (7-dof arms, floating-base quadrupeds on box terrain) on the GPU against
the FP64 generalized-coordinate oracle.

* one inner step per world (``na_art_substep``): coordinates within 1e-6 (metres / radians /
  quaternion components), velocities within 100x the oracle's own response to a float32-size
  perturbation of the coordinates, at least 1e-5 and at most 1e-4 relative (the largest response
  and every world above 1e-5 are printed), converged solves and equal contact counts;
* one step-doubling attempt (``na_art_attempt``): error norms within 1e-7 absolute plus 1e-3
  relative, doubled coordinates within 1e-6;
* the served boundary call (``na_step_articulated``) against the oracle's frame driver over several
  frames: equal accepted-step counts and clocks, coordinates within 1e-5.

States come from an oracle rollout (fixed 4 ms attempts committing the doubled result), so that
feet, links and terrain are in contact; both sides receive the same float32-rounded values.
"""

import numpy as np
import pytest
import torch

DEV = "cuda:0"
_ROLL = {}


def _model(na, worlds=16, seed=2):
    return na.scenes.articulated_mix(worlds, seed=seed, device=DEV)


def _rollout(O, model, attempts, dt=4e-3):
    key = (model.world_count, attempts, dt)
    if key not in _ROLL:
        om, oa = O.OracleModel(model), O.OracleArt(model)
        cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-8, max_iterations=100)
        q = model.host("joint_q").astype(np.float32).astype(np.float64)
        v = np.zeros(model.joint_dof_count)
        for _ in range(attempts):
            r = O.art_attempt_all(om, oa, cfg, q, v, np.full(model.world_count, dt))
            assert r["solve_ok"].min() == 1
            q, v = r["q_double"], r["v_double"]
        _ROLL[key] = (q.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64))
    return _ROLL[key]


def _t(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _art_substep(nalib, solver, q, v, h, body_f=None, control=None):
    W = solver.model.world_count
    qt, vt = _t(q), _t(v)
    qo, vo = torch.zeros_like(qt), torch.zeros_like(vt)
    st = torch.zeros((W, 4), dtype=torch.int32, device=DEV)
    hh = torch.full((W,), h, dtype=torch.float64, device=DEV)
    bf = _t(body_f, torch.float32) if body_f is not None else None
    ctl = solver._control_desc(control)
    nalib.check(nalib.lib().na_art_substep(solver._h, qt.data_ptr(), vt.data_ptr(), hh.data_ptr(),
                                           bf.data_ptr() if bf is not None else None, ctl, qo.data_ptr(),
                                           vo.data_ptr(), st.data_ptr(), nalib.stream_handle()), "na_art_substep")
    torch.cuda.synchronize()
    return qo.cpu().numpy(), vo.cpu().numpy(), st.cpu().numpy()


def _art_attempt(nalib, solver, q, v, dt, body_f=None, control=None, with_contacts=False):
    W = solver.model.world_count
    qt, vt = _t(q), _t(v)
    out = {k: torch.zeros_like(qt) for k in ("q_full", "q_double")}
    out.update({k: torch.zeros_like(vt) for k in ("v_full", "v_double")})
    err = torch.zeros(W, dtype=torch.float64, device=DEV)
    ok = torch.zeros(W, dtype=torch.int32, device=DEV)
    nc = torch.zeros(W, dtype=torch.int32, device=DEV)
    dtt = torch.full((W,), dt, dtype=torch.float64, device=DEV)
    bf = _t(body_f, torch.float32) if body_f is not None else None
    ctl = solver._control_desc(control)
    cs = torch.zeros((W, 2, 128, 16), dtype=torch.float32, device=DEV) if with_contacts else None
    cn = torch.zeros((W, 2), dtype=torch.int32, device=DEV) if with_contacts else None
    nalib.check(nalib.lib().na_art_attempt(solver._h, qt.data_ptr(), vt.data_ptr(), dtt.data_ptr(),
                                           bf.data_ptr() if bf is not None else None, ctl,
                                           out["q_full"].data_ptr(), out["v_full"].data_ptr(),
                                           out["q_double"].data_ptr(), out["v_double"].data_ptr(), err.data_ptr(),
                                           ok.data_ptr(), nc.data_ptr(), cs.data_ptr() if with_contacts else None,
                                           cn.data_ptr() if with_contacts else None, nalib.stream_handle()),
                "na_art_attempt")
    torch.cuda.synchronize()
    r = {k: t.cpu().numpy() for k, t in out.items()}
    r.update(err=err.cpu().numpy(), solve_ok=ok.cpu().numpy(), n_contacts=nc.cpu().numpy())
    if with_contacts:
        r.update(contacts=cs.cpu().numpy(), contact_counts=cn.cpu().numpy())
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("attempts", [0, 40])
def test_art_substep_matches_oracle(na, O, nalib, attempts):
    model = _model(na)
    q, v = _rollout(O, model, attempts)
    print(f"articulated substep: worlds={model.world_count} rollout attempts={attempts}", flush=True)
    solver = na.SolverSAPAdaptive(model, tol=1e-3, optimality_rel_tol=1e-9, max_iterations=200)
    qo, vo, st = _art_substep(nalib, solver, q, v, 2e-3)
    assert st[:, 2].min() == 1, f"GPU solves not converged: {st.tolist()}"
    om, oa = O.OracleModel(model), O.OracleArt(model)
    cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-12, max_iterations=200)
    rng = np.random.default_rng(0)
    worst, loose = 0.0, []
    for w in range(model.world_count):
        q0, q1 = oa.world_q_start[w], oa.world_q_start[w + 1]
        d0, d1 = oa.world_qd_start[w], oa.world_qd_start[w + 1]
        rq, rv, rst, ok = O.art_substep(om, oa, cfg, w, q[q0:q1], v[d0:d1], 2e-3)
        assert ok and rst[3] == st[w, 3], f"world {w}: contacts {st[w, 3]} vs {rst[3]}"
        dq = np.abs(qo[q0:q1] - rq).max()
        assert dq <= 1e-6, f"world {w}: dq={dq:.2e}"
        vscale = max(np.abs(rv).max(), 1.0)
        dv = np.abs(vo[d0:d1] - rv).max() / vscale
        sens = 0.0
        for _ in range(2):
            qp = q[q0:q1] * (1.0 + rng.normal(size=q1 - q0) * 6e-8)
            _, pv, pst, _ = O.art_substep(om, oa, cfg, w, qp, v[d0:d1], 2e-3)
            if pst[3] == rst[3]:
                sens = max(sens, np.abs(pv - rv).max() / vscale)
        worst = max(worst, sens)
        if dv > 1e-5:
            loose.append((w, float(dv), float(sens)))
        assert dv <= min(max(1e-5, 100.0 * sens), 1e-4), f"world {w}: dv={dv:.2e} oracle sensitivity={sens:.2e}"
    print(f"articulated substep ({attempts} rollout attempts): max oracle sensitivity s_w = {worst:.2e}; worlds "
          f"with dv > 1e-5 (world, dv, s_w): {[(w, f'{a:.1e}', f'{b:.1e}') for w, a, b in loose]}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [1e-3, 4e-3])
def test_art_attempt_matches_oracle(na, O, nalib, dt):
    model = _model(na)
    q, v = _rollout(O, model, 40)
    print(f"articulated attempt: worlds={model.world_count} dt={dt}", flush=True)
    solver = na.SolverSAPAdaptive(model, tol=1e-3)
    out = _art_attempt(nalib, solver, q, v, dt)
    om, oa = O.OracleModel(model), O.OracleArt(model)
    ref = O.art_attempt_all(om, oa, O.make_config(tol=1e-3, opt_rel_tol=1e-6), q, v, np.full(model.world_count, dt))
    assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
    assert np.array_equal(out["n_contacts"], ref["n_contacts"])
    de = np.abs(out["err"] - ref["err"])
    assert np.all(de <= 1e-7 + 1e-3 * ref["err"]), f"err gpu {out['err']} oracle {ref['err']}"
    dq = np.abs(out["q_double"] - ref["q_double"]).max()
    assert dq <= 1e-6, f"max |dq| = {dq:.2e}"


@pytest.mark.gpu
def test_art_served_frames_match_oracle_driver(na, O):
    model = _model(na, worlds=16, seed=7)
    q, v = _rollout(O, model, 30)
    print(f"articulated served frames: worlds={model.world_count}", flush=True)
    state = model.state()
    state.joint_q.copy_(_t(q, torch.float32))
    state.joint_qd.copy_(_t(v, torch.float32))
    solver = na.SolverSAPAdaptive(model, tol=1e-3, dt_inner_init=2e-3, max_substeps=64)
    om, oa = O.OracleModel(model), O.OracleArt(model)
    cfg = O.make_config(tol=1e-3, dt_init=2e-3, max_substeps=64, opt_rel_tol=1e-6)
    ref = O.OracleSolver(om, cfg)
    rq, rv = q.copy(), v.copy()
    for frame in range(3):
        solver.step(state, state, model.control(), None, 0.01)
        assert ref.step_articulated(oa, rq, rv, 0.01) == 0
        # the device hands float32 velocities across frames: give the oracle the same
        rv[:] = rv.astype(np.float32).astype(np.float64)
        torch.cuda.synchronize()
        assert np.allclose(solver.sim_time.cpu().numpy(), 0.01, rtol=0, atol=1e-12)
        sub = solver.substeps.cpu().numpy()
        assert np.array_equal(sub, ref.a["substeps"]), f"frame {frame}: {sub} vs {ref.a['substeps']}"
        dq = np.abs(state.joint_q.cpu().numpy().astype(np.float64) - rq).max()
        assert dq <= 1e-5, f"frame {frame}: max |dq| = {dq:.2e}"
    # body poses follow the coordinates (forward kinematics of the final joint_q)
    bq = state.body_q.cpu().numpy()
    for w in (0, 1):
        ref_bq = O.art_fk(om, oa, w, solver.body_q64.view(-1)[oa.world_q_start[w]:oa.world_q_start[w + 1]].cpu().numpy())
        b0, b1 = model.host("world_body_start")[w:w + 2]
        sign = np.sign(np.sum(bq[b0:b1, 3:] * ref_bq[:, 3:], axis=1, keepdims=True))
        assert np.abs(bq[b0:b1, :3] - ref_bq[:, :3]).max() <= 1e-5
        assert np.abs(bq[b0:b1, 3:] - sign * ref_bq[:, 3:]).max() <= 1e-5


def _floating_boxes(na, worlds, seed, with_pendulum):
    """Box worlds: each box is a FREE root; with_pendulum adds a far revolute pendulum per world
    (which makes the model articulated without touching the box)."""
    b = na.ModelBuilder()
    for w in range(worlds):
        rng = np.random.default_rng(seed + w)
        b.begin_world()
        qq = na.scenes.random_unit_quaternion(rng)
        body = b.add_body(xform=(0.0, 0.0, 0.051, *qq)) if not with_pendulum else b.add_link(
            xform=(0.0, 0.0, 0.051, *qq))
        b.add_shape_box(body, hx=0.05, hy=0.05, hz=0.05, cfg=na.ModelBuilder.ShapeConfig(ke=1e4, mu=0.5))
        if with_pendulum:
            b.add_joint_free(body)
            far = b.add_link(xform=(5.0, 5.0, 5.0, 0, 0, 0, 1))
            b.add_shape_sphere(far, radius=0.05)
            b.add_joint_revolute(-1, far, parent_xform=(5.0, 5.0, 5.0, 0, 0, 0, 1))
        b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=1e4, mu=0.5))
    return b.finalize(DEV)


@pytest.mark.gpu
def test_floating_base_kernel_equals_free_body_kernel(na, nalib):
    """A box on a FREE root joint steps through the articulated kernel (dense Cholesky Newton) as
    the same box does through the free-body kernel (PCG Newton), within the solves' tolerance."""
    W = 16
    free = _floating_boxes(na, W, 40, False)
    art = _floating_boxes(na, W, 40, True)
    print("floating base vs free body", flush=True)
    rng = np.random.default_rng(1)
    v6 = rng.normal(size=(W, 6)) * np.array([0.3, 0.3, 0.5, 1.0, 1.0, 1.0])
    v6[:, 2] = -np.abs(v6[:, 2])  # moving into the ground
    fs = na.SolverSAPAdaptive(free, tol=1e-3, optimality_rel_tol=1e-10, max_iterations=200)
    q64 = free.body_q.double().contiguous()
    qd = torch.as_tensor(v6, dtype=torch.float32, device=DEV).contiguous()
    f = torch.zeros_like(qd)
    qo = torch.zeros_like(q64)
    vo = torch.zeros_like(qd)
    st = torch.zeros((W, 6), dtype=torch.int32, device=DEV)
    hh = torch.full((W,), 2e-3, dtype=torch.float64, device=DEV)
    nalib.check(nalib.lib().na_substep(fs._h, q64.data_ptr(), qd.data_ptr(), f.data_ptr(), hh.data_ptr(), None,
                                       qo.data_ptr(), vo.data_ptr(), None, st.data_ptr(), nalib.stream_handle()),
                "substep")
    asol = na.SolverSAPAdaptive(art, tol=1e-3, optimality_rel_tol=1e-10, max_iterations=200)
    qa = art.joint_q.double().cpu().numpy().copy()
    va = np.zeros(art.joint_dof_count)
    qs, ds = art.host("joint_q_start"), art.host("joint_qd_start")
    wb = art.host("world_body_start")
    for w in range(W):
        b = int(wb[w])
        qa[qs[b]:qs[b] + 7] = q64[w].cpu().numpy()
        va[ds[b]:ds[b] + 6] = qd[w].cpu().numpy()
    aq, av, ast = _art_substep(nalib, asol, qa, va, 2e-3)
    torch.cuda.synchronize()
    assert st[:, 2].min().item() == 1 and ast[:, 2].min() == 1
    assert np.array_equal(st[:, 3].cpu().numpy(), ast[:, 3]), "contact counts"
    qo, vo = qo.cpu().numpy(), vo.cpu().numpy()
    for w in range(W):
        b = int(wb[w])
        assert np.abs(aq[qs[b]:qs[b] + 7] - qo[w]).max() <= 1e-6, f"world {w}: pose"
        assert np.abs(av[ds[b]:ds[b] + 6] - vo[w]).max() <= 1e-4 * max(1.0, np.abs(vo[w]).max()), f"world {w}: velocity"


@pytest.mark.gpu
def test_articulated_model_layout_is_validated(na):
    """na_model_create rejects articulated layouts the wave-per-world kernels cannot hold."""
    b = na.ModelBuilder()
    b.begin_world()
    parent = -1
    for k in range(17):  # 17 bodies in one world
        link = b.add_link()
        b.add_shape_sphere(link, radius=0.02)
        b.add_joint_revolute(parent, link, parent_xform=(0, 0, 0.1, 0, 0, 0, 1))
        parent = link
    b.end_world()
    with pytest.raises(RuntimeError, match="more than 16 bodies"):
        na.SolverSAPAdaptive(b.finalize(DEV))
    m = na.scenes.articulated_mix(2, seed=1, device=DEV)
    m.joint_type[3] = 2  # an unsupported joint type
    with pytest.raises(RuntimeError, match="joint types"):
        na.SolverSAPAdaptive(m)
    m = na.scenes.articulated_mix(2, seed=1, device=DEV)
    m.joint_parent[7] = 8  # quadruped trunk (body 7) attached to a later body
    with pytest.raises(RuntimeError, match="FREE joint must have the world as parent"):
        na.SolverSAPAdaptive(m)
    m = na.scenes.articulated_mix(2, seed=1, device=DEV)
    m.joint_parent[8] = 9  # a quadruped hip attached to its own (later) knee link
    with pytest.raises(RuntimeError, match="parent must precede its child"):
        na.SolverSAPAdaptive(m)
