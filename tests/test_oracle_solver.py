"""Oracle convex solve / integrator / frame loop pinned by closed forms and invariants.

The reference's inner solve (external ``sap_warp``) is absent, so the restated SAP problem is
pinned by what the mathematics fixes: free flight is the closed-form semi-implicit update
(``solver.py:50-94``); contact impulses are internal (momentum conservation between two free
bodies); a resting box carries its weight; the optimality residual meets the SAP criterion; a
warm start at the solution needs no factorisation (``test_adaptive_warmstart_gate.py:66-130``); a
drop never exceeds sqrt(2 g h) (``test_rigid_contact.py:766``); the frame loop lands on the
boundary, is deterministic, does more work at a tighter tolerance and constant work in fixed mode
(``test_adaptive_sap.py:61-140``).
"""

import math

import numpy as np
import pytest


def _state(m):
    q = m.host("body_q").astype(np.float32).astype(np.float64).copy()
    return q, np.zeros((m.body_count, 6)), np.zeros((m.body_count, 6))


def test_free_flight_closed_form(na, O):
    b = na.ModelBuilder()
    b.begin_world()
    body = b.add_body(xform=(0.1, -0.2, 5.0, 0.0, 0.0, math.sin(0.2), math.cos(0.2)))
    b.add_shape_box(body, hx=0.1, hy=0.2, hz=0.3)
    b.end_world()
    b.add_ground_plane()
    m = b.finalize("cpu")
    om = O.OracleModel(m)
    q, qd, f = _state(m)
    qd[0] = [0.3, -0.1, 0.5, 0.0, 0.0, 2.0]  # spin about the principal z axis: no gyroscopic torque
    h = 1e-3
    qo, qdo, st, ok = O.substep_world(om, O.make_config(), 0, q, qd, f, h)
    assert ok and st[3] == 0
    g = np.array([0.0, 0.0, -9.81], dtype=np.float32).astype(np.float64)
    assert np.allclose(qdo[0, :3], qd[0, :3] + h * g, atol=1e-12)
    assert np.allclose(qdo[0, 3:], qd[0, 3:], atol=1e-12)
    assert np.allclose(qo[0, :3], q[0, :3] + h * qdo[0, :3], atol=1e-12)
    # quaternion: normalize(q + h/2 (w,0) q) -> yaw advanced by ~h*wz
    yaw0 = 2 * math.atan2(q[0, 5], q[0, 6])
    yaw1 = 2 * math.atan2(qo[0, 5], qo[0, 6])
    assert yaw1 - yaw0 == pytest.approx(2 * math.atan(0.5 * h * 2.0), abs=1e-12)


def test_contact_impulses_conserve_momentum(na, O):
    """Two spheres meeting head-on without gravity: the contact impulse is internal."""
    b = na.ModelBuilder(gravity=0.0)
    b.begin_world()
    s0 = b.add_body(xform=(0.0, 0.0, 1.0, 0, 0, 0, 1))
    b.add_shape_sphere(s0, radius=0.1, cfg=na.ModelBuilder.ShapeConfig(ke=1e5, mu=0.6))
    s1 = b.add_body(xform=(0.195, 0.02, 1.0, 0, 0, 0, 1))
    b.add_shape_sphere(s1, radius=0.1, cfg=na.ModelBuilder.ShapeConfig(density=3000.0, ke=1e5, mu=0.6))
    b.end_world()
    m = b.finalize("cpu")
    om = O.OracleModel(m)
    q, qd, f = _state(m)
    qd[0, :3] = [1.0, 0.2, 0.0]
    qd[1, :3] = [-1.0, 0.0, 0.1]
    mass = m.host("body_mass").astype(np.float32).astype(np.float64)
    p0 = mass[0] * qd[0, :3] + mass[1] * qd[1, :3]
    qo, qdo, st, ok = O.substep_world(om, O.make_config(), 0, q, qd, f, 2e-3)
    assert ok and st[3] == 1
    p1 = mass[0] * qdo[0, :3] + mass[1] * qdo[1, :3]
    assert np.allclose(p1, p0, atol=1e-12)
    assert qdo[0, 0] < qd[0, 0] and qdo[1, 0] > qd[1, 0]  # they push apart


def test_resting_box_carries_its_weight_and_solution_is_optimal(na, O):
    m = na.scenes.box_drop(1, seed=0, device="cpu")
    om = O.OracleModel(m)
    cfg = O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=256)
    s = O.OracleSolver(om, cfg)
    q, qd, f = _state(m)
    for _ in range(150):
        assert s.step(q, qd, f, 0.01) == 0
    assert np.abs(qd).max() < 1e-4  # at rest
    h = 1e-3
    qo, qdo, st, ok = O.substep_world(om, cfg, 0, q, qd, f, h)
    assert ok and st[2] == 1 and st[3] == 4
    # m (v1 - v0) = m g h + sum gamma  ->  contacts carry the weight when v stays ~0
    mass = float(m.host("body_mass")[0])
    assert abs(qdo[0, 2] - qd[0, 2]) < 1e-6
    # penetration of a compliant rest: 4 corners share m g, phi ~ -m g / (4 k) (within the
    # dissipation-shifted equilibrium); box half extent 0.05
    z = q[0, 2]
    assert 0.05 - 9.81 * mass / (4 * 1e4) * 3 < z < 0.05


def test_warm_start_at_solution_needs_no_factorization(na, O):
    m = na.scenes.box_stack(2, seed=1, device="cpu")
    om = O.OracleModel(m)
    cfg = O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=256)
    s = O.OracleSolver(om, cfg)
    q, qd, f = _state(m)
    for _ in range(5):
        s.step(q, qd, f, 0.01)
    h = 2e-3
    qo, qdo, st, ok = O.substep_world(om, cfg, 0, q, qd, f, h)
    assert ok and st[1] >= 1
    b1 = int(om.world_body_start[1])
    qo2, qdo2, st2, ok2 = O.substep_world(om, cfg, 0, q, qd, f, h, guess=qdo)
    assert ok2 and st2[1] == 0 and np.abs(qdo2 - qdo).max() <= 1e-9
    # perturbed start factorises and reconverges to the same solution
    qo3, qdo3, st3, ok3 = O.substep_world(om, cfg, 0, q, qd, f, h, guess=qdo + 0.2)
    assert ok3 and st3[1] >= 1 and np.abs(qdo3 - qdo).max() <= 1e-6
    assert b1 == 4


def test_drop_speed_never_exceeds_free_fall(na, O):
    m = na.scenes.box_drop(4, seed=5, device="cpu")
    om = O.OracleModel(m)
    s = O.OracleSolver(om, O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=256))
    q, qd, f = _state(m)
    z0 = q[:, 2].copy()
    vmax = np.sqrt(2 * 9.81 * (z0 + 0.05)) * 1.01 + 1e-3
    for _ in range(80):
        s.step(q, qd, f, 0.01)
        assert np.all(np.linalg.norm(qd[:, :3], axis=1) <= vmax)
    assert np.all(q[:, 2] > 0.03)  # nothing tunnels through the ground


def test_frame_lands_on_boundary_and_is_deterministic(na, O):
    m = na.scenes.uneven_pile(3, seed=2, device="cpu", max_prims=6)
    om = O.OracleModel(m)
    runs = []
    for _ in range(2):
        s = O.OracleSolver(om, O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=256))
        q, qd, f = _state(m)
        for _ in range(3):
            assert s.step(q, qd, f, 1.0 / 60.0) == 0
            assert np.allclose(s.a["sim_time"], 1.0 / 60.0, rtol=1e-12, atol=1e-15)
        runs.append(q.copy())
    assert np.array_equal(runs[0], runs[1])


def test_tighter_tol_uses_more_steps_and_fixed_mode_is_constant(na, O):
    m = na.scenes.box_drop(4, seed=8, device="cpu")
    om = O.OracleModel(m)

    def run(tol, mode="adaptive"):
        s = O.OracleSolver(om, O.make_config(tol=tol, mode=mode, dt_init=0.0025, dt_max=0.01, max_substeps=256))
        q, qd, f = _state(m)
        per = []
        for _ in range(40):
            s.step(q, qd, f, 0.01)
            per.append(int(s.a["substeps"].sum()))
        return int(s.a["cum_accepted"][0]), per

    loose, _ = run(1e-2)
    tight, _ = run(1e-5)
    assert tight >= loose
    _, per = run(1e-3, mode="fixed")
    assert len(set(per)) == 1 and per[0] == 4 * 4


def test_box_drop_1000_accepted_steps_comes_to_rest(na, O):
    """The CPU-runnable small case: 64 one-box worlds (0.1 m, 1 kg, z0 ~ U[0.3, 0.8], random
    orientation, mu 0.5) at tol 1e-3 from dt0 = 1e-3 until every world has taken 1000 accepted
    steps.  Every box ends at rest on a face: COM height = half extent (within the contact
    compliance), no velocity, one box axis vertical."""
    m = na.scenes.box_drop(64, seed=0, device="cpu")
    om = O.OracleModel(m)
    s = O.OracleSolver(om, O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=256,
                                         opt_rel_tol=1e-6))
    q, qd, f = _state(m)
    per_world = np.zeros(64, dtype=np.int64)
    frames = 0
    while per_world.min() < 1000:
        assert s.step(q, qd, f, 0.01) == 0
        per_world += s.a["substeps"]
        frames += 1
        assert frames < 5000
    com_z = q[:, 2]  # the body origin is the box centre
    assert np.all(np.abs(com_z - 0.05) < 2e-3), com_z
    assert np.abs(qd).max() < 1e-3
    x, y, z, w = q[:, 3], q[:, 4], q[:, 5], q[:, 6]
    # third row of the rotation matrix: world z of the box axes
    rz = np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)
    assert np.all(np.abs(rz).max(axis=1) > 1 - 1e-4)


def _records(O, om, model, w, q_all):
    """The oracle's own contacts of world w at q_all as device-style records [128, 16]."""
    n, _, rows = O.collide_world(om, w, q_all, cap=128)
    sb = model.host("shape_body")
    rec = np.zeros((128, 16))
    for i, r in enumerate(rows):
        a, b = int(r[0]), int(r[1])
        rec[i, :5] = [a, b, sb[a], sb[b], r[2]]
        rec[i, 5:8], rec[i, 8:11] = r[3:6], r[6:9]
        rec[i, 11:15] = [r[9], r[10], r[11], r[12]]
    return rec, min(n, 128)


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
def test_attempt_on_given_contacts_equals_own_collision(na, O, mode):
    """Given the contact sets its own collision produces at q and at the midpoint state, the
    given-contact attempt (the entry the GPU parity tests feed with the device's contacts)
    reproduces the oracle's own attempt bit for bit."""
    m = na.scenes.uneven_pile(6, seed=12, device="cpu", max_prims=8)
    om = O.OracleModel(m)
    q, qd, f = _state(m)
    cfg = O.make_config(mode=mode, tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=128, max_iterations=100)
    drv = O.OracleSolver(om, cfg)
    for _ in range(25):  # land the piles so that there are contacts
        drv.step(q, qd, f, 0.01)
    dt = np.full(m.world_count, 2e-3)
    own = O.attempt_all(om, cfg, q, qd, f, dt)
    W = m.world_count
    contacts = np.zeros((W, 2, 128, 16))
    counts = np.zeros((W, 2), np.int32)
    wb = m.host("world_body_start")
    for w in range(W):
        contacts[w, 0], counts[w, 0] = _records(O, om, m, w, q)
        if mode == "adaptive":  # the midpoint state of the attempt: full step, then the first half step
            qf, vf, _, _ = O.substep_world(om, cfg, w, q, qd, f, dt[w])
            g = 0.5 * (qd[wb[w]:wb[w + 1]] + vf)
            qm, _, _, _ = O.substep_world(om, cfg, w, q, qd, f, 0.5 * dt[w], guess=g)
            qmid = q.copy()
            qmid[wb[w]:wb[w + 1]] = qm
            contacts[w, 1], counts[w, 1] = _records(O, om, m, w, qmid)
    assert counts[:, 0].sum() > 0
    given = O.attempt_all_given(om, cfg, q, qd, f, dt, contacts, counts)
    assert np.array_equal(given["err"], own["err"])
    assert np.array_equal(given["q_double"], own["q_double"]) and np.array_equal(given["q_full"], own["q_full"])
    assert np.array_equal(given["solve_ok"], own["solve_ok"])


def test_trend_of_a_free_fall_is_gravity(na, O):
    """The full step's seed acceleration (warm_start "extrapolate") is the mean acceleration of the
    last accepted step, (v_new - v_old) / step: for boxes falling freely (no contact within the
    frame) it equals gravity, and the reference seeds keep no trend."""
    m = na.scenes.box_drop(4, seed=3, device="cpu")
    om = O.OracleModel(m)
    for seeds in ("extrapolate", "reference"):
        q, qd, f = _state(m)
        drv = O.OracleSolver(om, O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=64,
                                               warm_start=seeds))
        drv.step(q, qd, f, 0.01)
        assert (drv.a["substeps"] > 1).all()
        tr = drv.a["trend"][:m.body_count]
        if seeds == "reference":
            assert not tr.any()
            continue
        g = np.array([0.0, 0.0, -9.81], dtype=np.float32).astype(np.float64)
        assert np.abs(tr[:, :3] - g).max() < 1e-9, tr
        assert np.abs(tr[:, 3:]).max() < 1e-9, tr


def test_extrapolated_seeds_reach_the_reference_seeds_solution(na, O):
    """The seeds move where Newton starts, not the convex problems: an attempt seeded by the
    extrapolation (with the trend of the settling frames) and one seeded as the reference
    (solver_sap_adaptive.py:756-771) agree to the solve tolerance (optimality 1e-10), on settled
    piles where the seeds differ by up to the bodies' velocity changes."""
    m = na.scenes.uneven_pile(6, seed=12, device="cpu", max_prims=12)
    om = O.OracleModel(m)
    q, qd, f = _state(m)
    cfg = O.make_config(tol=1e-3, dt_init=1e-3, dt_max=0.01, max_substeps=128, max_iterations=100)
    drv = O.OracleSolver(om, cfg)
    for _ in range(25):
        drv.step(q, qd, f, 0.01)
    trend = drv.a["trend"][:m.body_count].copy()
    assert np.abs(trend).max() > 1e-3  # the piles still move: the seeds differ
    dt = np.full(m.world_count, 2e-3)
    ext = O.attempt_all(om, cfg, q, qd, f, dt, trend=trend)
    ref = O.attempt_all(om, O.make_config(tol=1e-3, max_iterations=100, warm_start="reference"), q, qd, f, dt)
    assert ext["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
    assert np.abs(ext["q_double"] - ref["q_double"]).max() < 1e-9
    assert np.abs(ext["qd_double"] - ref["qd_double"]).max() < 1e-6
    assert np.abs(ext["err"] - ref["err"]).max() < 1e-9
