"""The oracle's MuJoCo soft-constraint inner step (contact law 1), the inner step of the reference's
SolverMuJoCoAdaptive (mujoco_warp.step, solver_mujoco_adaptive.py:585-605; mujoco-warp 3.8.1 is
absent, so parity with it is unpinned).  What pins the restatement:

* the reference's own recorded results (results/v1_single_drop.json, copied as data into
  tests/golden/v1_single_drop.json): a sphere driven into a stiff ground contact by SolverMuJoCo
  with the euler integrator.  The gold run (dt 2e-5) and the five fixed steps are reproduced to
  1e-7 m in the lowest sphere height (the reference's float32 rounding is ~1e-8 m).  This fixes
  the reference acceleration (solref time constant and damping ratio with the 2h floor), the
  impedance, the margin convention and the regulariser of a pyramidal contact at mu = 1;
* the impedance curve at hand-computed points, and the pair mixing of solvers/mujoco/kernels.py
  on hand-made priorities, solmix and solref values;
* a frictionless contact with a flat impedance moves as the scalar recurrence of MuJoCo's
  constraint dynamics, a = d (aref) (exactly), and as h -> 0 that recurrence is the damped
  oscillator of natural frequency 1/(timeconst dampratio) and damping ratio dampratio;
* a pyramidal contact at rest holds the penetration where its four rows carry the weight, a fixed
  point of the impedance curve, for mu 0.5 and 1 and impratio 1 and 2;
* the optimality conditions at the oracle's solutions, re-evaluated with an independent
  evaluation of the pyramid rows, on worlds with sliding contacts; an incline below tan = mu holds
  a box, one above does not;
* a planted mistake in the rows (-DOR_PLANT=4: mu dropped from the row direction) fails a check.

What stays unpinned: the regulariser's mu dependence away from mu = 1 (restated from
mujoco_warp's pyramidal rows), and MuJoCo's own collision (this build generates the contacts).
"""

import json
import math
import os

import numpy as np
import pytest

G = 9.81
HERE = os.path.dirname(os.path.abspath(__file__))


def _ball(na, ke=1e5, kd=100.0, r=0.1, mu=0.5, z=None, margin=0.0):
    b = na.ModelBuilder()
    b.begin_world()
    body = b.add_body(xform=(0.0, 0.0, r if z is None else z, 0, 0, 0, 1))
    b.add_shape_sphere(body, radius=r, cfg=na.ModelBuilder.ShapeConfig(ke=ke, kd=kd, mu=mu, gap=0.01, margin=margin))
    b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=ke, kd=kd, mu=mu))
    return b.finalize("cpu")


def _cfg(O, **kw):
    base = dict(mode="fixed", contact_law="mujoco", max_iterations=100, opt_rel_tol=1e-12)
    base.update(kw)
    return O.make_config(**base)


def _run(O, m, q, qd, h, n, cfg=None, sample=1):
    om = O.OracleModel(m)
    cfg = cfg or _cfg(O)
    f = np.zeros_like(qd)
    out = []
    for k in range(n):
        q, qd, st, ok = O.substep_world(om, cfg, 0, q, qd, f, h)
        assert ok
        if (k + 1) % sample == 0:
            out.append((q.copy(), qd.copy()))
    return out


# ---------------------------------------------------------------- the reference's recorded run
def _single_drop(na, O, g, dt, steps_per_frame):
    s = g["scene"]
    b = na.ModelBuilder()
    cfg = na.ModelBuilder.ShapeConfig(ke=s["ke"], kd=s["kd"], kf=0.0, mu=s["mu"], margin=s["margin"])
    body = b.add_body(xform=(0.0, 0.0, s["z0"], 0.0, 0.0, 0.0, 1.0))
    b.add_shape_sphere(body, radius=s["radius"], cfg=cfg)
    b.add_ground_plane()
    m = b.finalize("cpu")
    q = m.host("body_q").astype(np.float64).copy()
    qd = np.zeros((1, 6))
    qd[0, 2] = s["vz"]
    frames = int(round(s["duration"] / s["dt_outer"]))
    out = _run(O, m, q, qd, dt, frames * steps_per_frame, sample=steps_per_frame)
    return min(float(x[0][0, 2]) for x in out)  # sampled at the frame boundaries, as the script does


def test_single_drop_gold_matches_the_reference_record(na, O):
    """Gold run: 20 000 steps of 2e-5 s; the lowest height of the sphere at the 10 ms frames."""
    g = json.load(open(os.path.join(HERE, "golden", "v1_single_drop.json")))
    z = _single_drop(na, O, g, g["gold_dt"], int(round(g["scene"]["dt_outer"] / g["gold_dt"])))
    assert abs(z - g["gold_zmin"]) <= 1e-7, f"z_min {z:.9f} against the recorded {g['gold_zmin']:.9f}"


@pytest.mark.parametrize("row", range(5))
def test_single_drop_fixed_steps_match_the_reference_record(na, O, row):
    g = json.load(open(os.path.join(HERE, "golden", "v1_single_drop.json")))
    r = g["fixed"][row]
    n = int(round(g["scene"]["dt_outer"] / r["fixed_dt"]))
    z = _single_drop(na, O, g, r["fixed_dt"], n)
    err_mm = abs(z - g["gold_zmin"]) * 1e3
    assert abs(err_mm - r["pen_err_mm"]) <= 1e-4, f"dt {r['fixed_dt']}: {err_mm:.7f} mm against {r['pen_err_mm']:.7f}"


# ---------------------------------------------------------------- impedance and mixing
@pytest.mark.parametrize("solimp, pos, want", [
    ((0.9, 0.95, 0.001, 0.5, 2.0), -0.0002, 0.9 + 0.05 * (0.2 ** 2 / 0.5)),     # x = 0.2 < mid: x^p / mid^(p-1)
    ((0.9, 0.95, 0.001, 0.5, 2.0), 0.0007, 0.9 + 0.05 * (1 - 0.3 ** 2 / 0.5)),  # x = 0.7 > mid: 1 - (1-x)^p/(1-mid)^(p-1)
    ((0.9, 0.95, 0.001, 0.5, 2.0), -0.002, 0.95),                               # |pos| >= width: dmax
    ((0.5, 0.9, 0.01, 0.3, 1.0), -0.004, 0.5 + 0.4 * 0.4),                      # power 1: linear
    ((0.8, 0.8, 0.01, 0.5, 3.0), -0.003, 0.8),                                  # dmin == dmax: flat
    ((0.2, 0.6, 0.01, 0.25, 3.0), -0.001, 0.2 + 0.4 * (0.1 ** 3 / 0.25 ** 2)),
    ((0.2, 0.6, 0.01, 0.25, 3.0), -0.006, 0.2 + 0.4 * (1 - 0.4 ** 3 / 0.75 ** 2)),
    ((0.0, 1.5, 0.01, 0.5, 2.0), -0.02, 0.9999),                                # clamped to (1e-4, 0.9999)
])
def test_impedance_curve(O, solimp, pos, want):
    assert O.mj_impedance(solimp, pos) == pytest.approx(want, rel=1e-12, abs=1e-15)


def _mix_model(na, O, priority, solmix, solref, solimp, mu, condim):
    b = na.ModelBuilder()
    b.begin_world()
    for k in range(len(mu)):
        body = b.add_body(xform=(0.3 * k, 0.0, 0.5, 0, 0, 0, 1))
        b.add_shape_sphere(body, radius=0.1, cfg=na.ModelBuilder.ShapeConfig(mu=mu[k]))
    b.end_world()
    m = b.finalize("cpu")
    import torch
    m.mujoco.geom_priority = torch.tensor(priority, dtype=torch.int32)
    m.mujoco.geom_solmix = torch.tensor(solmix, dtype=torch.float32)
    m.mujoco.solref = torch.tensor(solref, dtype=torch.float32)
    m.mujoco.geom_solimp = torch.tensor(solimp, dtype=torch.float32)
    m.mujoco.condim = torch.tensor(condim, dtype=torch.int32)
    return O.OracleModel(m)


def test_pair_mixing(na, O):
    """contact_params (solvers/mujoco/kernels.py:127-190) on hand-made shapes."""
    si = [(0.9, 0.95, 0.001, 0.5, 2.0), (0.5, 0.7, 0.01, 0.3, 3.0), (0.8, 0.9, 0.002, 0.4, 1.0),
          (0.6, 0.99, 0.005, 0.6, 2.0)]
    om = _mix_model(na, O, priority=[0, 0, 2, 0], solmix=[1.0, 3.0, 1.0, 0.0],
                    solref=[(0.02, 1.0), (0.04, 0.5), (0.01, 2.0), (-1000.0, -20.0)], solimp=si,
                    mu=[0.5, 0.8, 0.1, 0.3], condim=[3, 1, 1, 3])
    r = O.mj_contact_params(om, 0, 1)  # equal priority: mix = 1 / (1 + 3) = 0.25
    assert r["solref"] == pytest.approx([0.25 * 0.02 + 0.75 * 0.04, 0.25 * 1.0 + 0.75 * 0.5], rel=1e-6)
    assert r["solimp"] == pytest.approx([0.25 * a + 0.75 * b for a, b in zip(si[0], si[1])], rel=1e-6)
    assert r["mu"] == pytest.approx(0.8, rel=1e-6) and r["condim"] == 3
    r = O.mj_contact_params(om, 1, 2)  # shape 2 has the higher priority and decides alone
    assert r["solref"] == pytest.approx([0.01, 2.0], rel=1e-6)
    assert r["solimp"] == pytest.approx(si[2], rel=1e-6)
    assert r["mu"] == pytest.approx(0.1, rel=1e-6) and r["condim"] == 1
    r = O.mj_contact_params(om, 0, 3)  # solmix 0 on one side: the other side's values; solref min
    assert r["solimp"] == pytest.approx(si[0], rel=1e-6)
    assert r["solref"] == pytest.approx([-1000.0, -20.0], rel=1e-6)
    assert r["mu"] == pytest.approx(0.5, rel=1e-6) and r["condim"] == 3
    om = _mix_model(na, O, priority=[0, 0], solmix=[0.0, 0.0], solref=[(0.02, 1.0), (0.06, 0.5)],
                    solimp=si[:2], mu=[0.0, 0.0], condim=[3, 3])
    r = O.mj_contact_params(om, 0, 1)  # both solmix ~0: 0.5; friction floored at MJ_MINMU
    assert r["solref"] == pytest.approx([0.04, 0.75], rel=1e-6)
    assert r["mu"] == pytest.approx(1e-5, rel=1e-6)


# ---------------------------------------------------------------- constraint dynamics
def _frictionless_flat(na, tc=0.02, dr=0.5, d=0.95, r=0.1, pos0=-0.01):
    import torch
    m = _ball(na, r=r, z=r + pos0)
    m._host["gravity"] = np.zeros(3)
    S = m.shape_count
    m.mujoco.condim = torch.ones(S, dtype=torch.int32)
    m.mujoco.geom_solimp = torch.tensor([[d, d, 0.001, 0.5, 2.0]] * S, dtype=torch.float32)
    m.mujoco.solref = torch.tensor([[tc, dr]] * S, dtype=torch.float32)
    return m


def _recurrence(tc, dr, d, pos0, h, n):
    """MuJoCo's constraint dynamics of one frictionless row at constant impedance d (the
    regulariser (1 - d)/d of the row's inverse weight): while active, a = d aref with
    aref = -B v - K d pos, B = 2/(d tc), K = 1/(d^2 tc^2 dr^2); inactive once the row would pull
    (aref < 0) or separated (pos >= 0); Euler: v += h a, pos += h v."""
    tc = max(tc, 2 * h)
    B, K = 2.0 / (d * tc), 1.0 / (d * d * tc * tc * dr * dr)
    pos, v, out = pos0, 0.0, []
    for _ in range(n):
        aref = -B * v - K * d * pos
        a = d * aref if (pos < 0 and aref > 0) else 0.0
        v += h * a
        pos += h * v
        out.append(pos)
    return np.array(out)


def test_frictionless_row_follows_the_constraint_dynamics(na, O):
    tc, dr, d, r, pos0 = 0.02, 0.5, 0.95, 0.1, -0.01
    m = _frictionless_flat(na, tc, dr, d, r, pos0)
    # the model holds float32 values: use the rounded ones in the recurrence
    tc32, dr32, d32, r32 = (float(np.float32(x)) for x in (tc, dr, d, r))
    q0 = m.host("body_q").astype(np.float32).astype(np.float64)
    for h in (1e-3, 2e-4):
        n = int(round(0.03 / h))
        out = _run(O, m, q0.copy(), np.zeros((1, 6)), h, n)
        z = np.array([x[0][0, 2] for x in out]) - r32
        ref = _recurrence(tc32, dr32, d32, q0[0, 2] - r32, h, n)
        assert np.abs(z - ref).max() <= 1e-12, f"h {h}: {np.abs(z - ref).max():.3e}"


def test_recurrence_is_the_solref_oscillator():
    """As h -> 0 the row's motion is x'' + 2 zeta w0 x' + w0^2 x = 0 with w0 = 1/(tc dr) and
    zeta = dr: the meaning of solref (time constant, damping ratio)."""
    tc, dr, d, pos0 = 0.02, 0.5, 0.95, -0.01
    w0, z = 1.0 / (tc * dr), dr
    wd = w0 * math.sqrt(1 - z * z)
    x = lambda t: pos0 * np.exp(-z * w0 * t) * (np.cos(wd * t) + z / math.sqrt(1 - z * z) * np.sin(wd * t))
    xd = lambda t: -pos0 * w0 / math.sqrt(1 - z * z) * np.exp(-z * w0 * t) * np.sin(wd * t)
    # the row stays active while its reference acceleration pushes: -2 zeta w0 x' - w0^2 x > 0
    tt = np.linspace(0.0, math.pi / wd, 100001)
    t_end = 0.95 * tt[np.argmax(-2 * z * w0 * xd(tt) - w0 * w0 * x(tt) <= 0)]
    errs = []
    for h in (1e-4, 1e-5):
        n = int(t_end / h)
        got = _recurrence(tc, dr, d, pos0, h, n)
        t = h * np.arange(1, n + 1)
        errs.append(np.abs(got - x(t)).max() / abs(pos0))
    assert errs[1] <= 2e-3 and errs[1] < 0.2 * errs[0], errs  # first order in h


def _impedance(si, pos):
    dmin, dmax, width, mid, p = si
    x = abs(pos) / width
    if x >= 1:
        return dmax
    y = x ** p / mid ** (p - 1) if x < mid else 1 - (1 - x) ** p / (1 - mid) ** (p - 1)
    return dmin + y * (dmax - dmin)


@pytest.mark.parametrize("mu, impratio", [(0.5, 1.0), (1.0, 1.0), (0.5, 2.0)])
def test_pyramid_rest_penetration(na, O, mu, impratio):
    """At rest the four rows of a pyramidal contact carry the weight: with y_n = h K imp pos for
    every row, m g h = -4 D h K imp pos, D = 1/R, R = (1 + mu^2) 2 mu^2 / (impratio m) (1 - imp)/imp:
    pos = -g (1 + mu^2) mu^2 (1 - imp) / (2 K imp^2 impratio), imp = imp(pos) (a fixed point)."""
    ke, kd, r = 1e5, 100.0, 0.1
    si = (0.9, 0.95, 0.001, 0.5, 2.0)
    tc, dr = 2.0 / kd, kd / 2.0 * math.sqrt(1.0 / ke)
    K = 1.0 / (0.95 ** 2 * tc * tc * dr * dr)
    f = lambda p: p + G * (1 + mu * mu) * mu * mu * (1 - _impedance(si, p)) / (2 * K * _impedance(si, p) ** 2 * impratio)
    lo, hi = -0.05, -1e-12
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
    pos = 0.5 * (lo + hi)
    r32 = float(np.float32(r))  # the model's (float32) radius
    m = _ball(na, ke=ke, kd=kd, r=r, mu=mu, z=r + pos)
    q = m.host("body_q").astype(np.float64).copy()
    q[0, 2] = r32 + pos
    out = _run(O, m, q, np.zeros((1, 6)), 1e-3, 200, cfg=_cfg(O, impratio=impratio))
    z = np.array([x[0][0, 2] for x in out]) - r32
    v = np.array([x[1][0] for x in out])
    assert np.abs(z - pos).max() <= 1e-9 * (1 + abs(pos) * 1e3), (pos, np.abs(z - pos).max())
    assert np.abs(v).max() <= 1e-8


# ---------------------------------------------------------------- friction: independent checks
def _pyramid_gamma(y, mu, D):
    """Impulse of the pyramid rows of one contact: minus the gradient of
    sum_rows 1/2 D min(c.y, 0)^2, rows c = n + s mu t_k (mu == 0: n alone)."""
    rows = [np.array([0.0, 0.0, 1.0])] if mu == 0 else \
        [np.array([s * mu * (k == 0), s * mu * (k == 1), 1.0]) for k in (0, 1) for s in (1, -1)]
    g = np.zeros(3)
    for c in rows:
        x = c @ y
        if x < 0:
            g -= D * x * c
    return g


def _slab(na, theta_tilt=None, ke=1e5, mu=0.5, hx=0.15, hz=0.02):
    b = na.ModelBuilder()
    b.begin_world()
    body = b.add_body(xform=(0.0, 0.0, hz - 2e-4, 0, 0, 0, 1))
    b.add_shape_box(body, hx=hx, hy=hx, hz=hz, cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu, gap=0.01))
    b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu))
    return b.finalize("cpu")


def _incline(theta):
    return np.array([G * math.sin(theta), 0.0, -G * math.cos(theta)])


def _sliding_worlds(na):
    out = []
    m = _ball(na, z=0.1 - 2e-5)
    out.append((m, np.array([0.0, 0.0, -G]), m.host("body_q").astype(np.float64).copy(),
                np.array([[0.8, 0.3, -0.02, 0.0, 2.0, 0.5]])))
    m = _slab(na)
    out.append((m, _incline(math.radians(35)), m.host("body_q").astype(np.float64).copy(),
                np.array([[0.6, -0.1, 0.0, 0.1, 0.2, 0.0]])))
    b = na.ModelBuilder()
    b.begin_world()
    cfg = na.ModelBuilder.ShapeConfig(ke=1e5, mu=0.4, gap=0.01)
    for k in range(2):
        body = b.add_body(xform=(0.02 * k, 0.0, 0.05 + 0.1 * k - 2e-4 * (k + 1), 0, 0, 0, 1))
        b.add_shape_box(body, hx=0.05, hy=0.05, hz=0.05, cfg=cfg)
    b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=1e5, mu=0.4))
    m = b.finalize("cpu")
    out.append((m, np.array([0.0, 0.0, -G]), m.host("body_q").astype(np.float64).copy(),
                np.array([[0.7, 0.0, 0.0, 0.0, 0.0, 0.3], [0.0, 0.4, 0.0, 0.0, 0.0, 0.0]])))
    return out


def check_pyramid_optimality_independent(na, O):
    """g = A (v - v*) - J^T gamma(v), gamma from the rows above, meets the optimality criterion at
    the oracle's solutions (1e-10 relative; the bar is 1e-8) with a sliding contact in each world."""
    for m, g, q, qd in _sliding_worlds(na):
        m._host["gravity"] = np.asarray(g, dtype=np.float64)
        om = O.OracleModel(m)
        P = O.substep_problem(om, _cfg(O, opt_rel_tol=1e-10), 0, q, qd, np.zeros_like(qd), 1e-3)
        assert P["converged"] and len(P["mu"]) > 0
        A, v, J = P["A"], P["v"], P["J"]
        jt = np.zeros_like(v)
        slide = 0
        for c in range(len(P["mu"])):
            w = J[3 * c:3 * c + 3] @ v
            slide += math.hypot(w[0], w[1]) > 1e-3 and P["D"][c] > 0
            jt += J[3 * c:3 * c + 3].T @ _pyramid_gamma(w - P["what"][c], P["mu"][c], P["D"][c])
        assert slide > 0
        gr = A @ (v - P["vstar"]) - jt
        D = 1.0 / np.sqrt(np.diag(A))
        gn = np.linalg.norm(D * gr)
        scale = max(np.linalg.norm(D * (A @ v)), np.linalg.norm(D * jt))
        assert gn <= 1e-12 + 1e-8 * scale, f"|D g| = {gn:.3e} against {1e-8 * scale:.3e}"


def check_incline_hold_and_slide(na, O):
    """The slide direction of this incline is a pyramid axis (MuJoCo's contact frame), whose rows
    bound the friction at mu gamma_n: at tan(theta) < mu a box only creeps (soft friction rows
    hold it at a bounded slip velocity that does not grow), at tan(theta) > mu it accelerates at
    close to g (sin - mu cos)."""
    h = 1e-3
    for theta, held in ((math.radians(20), True), (math.radians(35), False)):
        m = _slab(na)
        m._host["gravity"] = _incline(theta)
        out = _run(O, m, m.host("body_q").astype(np.float64).copy(), np.zeros((1, 6)), h, 400)
        vx = np.array([x[1][0, 0] for x in out])
        if held:
            assert np.abs(vx[100:]).max() <= 0.02, f"held box moves at {np.abs(vx[100:]).max():.2e} m/s"
            assert vx[300:].mean() <= 1.05 * vx[100:200].mean() + 1e-4, "held box accelerates"
        else:
            a = (vx[-1] - vx[199]) / (200 * h)
            a_coulomb = G * (math.sin(theta) - 0.5 * math.cos(theta))
            assert abs(a / a_coulomb - 1) <= 0.1, f"sliding box accelerates at {a:.3f} against {a_coulomb:.3f}"


CHECKS = [check_pyramid_optimality_independent, check_incline_hold_and_slide]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_pyramid_friction_pin(na, O, check):
    check(na, O)


def test_planted_row_mistake_is_caught(na, O):
    """An oracle whose pyramid rows drop mu from the row direction (-DOR_PLANT=4) fails a check."""
    path = O.build_planted(4)
    failed = []
    with O.use_library(path):
        for check in CHECKS:
            try:
                check(na, O)
            except AssertionError:
                failed.append(check.__name__)
    assert failed, "the planted row mistake passed every check"
