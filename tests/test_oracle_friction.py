"""Friction of the oracle's convex contact solve, pinned by what the mechanics fixes.

No reference test pins friction (the reference's solver, sap_warp, is absent), so these checks
use closed forms and an independent re-evaluation, each chosen to fail on a plausible mistake in
the sliding branch of the friction-cone projection:

* impulses measured from momentum changes (not read from the solver): while a contact slides,
  the tangential impulse opposes the slip and |gamma_t| = mu gamma_n (a sphere sliding on the
  ground; a box sliding down an incline, summed over its contacts);
* angular momentum about the contact point: a sphere launched sliding without spin ends rolling
  at exactly 5/7 of its launch speed, whatever the normal force does meanwhile;
* an incline below tan(theta) = mu holds a box (its creep bounded through R_t = sigma w), and a
  sphere rolls down one at a = 5/7 g sin(theta) (tan(theta) < 7 mu / 2);
* an independent evaluation of the SAP gradient at the oracle's solutions, with the friction
  cone projection computed as the textbook Euclidean projection onto a second-order cone in
  R^(1/2)-scaled coordinates (itself checked against the variational characterisation of a
  projection on sampled cone points), meets the
  optimality criterion on three worlds with sliding contacts;
* planted mistakes in the sliding branch (oracle built with -DOR_PLANT=k) fail these checks.

The normal direction of the convex model is compliant, and a sliding contact's normal impulse is
computed as if approaching at mu |v_t| (the convex relaxation), so the normal force of a sliding
body varies; the checks above hold regardless of that history.
"""

import math

import numpy as np
import pytest

G = 9.81
MU = 0.5


def _ball(na, ke=1e5, r=0.1, mu=MU):
    b = na.ModelBuilder()
    b.begin_world()
    body = b.add_body(xform=(0.0, 0.0, r, 0, 0, 0, 1))
    b.add_shape_sphere(body, radius=r, cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu, gap=0.01))
    b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu))
    return b.finalize("cpu")


def _slab(na, ke=1e5, hx=0.15, hz=0.02, mu=MU):
    b = na.ModelBuilder()
    b.begin_world()
    body = b.add_body(xform=(0.0, 0.0, hz, 0, 0, 0, 1))
    b.add_shape_box(body, hx=hx, hy=hx, hz=hz, cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu, gap=0.01))
    b.end_world()
    b.add_ground_plane(cfg=na.ModelBuilder.ShapeConfig(ke=ke, mu=mu))
    return b.finalize("cpu")


def _incline(theta):
    """Gravity of a ground tilted by theta about y: x runs down the slope, z is its normal."""
    return np.array([G * math.sin(theta), 0.0, -G * math.cos(theta)])


def _steps(O, m, gravity, q, qd, n, h=1e-3):
    """n inner steps of the oracle's world 0 from (q, qd) under `gravity`; returns the velocities
    before and after every step and the final pose."""
    m._host["gravity"] = np.asarray(gravity, dtype=np.float64)
    om = O.OracleModel(m)
    cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-10, max_iterations=100)
    f = np.zeros((1, 6))
    vel = [qd[0].copy()]
    for _ in range(n):
        q, qd, st, ok = O.substep_world(om, cfg, 0, q, qd, f, h)
        assert ok
        vel.append(qd[0].copy())
    return np.array(vel), q, om


def _impulses(m, vel, gravity, h=1e-3):
    """Contact impulse of every step from the momentum change beyond free motion."""
    mass = float(m.host("body_mass")[0])
    return mass * (vel[1:, :3] - vel[:-1, :3] - h * np.asarray(gravity)[None, :])


# ---------------------------------------------------------------- checks (also run on planted builds)
def check_sphere_slides_then_rolls(na, O):
    """Sphere launched at 1 m/s without spin, from the compliant equilibrium penetration."""
    r, ke = 0.1, 1e5
    m = _ball(na, ke=ke, r=r)
    mass, I = float(m.host("body_mass")[0]), float(m.host("body_inertia")[0][1, 1])
    q = m.host("body_q").astype(np.float64).copy()
    q[0, 2] = r - mass * G / ke
    qd = np.array([[1.0, 0, 0, 0, 0, 0]])
    g = np.array([0.0, 0.0, -G])
    vel, _, _ = _steps(O, m, g, q, qd, 200)
    gam = _impulses(m, vel, g)
    slip = vel[:-1, 0] - vel[:-1, 4] * r  # contact-point slip velocity before each step
    sliding = (slip > 0.05) & (vel[1:, 0] - vel[1:, 4] * r > 0.05) & (gam[:, 2] > 1e-3 * mass * G * 1e-3)
    assert sliding.sum() >= 10, "the sphere slides for a while"
    assert np.all(gam[sliding, 0] < 0), "friction opposes the slip"
    # |gamma_t| - mu gamma_n against the weight impulse of a step (the solves converge to 1e-10)
    dev = np.abs(-gam[sliding, 0] - MU * gam[sliding, 2]) / (mass * G * 1e-3)
    assert dev.max() <= 1e-6, f"| |gamma_t| - mu gamma_n | up to {dev.max():.2e} m g h"
    assert np.abs(gam[:, 1]).max() <= 1e-9
    # angular momentum about the contact point: m v r + I w is conserved -> rolling at 5/7 v0
    L = mass * vel[:, 0] * r + I * vel[:, 4]
    assert np.abs(L / L[0] - 1).max() <= 2e-3, f"angular momentum about the contact drifts {np.abs(L / L[0] - 1).max()}"
    assert abs(vel[-1, 0] / (5.0 / 7.0) - 1) <= 3e-3, f"rolling speed {vel[-1, 0]:.5f} vs 5/7"
    assert abs(vel[-1, 0] - vel[-1, 4] * r) <= 2e-3, "rolls without slipping at the end"


def check_box_on_incline(na, O):
    """Box released on an incline: held below tan(theta) = mu (creep bounded by the stiction
    regularisation), sliding above it with |gamma_t| = mu gamma_n summed over its contacts."""
    ke, h = 1e5, 1e-3
    # held: tan 20 deg = 0.36 < mu.  Creep: each of the 4 corners carries ~m g sin(theta) h / 4 of
    # tangential impulse, which the stiction regularisation turns into a slip of R_t gamma_t with
    # R_t = sigma w (sigma 1e-3, w <= 50 / kg here): far below 1e-4 m/s
    m = _slab(na, ke=ke)
    g = _incline(math.radians(20))
    q = m.host("body_q").astype(np.float64).copy()
    vel, q1, _ = _steps(O, m, g, q, np.zeros((1, 6)), 300)
    vel2, q2, _ = _steps(O, m, g, q1, vel[-1:].copy(), 300)
    assert np.abs(vel2[:, :3]).max() <= 1e-4, f"held box moves at {np.abs(vel2[:, :3]).max():.2e} m/s"
    assert abs(q2[0, 0] - q1[0, 0]) <= 3e-5, "held box creeps"
    # sliding: tan 35 deg = 0.70 > mu
    g = _incline(math.radians(35))
    vel, _, _ = _steps(O, _slab(na, ke=ke), g, q, np.zeros((1, 6)), 400)
    gam = _impulses(_slab(na, ke=ke), vel, g)
    late = slice(150, 400)
    assert np.all(vel[151:, 0] > 0.1), "the box slides"
    mass = float(_slab(na, ke=ke).host("body_mass")[0])
    assert np.all(gam[late, 0] < 0), "friction opposes the slide"
    dev = np.abs(-gam[late, 0] - MU * gam[late, 2]) / (mass * G * h)
    assert dev.max() <= 1e-6, f"| |gamma_t| - mu gamma_n | up to {dev.max():.2e} m g h"


def check_sphere_rolls_down_incline(na, O):
    """tan(theta) < 7 mu / 2: the sphere rolls without slipping, a = 5/7 g sin(theta)."""
    r, ke, th = 0.1, 1e5, math.radians(30)
    m = _ball(na, ke=ke, r=r)
    mass = float(m.host("body_mass")[0])
    q = m.host("body_q").astype(np.float64).copy()
    q[0, 2] = r - mass * G * math.cos(th) / ke
    vel, _, _ = _steps(O, m, _incline(th), q, np.zeros((1, 6)), 200)
    a = 5.0 / 7.0 * G * math.sin(th)
    assert abs(vel[-1, 0] / (a * 0.2) - 1) <= 3e-3, f"v = {vel[-1, 0]:.5f} vs {a * 0.2:.5f}"
    assert abs(vel[-1, 0] - vel[-1, 4] * r) <= 3e-3, "rolls without slipping"


def soc_project(y, mu, Rt, Rn):
    """argmin over the friction cone {|g_t| <= mu g_n} of 1/2 (g - y)^T R (g - y), R = diag(Rt, Rt,
    Rn): with x = R^(1/2) g the cone is {|x_t| <= m x_n}, m = mu sqrt(Rt / Rn), and the problem is
    the Euclidean projection of z = R^(1/2) y onto that circular cone: z inside -> z; z in the
    polar cone (m |z_t| <= -z_n) -> 0; otherwise its component along the boundary ray
    d = (m z_t / |z_t|, 1) / sqrt(1 + m^2)."""
    s = np.array([math.sqrt(Rt), math.sqrt(Rt), math.sqrt(Rn)])
    z = s * y
    m = mu * math.sqrt(Rt / Rn)
    zt = math.hypot(z[0], z[1])
    if zt <= m * z[2]:
        x = z
    elif m * zt <= -z[2]:
        x = np.zeros(3)
    else:
        d = np.array([m * z[0] / zt, m * z[1] / zt, 1.0]) / math.sqrt(1.0 + m * m)
        x = (z @ d) * d
    return x / s


def _sap_gradient(P):
    """g = A (v - v*) - J^T gamma(v) and the optimality scale of the SAP criterion."""
    A, vs, J, v = P["A"], P["vstar"], P["J"], P["v"]
    jt = np.zeros_like(v)
    for c in range(len(P["mu"])):
        u = J[3 * c:3 * c + 3] @ v
        y = np.array([-u[0] / P["Rt"][c], -u[1] / P["Rt"][c], -(u[2] - P["vhat"][c]) / P["Rn"][c]])
        jt += J[3 * c:3 * c + 3].T @ soc_project(y, P["mu"][c], P["Rt"][c], P["Rn"][c])
    g = A @ (v - vs) - jt
    D = 1.0 / np.sqrt(np.diag(A))
    return np.linalg.norm(D * g), max(np.linalg.norm(D * (A @ v)), np.linalg.norm(D * jt))


def _sliding_worlds(na, O):
    """(model, q, qd) of three worlds with sliding contacts: a sphere, a box on a 35 deg incline,
    and a two-box stack whose bottom box slides."""
    out = []
    m = _ball(na)
    q = m.host("body_q").astype(np.float64).copy()
    q[0, 2] -= float(m.host("body_mass")[0]) * G / 1e5
    out.append((m, np.array([0.0, 0.0, -G]), q, np.array([[0.8, 0.1, -0.02, 0.0, 2.0, 0.5]])))
    m = _slab(na)
    q = m.host("body_q").astype(np.float64).copy()
    q[0, 2] -= 2e-4
    out.append((m, _incline(math.radians(35)), q, np.array([[0.6, -0.1, 0.0, 0.1, 0.2, 0.0]])))
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


def check_sap_optimality_independent(na, O):
    """The SAP gradient at the oracle's solutions, re-evaluated from the problem data with the
    cone projection above, meets the optimality criterion the oracle solved to (1e-10 relative);
    the bar here is 1e-8."""
    for m, g, q, qd in _sliding_worlds(na, O):
        m._host["gravity"] = np.asarray(g, dtype=np.float64)
        om = O.OracleModel(m)
        cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-10, max_iterations=100)
        P = O.substep_problem(om, cfg, 0, q, qd, np.zeros_like(qd), 1e-3)
        assert P["converged"] and len(P["mu"]) > 0
        # at least one contact is sliding at the solution
        slide = 0
        for c in range(len(P["mu"])):
            u = P["J"][3 * c:3 * c + 3] @ P["v"]
            slide += math.hypot(u[0], u[1]) > 1e-3
        assert slide > 0
        gn, scale = _sap_gradient(P)
        assert gn <= 1e-12 + 1e-8 * scale, f"|D g| = {gn:.3e} against {1e-8 * scale:.3e}"


CHECKS = [check_sphere_slides_then_rolls, check_box_on_incline, check_sphere_rolls_down_incline,
          check_sap_optimality_independent]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_friction_pin(na, O, check):
    check(na, O)


def test_soc_projection_is_the_r_norm_projection():
    """The cone projection of the independent check meets the variational characterisation of the
    projection onto a closed convex set in the R inner product: g* lies in the cone and
    <R (y - g*), g - g*> <= 0 for every g of the cone (sampled: interior, boundary, apex, and the
    ray through g*), and is not beaten by any sample."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        y = rng.normal(size=3) * rng.choice([0.1, 1.0, 10.0])
        mu, Rt, Rn = rng.uniform(0.2, 1.2), 10 ** rng.uniform(-4, -1), 10 ** rng.uniform(-2, 1)
        R = np.array([Rt, Rt, Rn])
        got = soc_project(y, mu, Rt, Rn)
        scale = 1e-12 * (1.0 + np.abs(got).max() + np.abs(y).max()) ** 2 * R.max()
        assert mu * got[2] - math.hypot(got[0], got[1]) >= -1e-12 * (1 + np.abs(got).max())
        phi = rng.uniform(0, 2 * math.pi, 400)
        gn = np.abs(rng.normal(size=400)) * (1 + np.abs(y).max())
        frac = np.concatenate([rng.uniform(0, 1, 300), np.ones(100)])  # interior and boundary
        cone = np.stack([mu * gn * frac * np.cos(phi), mu * gn * frac * np.sin(phi), gn], axis=1)
        cone = np.concatenate([cone, [np.zeros(3)], got[None, :] * rng.uniform(0, 3, (50, 1))])
        vi = (cone - got) @ (R * (y - got))
        assert vi.max() <= scale, (y, mu, Rt, Rn, got, vi.max())
        cost = 0.5 * ((cone - y) ** 2 * R).sum(axis=1)
        assert 0.5 * np.sum(R * (got - y) ** 2) <= cost.min() + scale


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_sliding_mistakes_are_caught(na, O, plant):
    """An oracle built with a planted mistake in the sliding branch (1: mu mu_hat dropped from the
    denominator, 2: the mu_hat |y_t| term of the normal impulse dropped, 3: mu^2 for mu in the
    tangential impulse) fails at least one of the checks."""
    path = O.build_planted(plant)
    failed = []
    with O.use_library(path):
        for check in CHECKS:
            try:
                check(na, O)
            except AssertionError:
                failed.append(check.__name__)
    assert failed, f"planted mistake {plant} passed every friction check"
