"""Controller known-answer tests on the CPU oracle.

Cases restate the reference's pure-kernel contract tests:
``newton/tests/test_adaptive_floor_nan_guard.py:53-92`` (Fix A),
``test_adaptive_time_rebase.py:34-76`` (Fix B, bit-exact float32),
``test_adaptive_reset_worlds.py:64-130`` (Fix C, bit-exact),
``test_adaptive_cubic_linesearch.py:36-70`` (cubic Hermite minimiser),
plus the Drake rule's closed form (``solver_mujoco_adaptive.py:120-198``).
"""

import ctypes as C

import numpy as np
import pytest

TOL = 1.0e-3
DT_MIN = 1.0e-6
DIVERGENCE = 1.0e9
SENTINEL = 1.0e10


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def calc_adjusted_f32(O, err, dt, tol=TOL, dt_min=DT_MIN):
    n = len(err)
    err = np.asarray(err, np.float32)
    dt = np.asarray(dt, np.float32)
    ideal = np.zeros(n, np.float32)
    acc = np.zeros(n, np.uint8)
    com = np.zeros(n, np.uint8)
    div = np.zeros(n, np.uint8)
    O.lib().or_calc_adjusted_step_f32(n, _p(err, C.c_float), _p(dt, C.c_float), _p(ideal, C.c_float),
                                      _p(acc, C.c_uint8), _p(com, C.c_uint8), _p(div, C.c_uint8),
                                      C.c_float(tol), C.c_float(dt_min), C.c_float(DIVERGENCE))
    return acc.astype(bool), com.astype(bool), div.astype(bool), ideal


# ---------------------------------------------------------------- Fix A (floor NaN guard)
def test_floor_diverged_is_held_not_committed(O):
    acc, com, div, _ = calc_adjusted_f32(O, [SENTINEL], [DT_MIN])
    assert acc[0] and not com[0] and div[0]


def test_floor_finite_over_tol_commits_progress(O):
    acc, com, div, _ = calc_adjusted_f32(O, [10 * TOL], [DT_MIN])
    assert acc[0] and com[0] and not div[0]


def test_normal_within_tol_commits(O):
    acc, com, div, _ = calc_adjusted_f32(O, [0.5 * TOL], [10 * DT_MIN])
    assert acc[0] and com[0] and not div[0]


def test_above_floor_diverged_rejects_and_retries(O):
    acc, com, div, ideal = calc_adjusted_f32(O, [SENTINEL], [10 * DT_MIN])
    assert not acc[0] and not com[0] and not div[0]
    assert ideal[0] < 10 * DT_MIN


def test_drake_rule_closed_form(O):
    """new = 0.9 dt sqrt(tol/e), deadband (0.9, 1.2) dt keeps dt, clamp to [0.1, 5] dt."""
    dt = np.full(6, 1e-3, np.float32)
    err = np.array([TOL / 4, TOL, 4 * TOL, 1e-12, 1e3, 0.7 * TOL], np.float32)
    acc, com, div, ideal = calc_adjusted_f32(O, err, dt)
    for i in range(6):
        e = np.float32(err[i])
        ns = np.float32(0.9) * dt[i] * np.sqrt(np.float32(TOL) / max(e, np.float32(1e-30)))
        if 0.9 * dt[i] < ns < 1.2 * dt[i]:
            ns = dt[i]
        ns = min(max(ns, np.float32(0.1) * dt[i]), np.float32(5.0) * dt[i])
        assert ideal[i] == pytest.approx(float(ns), rel=1e-6)
        assert acc[i] == (e <= TOL or ns >= dt[i])
    assert ideal[3] == pytest.approx(5e-3, rel=1e-6)  # grow clamp
    assert ideal[4] == pytest.approx(1e-4, rel=1e-6)  # shrink clamp
    assert ideal[5] == np.float32(1e-3)  # 0.9*sqrt(1/0.7) = 1.076 lies in the deadband
    assert not acc[2] and acc[0] and acc[1]


def test_boundary_stalled_world_accepts_without_touching_ideal(O):
    err = np.array([0.0], np.float32)
    dt = np.array([0.0], np.float32)
    n = 1
    ideal = np.array([0.0042], np.float32)
    acc = np.zeros(n, np.uint8); com = np.zeros(n, np.uint8); div = np.zeros(n, np.uint8)
    O.lib().or_calc_adjusted_step_f32(n, _p(err, C.c_float), _p(dt, C.c_float), _p(ideal, C.c_float),
                                      _p(acc, C.c_uint8), _p(com, C.c_uint8), _p(div, C.c_uint8),
                                      C.c_float(TOL), C.c_float(DT_MIN), C.c_float(DIVERGENCE))
    assert acc[0] and com[0] and ideal[0] == np.float32(0.0042)


# ---------------------------------------------------------------- Fix B (float32 rebase)
def rebase_f32(O, sim, nxt):
    sim = np.asarray(sim, np.float32).copy()
    nxt = np.asarray(nxt, np.float32).copy()
    O.lib().or_rebase_time_f32(len(sim), _p(sim, C.c_float), _p(nxt, C.c_float))
    return sim, nxt


def test_rebase_next_time_zeroed(O):
    _, nxt = rebase_f32(O, [1.0e6, 5.0], [1.0e6, 5.0])
    assert np.all(nxt == 0.0)


def test_rebase_remaining_preserved_exactly(O):
    nb = np.float32(1048576.0)
    sb = np.float32(1048576.0) - np.float32(1.0e-3)
    sim, nxt = rebase_f32(O, [sb], [nb])
    assert np.float32(nxt[0]) - np.float32(sim[0]) == np.float32(nb) - np.float32(sb)


def test_rebase_sim_becomes_residual(O):
    sv, nv = [1.0e6, 5.0 + 1.0e-4], [1.0e6, 5.0]
    sim, _ = rebase_f32(O, sv, nv)
    for i in range(2):
        assert sim[i] == np.float32(np.float32(sv[i]) - np.float32(nv[i]))


def test_rebase_first_step_noop_and_overshoot(O):
    sim, nxt = rebase_f32(O, [0.0, 10.0 + 2.0e-6], [0.0, 10.0])
    assert sim[0] == 0.0 and nxt[0] == 0.0
    assert nxt[1] == 0.0 and sim[1] == np.float32(np.float32(10.0 + 2.0e-6) - np.float32(10.0)) and sim[1] > 0


# ---------------------------------------------------------------- Fix C (masked reset)
def reset_f32(O, mask, dt_init, seeds):
    n = len(mask)
    a = {k: np.asarray(v, np.float32).copy() for k, v in seeds.items() if k not in ("diverged", "accepted")}
    b = {k: np.asarray(seeds[k], np.uint8).copy() for k in ("diverged", "accepted")}
    m = np.asarray(mask, np.uint8)
    O.lib().or_reset_worlds_f32(n, _p(m, C.c_uint8), C.c_float(dt_init), _p(a["ideal_dt"], C.c_float),
                                _p(a["dt"], C.c_float), _p(a["dt_half"], C.c_float), _p(a["sim_time"], C.c_float),
                                _p(a["next_time"], C.c_float), _p(b["diverged"], C.c_uint8),
                                _p(b["accepted"], C.c_uint8))
    a.update({k: v.astype(bool) for k, v in b.items()})
    return a


SEED2 = {"ideal_dt": [0.005, 0.005], "dt": [0.005, 0.005], "dt_half": [0.0025, 0.0025], "sim_time": [12.3, 12.3],
         "next_time": [12.3, 12.3], "diverged": [True, True], "accepted": [True, True]}


def test_reset_masked_world_restored(O):
    out = reset_f32(O, [True, False], 0.01, SEED2)
    assert out["ideal_dt"][0] == np.float32(0.01) and out["dt"][0] == np.float32(0.01)
    assert out["dt_half"][0] == np.float32(0.01) * np.float32(0.5)
    assert out["sim_time"][0] == 0.0 and out["next_time"][0] == 0.0
    assert not out["diverged"][0] and not out["accepted"][0]
    assert out["ideal_dt"][1] == np.float32(0.005) and out["sim_time"][1] == np.float32(12.3)
    assert out["diverged"][1] and out["accepted"][1]


def test_reset_all_false_is_noop(O):
    out = reset_f32(O, [False, False], 0.01, SEED2)
    for i in range(2):
        assert out["dt"][i] == np.float32(0.005) and out["next_time"][i] == np.float32(12.3)


# ---------------------------------------------------------------- SAP _adapt_dt
def adapt_f32(O, err, sim, nxt, dt, mode=1, tol=TOL, dt_min=DT_MIN, dt_max=1.0):
    n = len(err)
    err = np.asarray(err, np.float32); sim = np.asarray(sim, np.float32).copy()
    nxt = np.asarray(nxt, np.float32); dt = np.asarray(dt, np.float32).copy()
    dth = dt * np.float32(0.5); ideal = dt.copy(); div = np.zeros(n, np.uint8); acc = np.zeros(n, np.uint8)
    ae = np.zeros(n, np.float32); sub = np.zeros(n, np.int32); cum = np.zeros(1, np.int32)
    O.lib().or_adapt_dt_f32(n, _p(err, C.c_float), _p(sim, C.c_float), _p(nxt, C.c_float), _p(dt, C.c_float),
                            _p(dth, C.c_float), _p(ideal, C.c_float), _p(div, C.c_uint8), _p(acc, C.c_uint8),
                            _p(ae, C.c_float), _p(sub, C.c_int32), _p(cum, C.c_int32), mode, C.c_float(tol),
                            C.c_float(dt_min), C.c_float(dt_max), C.c_float(DIVERGENCE))
    return dict(sim=sim, dt=dt, ideal=ideal, div=div.astype(bool), acc=acc.astype(bool), sub=sub, cum=int(cum[0]))


def test_adapt_done_world_noop(O):
    o = adapt_f32(O, [0.0, 0.0], [0.01, 0.0], [0.01, 0.01], [1e-3, 0.0])
    assert not o["acc"].any() and o["cum"] == 0 and o["sim"][0] == np.float32(0.01)


def test_adapt_fixed_mode_guard(O):
    o = adapt_f32(O, [SENTINEL, 0.0], [0.0, 0.0], [0.01, 0.01], [1e-3, 1e-3], mode=0)
    assert not o["acc"][0] and o["div"][0] and o["sim"][0] == np.float32(0.01)
    assert o["acc"][1] and o["sim"][1] == np.float32(1e-3) and o["cum"] == 1


def test_adapt_floor_recovers(O):
    """At the floor a finite error is accepted and ideal_dt follows the Drake rule (no trap)."""
    o = adapt_f32(O, [1e-8], [0.0], [0.01], [DT_MIN])
    assert o["acc"][0] and o["ideal"][0] == pytest.approx(5 * DT_MIN, rel=1e-6)


def test_adapt_reject_shrinks_and_clamps(O):
    o = adapt_f32(O, [1.0], [0.0], [0.01], [1e-3])
    assert not o["acc"][0] and o["ideal"][0] == pytest.approx(1e-4, rel=1e-6) and o["sim"][0] == 0.0


# ---------------------------------------------------------------- cubic Hermite minimiser
def test_cubic_minimizer_known_cubic(O):
    q = lambda a: a ** 3 - 1.05 * a ** 2 + 0.3 * a
    qp = lambda a: 3.0 * a ** 2 - 2.1 * a + 0.3
    for h in (1.0, 1.5):
        got = O.cubic_min(q(0.0), qp(0.0), q(h), qp(h), h)
        assert got is not None and abs(got - 0.5) < 1e-12


def test_cubic_minimizer_degenerate_falls_back(O):
    assert O.cubic_min(0.0, -1.0, -1.0, -1.0, 1.0) is None


def test_cubic_minimizer_exact_on_quadratic(O):
    """A quadratic is its own cubic Hermite interpolant: minimiser -b/2a."""
    for a, b in ((2.0, -1.0), (0.5, -0.3), (3.0, -2.5)):
        f = lambda x: a * x * x + b * x
        fp = lambda x: 2 * a * x + b
        h = 1.0
        got = O.cubic_min(f(0), fp(0), f(h), fp(h), h)
        assert got == pytest.approx(-b / (2 * a), abs=1e-12)
