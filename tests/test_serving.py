"""Continuous-batching rollout serving.

CPU: the arena allocator, request extraction and validation.
GPU (through ``na_serve``):
* a served rollout of F frames leaves bit-identical states to F boundary calls (``na_step``) of
  the same worlds, also when a small device-time budget interrupts and resumes the worlds;
* a request's result is bit-identical alone and in a mix where it is admitted later, into a
  recycled slot and arena range, next to other rollouts of different sizes and tolerances;
* admission never exceeds the HBM budget, and horizons in frames and in accepted steps end the
  rollouts where they say.
"""

import importlib

import numpy as np
import pytest
import torch

na = importlib.import_module("newton-adaptive_amd")
serving = importlib.import_module("newton-adaptive_amd.serving")
DEV = "cuda:0"


# ---------------------------------------------------------------- CPU
def test_arena_first_fit_and_coalescing():
    a = serving._Arena(100)
    s1, s2, s3 = a.alloc(30), a.alloc(30), a.alloc(30)
    assert (s1, s2, s3) == (0, 30, 60) and a.alloc(20) is None
    a.release(s2, 30)
    assert a.alloc(31) is None and a.alloc(30) == 30
    a.release(s1, 30)
    a.release(30, 30)
    a.release(s3, 30)
    assert a.free == [(0, 100)] and a.free_total == 100
    assert a.alloc(0) == 0


def test_requests_from_model_local_indices():
    m = na.scenes.uneven_pile(5, seed=2, device="cpu", max_prims=6)
    reqs = na.requests_from_model(m, horizon_frames=3, tol=[1e-2, 1e-3, 1e-4, 1e-3, 1e-3])
    wb = m.host("world_body_start")
    assert [r.body_count for r in reqs] == list(wb[1:] - wb[:-1])
    for r in reqs:
        sb = r.shapes["shape_body"]
        assert sb.min() >= 0 and sb.max() < r.body_count
        assert r.bodies["body_inertia"].shape == (r.body_count, 3, 3)
    assert reqs[0].tol == 1e-2 and reqs[2].tol == 1e-4
    big = serving.ServingEngine.request_bytes(reqs[0])
    with pytest.raises(ValueError):
        serving.check_request(serving.RolloutRequest(reqs[0].bodies, reqs[0].shapes), 5, 1 << 30)
    with pytest.raises(ValueError):
        serving.check_request(reqs[0], 5, big - 1)
    serving.check_request(reqs[0], 5, big)


# ---------------------------------------------------------------- GPU
def _boundary_calls(model, frames, **kw):
    state = model.state()
    solver = na.SolverSAPAdaptive(model, **kw)
    for _ in range(frames):
        solver.step(state, state, model.control(), None, 0.01)
    torch.cuda.synchronize()
    return state.body_q.cpu().numpy(), state.body_qd.cpu().numpy(), solver.accepted_steps()


KW = dict(tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)


@pytest.mark.gpu
@pytest.mark.parametrize("budget_ms", [0.0, 0.02])
def test_served_rollout_equals_boundary_calls(budget_ms):
    model = na.scenes.uneven_pile(48, seed=11, device=DEV, max_prims=12)
    frames = 8
    q_ref, qd_ref, acc_ref = _boundary_calls(model, frames, **KW)
    eng = na.ServingEngine(model, max_worlds=48, body_capacity=600, shape_capacity=600, budget_ms=budget_ms,
                           device=DEV, **KW)
    reqs = na.requests_from_model(model, horizon_frames=frames)
    summary = eng.run(reqs, realtime=False)
    print("serve summary", summary, flush=True)
    assert summary["completed"] == 48 and summary["failed"] == 0
    if budget_ms > 0:
        assert summary["rounds"] > 1  # worlds were interrupted and resumed
    wb = model.host("world_body_start")
    for r in reqs:
        b0, b1 = int(wb[r.rid]), int(wb[r.rid + 1])
        assert r.frames == frames
        assert np.array_equal(r.body_q, q_ref[b0:b1]), f"world {r.rid}: pose differs"
        assert np.array_equal(r.body_qd, qd_ref[b0:b1]), f"world {r.rid}: velocity differs"
    assert sum(r.accepted for r in reqs) == acc_ref


@pytest.mark.gpu
def test_request_alone_equals_in_mix():
    model = na.scenes.uneven_pile(40, seed=21, device=DEV, max_prims=16)
    tols = [1e-2, 1e-3, 1e-4, 1e-3] * 10
    target = 22
    mk = lambda: na.requests_from_model(model, horizon_accepted=[40 + 3 * (i % 7) for i in range(40)], tol=tols,
                                        arrivals=[0.0] * 40)
    alone = mk()[target]
    e1 = na.ServingEngine(model, max_worlds=4, body_capacity=100, shape_capacity=100, budget_ms=2.0, device=DEV,
                          **KW)
    s1 = e1.run([alone], realtime=False)
    assert s1["completed"] == 1
    mix = mk()
    e2 = na.ServingEngine(model, max_worlds=8, body_capacity=160, shape_capacity=160, budget_ms=3.0, device=DEV,
                          **KW)
    s2 = e2.run(mix, realtime=False)
    print("mix summary", s2, flush=True)
    assert s2["completed"] == 40 and s2["rounds"] > 4
    r = mix[target]
    assert r.slot >= 0 and r.admitted_at >= 0
    assert r.accepted == alone.accepted and r.accepted >= r.horizon_accepted
    assert np.array_equal(r.body_q, alone.body_q) and np.array_equal(r.body_qd, alone.body_qd)
    # horizons in accepted steps end the rollout as soon as they are reached
    for q in mix:
        assert q.accepted == q.horizon_accepted


@pytest.mark.gpu
def test_admission_respects_hbm_budget():
    model = na.scenes.uneven_pile(24, seed=5, device=DEV, max_prims=10)
    reqs = na.requests_from_model(model, horizon_frames=2)
    per = max(serving.ServingEngine.request_bytes(r) for r in reqs)
    eng = na.ServingEngine(model, max_worlds=24, body_capacity=400, shape_capacity=400, budget_ms=1.0,
                           hbm_budget_bytes=3 * per, device=DEV, **KW)
    eng.submit(reqs)
    peak = 0
    while eng.pending or eng.resident:
        eng.admit(float("inf"))
        peak = max(peak, len(eng.resident))
        assert eng.resident_bytes <= eng.hbm_budget_bytes
        eng.launch()
        torch.cuda.synchronize()
        eng.retire(0.0)
    smallest = min(serving.ServingEngine.request_bytes(r) for r in reqs)
    assert 3 <= peak <= 3 * per // smallest and peak < 24
    assert eng.stats.completed == 24 and all(r.frames == 2 for r in reqs)


@pytest.mark.gpu
def test_static_lockstep_batches_equal_continuous():
    """The fixed-batch baseline (one launch per iteration, 4-byte readback, batch held until its
    longest rollout ends) gives every request the same arithmetic as continuous batching."""
    model = na.scenes.uneven_pile(24, seed=31, device=DEV, max_prims=10)

    def mk():
        return na.requests_from_model(model, horizon_frames=[3 if i % 2 else None for i in range(24)],
                                      horizon_accepted=[None if i % 2 else 25 + 4 * (i % 5) for i in range(24)],
                                      tol=[[1e-2, 1e-3, 1e-4][i % 3] for i in range(24)])

    a, b = mk(), mk()
    sa = na.ServingEngine(model, max_worlds=8, body_capacity=160, shape_capacity=160, budget_ms=2.0, device=DEV,
                          **KW).run(a, realtime=False)
    sb = na.ServingEngine(model, max_worlds=8, body_capacity=160, shape_capacity=160, schedule="static", device=DEV,
                          **KW).run(b, realtime=False)
    print("continuous", sa, "static", sb, flush=True)
    assert sa["completed"] == sb["completed"] == 24
    assert sb["lockstep_iterations"] >= max(r.accepted for r in b)
    for x, y in zip(a, b):
        assert x.accepted == y.accepted and x.frames == y.frames
        assert np.array_equal(x.body_q, y.body_q) and np.array_equal(x.body_qd, y.body_qd)
    for r in b:
        assert (r.frames == 3) if r.horizon_frames else (r.accepted == r.horizon_accepted)
