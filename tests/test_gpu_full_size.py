"""The stiff-stack, uneven-pile and articulated workloads at their full sizes on the GPU.

Size-independent properties, plus the FP64 oracle on the whole batch at the end state:

* 1024 stiff stacks (4 boxes, ke 1e6, mu ~ U[0.3, 1]) at tol 1e-4, run until every world has
  2000 accepted steps (about 2000 frames of 10 ms: a settled stack takes the 10 ms cap).  Every
  world lands on its boundary unless it spent all max_substeps steps of the frame: a landing
  remainder below dt_min sizes the next frame's seed below the floor, and at the floor the
  controller accepts without growing the working dt (solver_sap_adaptive.py:110-133, 292-308),
  so that frame ends at the cap and the next one carries the remainder.  The FP64 oracle driver
  shows 21 such frames in 2000 over the 1024 stacks.  No world diverges, and every accepted
  step's error norm is within tol (the Drake rule accepts only e <= tol).  At the end every
  stack still stands (xy drift < 5 mm, z order kept, residual speed < 0.05 m/s; the FP64 oracle
  driver over 64 of these stacks drifts 0.7 mm with speed 0.016 m/s).  The first 64 worlds end within 2 mm of the oracle's
  own rollout, and a step-doubling attempt over all 1024 worlds matches the oracle to the bars of
  the 64-world attempt tests (error norm 1e-6, doubled pose 1e-5) -- every world, on the contact
  sets the device solved with, and end to end wherever both sides found the same contacts.
* 4096 uneven piles after the 20 warm-up frames of the throughput workload: the persistent
  per-frame form and the reference's per-iteration host loop give bit-identical states, and an
  attempt over all 4096 worlds matches the oracle to the same bars.  The oracle runs with the
  solver's Newton cap of 100: under the reference's 30 it stops short on one of these worlds.
* 8192 articulated worlds (arms and floating quadrupeds on box terrain) after 10 frames: an
  attempt over all worlds matches the oracle to the articulated bars (error norm within
  1e-7 + 1e-3 e, doubled coordinates 1e-6), every world on the device's contact sets and end to
  end wherever both found the same contacts (``tests/parity.py``).
"""

import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GE = importlib.import_module("__graft_entry__")
P = importlib.import_module("parity")
KW = dict(dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)


def _landed_or_capped(sol, dt_outer=0.01):
    """Every world reached its boundary or spent the frame's whole substep budget."""
    off = (sol.sim_time.double() - dt_outer).abs() > 1e-12
    return not bool((off & (sol.substeps < KW["max_substeps"])).any())


def _attempt_matches_oracle(O, model, solver, q64, qd, tol, dts):
    f = torch.zeros_like(qd)
    om = O.OracleModel(model)
    for dt in dts:
        dtt = torch.full((model.world_count,), dt, dtype=torch.float64, device=DEV)
        out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True)
        # the oracle gets the solver's Newton cap (100); the reference's 30 is too few for some
        # settled-pile worlds at dt 4e-3
        cfg = O.make_config(tol=tol, max_iterations=solver._cfg.max_iterations)
        qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
        ref = O.attempt_all(om, cfg, qn, qdn, fn, dtn)
        assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
        print(f"  attempt dt {dt}", flush=True)
        P.check_free_attempts(O, om, cfg, out, ref, qn, qdn, fn, dtn, model.host)


def test_stiff_stacks_1024_at_tol_1e4_for_2000_accepted_steps(na, O):
    tol = 1e-4
    m = na.scenes.box_stack(1024, seed=0, device=DEV)
    state = m.state()
    q0 = state.body_q.double().clone()
    solver = na.SolverSAPAdaptive(m, tol=tol, **KW)
    per_world = torch.zeros(m.world_count, dtype=torch.int64, device=DEV)
    short = torch.zeros((), dtype=torch.int64, device=DEV)
    short_uncapped = torch.zeros((), dtype=torch.int64, device=DEV)
    err_max = torch.zeros((), dtype=torch.float64, device=DEV)
    div = torch.zeros((), dtype=torch.int64, device=DEV)
    frames = 0
    print("stiff stacks: 1024 worlds, tol 1e-4, until 2000 accepted steps per world", flush=True)
    while True:
        solver.step(state, state, m.control(), None, 0.01)
        per_world += solver.substeps
        off = (solver.sim_time.double() - 0.01).abs() > 1e-12
        short += off.sum()
        short_uncapped += (off & (solver.substeps < KW["max_substeps"])).sum()
        err_max = torch.maximum(err_max, solver.last_error.double().max())
        div += solver.diverged.to(torch.int64).sum()
        frames += 1
        if frames % 100 == 0 and int(per_world.min().item()) >= 2000:
            break
        assert frames < 4000
    torch.cuda.synchronize()
    print(f"  {frames} frames, accepted per world {per_world.min().item()}..{per_world.max().item()}, "
          f"{short.item()} world-frames ended at the substep cap", flush=True)
    assert short_uncapped.item() == 0 and div.item() == 0
    assert short.item() <= 200
    assert err_max.item() <= tol * (1 + 1e-9), err_max.item()
    q = state.body_q.double()
    assert bool(torch.isfinite(q).all())
    assert (q[:, :2] - q0[:, :2]).norm(dim=1).max().item() < 5e-3
    z = q[:, 2].view(-1, 4)
    assert bool((z[:, 1:] > z[:, :-1]).all())
    assert (q0[:, 2] - q[:, 2]).abs().max().item() < 1e-2
    assert state.body_qd.abs().max().item() < 0.05

    # the first 64 worlds against the oracle's own boundary-call rollout over as many frames
    sub = na.scenes.box_stack(64, seed=0, device="cpu")
    drv = O.OracleSolver(O.OracleModel(sub), O.make_config(tol=tol, dt_init=1e-3, dt_max=0.01, max_substeps=256))
    oq = sub.host("body_q").astype(np.float64).copy()
    oqd = np.zeros((sub.body_count, 6))
    for _ in range(frames):
        drv.step(oq, oqd, np.zeros_like(oqd), 0.01)
    dq = np.abs(solver.body_q64[: sub.body_count].cpu().numpy() - oq).max()
    assert dq < 2e-3, f"max |q - q_oracle| over 64 stacks after {frames} frames = {dq:.2e}"

    _attempt_matches_oracle(O, m, solver, solver.body_q64.clone(), state.body_qd.clone(), tol, (1e-3, 1e-2))


def test_uneven_piles_4096_forms_identical_and_attempt_matches_oracle(na, O):
    m = na.scenes.uneven_pile(4096, seed=0, device=DEV)
    st = m.state()
    warm = na.SolverSAPAdaptive(m, tol=1e-3, **KW)
    print("uneven piles: 4096 worlds, 20 warm-up frames", flush=True)
    for _ in range(20):
        warm.step(st, st, m.control(), None, 0.01)
    torch.cuda.synchronize()
    q0, qd0 = st.body_q.clone(), st.body_qd.clone()
    runs = {}
    for schedule in ("served", "plain"):
        print(f"  3 frames, schedule {schedule}", flush=True)
        s2 = na.State(q0.clone(), qd0.clone(), torch.zeros_like(qd0))
        sol = na.SolverSAPAdaptive(m, tol=1e-3, schedule=schedule, **KW)
        for _ in range(3):
            sol.step(s2, s2, m.control(), None, 0.01)
            assert _landed_or_capped(sol)
        torch.cuda.synchronize()
        runs[schedule] = (s2.body_q.cpu().numpy(), s2.body_qd.cpu().numpy(), sol.accepted_steps())
    assert np.array_equal(runs["served"][0], runs["plain"][0])
    assert np.array_equal(runs["served"][1], runs["plain"][1])
    assert runs["served"][2] == runs["plain"][2]
    assert np.all(np.isfinite(runs["served"][0]))
    print("  attempt over all worlds against the oracle", flush=True)
    _attempt_matches_oracle(O, m, warm, warm.body_q64.clone(), qd0, 1e-3, (1e-3, 4e-3))


def test_articulated_8192_attempt_matches_oracle(na, O, nalib):
    art = importlib.import_module("test_gpu_articulation")
    m = na.scenes.articulated_mix(8192, seed=0, device=DEV)
    st = m.state()
    sol = na.SolverSAPAdaptive(m, tol=1e-3, **KW)
    print("articulated: 8192 worlds, 10 frames", flush=True)
    for _ in range(10):
        sol.step(st, st, m.control(), None, 0.01)
        assert _landed_or_capped(sol)
    torch.cuda.synchronize()
    q = st.joint_q.cpu().numpy().astype(np.float64)
    v = st.joint_qd.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(q)) and np.all(np.isfinite(v))
    om, oa = O.OracleModel(m), O.OracleArt(m)
    for dt in (1e-3, 4e-3):
        print(f"  attempt dt {dt} against the oracle", flush=True)
        out = art._art_attempt(nalib, sol, q, v, dt, with_contacts=True)
        cfg = O.make_config(tol=1e-3, opt_rel_tol=1e-6)
        dts = np.full(m.world_count, dt)
        ref = O.art_attempt_all(om, oa, cfg, q, v, dts)
        assert out["solve_ok"].min() == 1 and ref["solve_ok"].min() == 1
        P.check_art_attempts(O, om, oa, cfg, out, ref, q, v, dts)
