"""Accepted world-steps/s of the served adaptive step on one MI355X, against the lock-step form.

Workload (default): 4096 uneven worlds, each N ~ U{1..32} primitives from {sphere, box, capsule}
(size lognormal(ln 0.08 m, 0.4)) dropped into a 1 m^2 walled bin, step-doubling error control at
tol 1e-3, dt_outer = 10 ms, after 20 settling frames (part of the workload: the piles land).  A
"step" is one 10 ms boundary frame of every world; the timed region runs exactly `--steps` frames
of every world from the settled state.  `--warmup` is the number of untimed repeats of the timed
launch from the same settled state before the timed one, so the timed window never depends on
it.  The frames run in three forms that leave bit-identical states:

* served (`value`): the rollout form of the serving engine -- one persistent launch (`na_serve`)
  in which waves take worlds from an atomic list and march each one through all its frames at
  its own adaptive dt, with no barrier between worlds or frames;
* plain: the reference's host loop -- per frame, one launch over all worlds per step-doubling
  iteration and a 4-byte flag read back per iteration, worlds in lock step;
* frames: one persistent launch per frame (`na_step`), worlds synchronised at every boundary.

The JSON line also carries the roofline of the dominant kernel (algorithmic bytes of the work
recorded by the per-world device counters over the launch time measured with HIP events on the
launch stream), a serving run with Poisson arrivals (p50/p99 rollout latency), a CPU
baseline -- the FP32 CPU restatement of the same algorithm (FP64 residuals, as on the device) timed
on all host cores and on one, on bounded samples of the same window -- and, from the same CPU leg,
the precision of the device's attempts against the FP64 restatement on a sample of the timed
worlds.
"""

from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
KW = dict(dt_inner_init=1e-3, max_substeps=256)
SETTLE_FRAMES = 20  # the piles land before the timed window (part of the named workload)


def algorithmic_bytes(nb, ns, work, adaptive=True):
    """Algorithmic HBM bytes of the step-doubling attempts recorded in `work` [W, >=6]
    (attempts, Newton iterations, CG iterations, contacts, pairs, solves) for worlds of `nb`
    bodies and `ns` segment shapes: per world-attempt 3 inner steps (body state 104 B/body r+w,
    mass properties 92 B/body, per Newton iteration 92 B/contact + 48 B/body), 2 collision
    passes (176 B/shape, 104 B/pair, 80 B/contact), the error norm (12 B/coord + 8), 64 B of
    controller scalars and the masked commit (3 x (4 nq + 4 nv + 52 nb))."""
    import numpy as np

    w = work.cpu().numpy().astype(np.float64)
    att, newton, contacts, pairs = w[:, 0], w[:, 1], w[:, 3], w[:, 4]
    nb = np.asarray(nb, dtype=np.float64)
    ns = np.asarray(ns, dtype=np.float64)
    nq, nv = 7 * nb, 6 * nb
    collisions = np.maximum(2 * att if adaptive else att, 1)
    cbar = contacts / collisions
    per_attempt = 3 * (104 + 92) * nb + 176 * ns * (2 if adaptive else 1) + 12 * nq + 8 + 64 \
        + 3 * (4 * nq + 4 * nv + 52 * nb)
    b = att * per_attempt + 104 * pairs + 80 * contacts + (92 * cbar + 48 * nb) * newton
    return float(b.sum())


def cpu_baseline(model_fn, q0, qd0, tol, dt_outer, samples, frames, threads, log, warm_start="extrapolate"):
    """Time the FP32 CPU restatement of the path (FP32 geometry, contact rows, Hessian and dense
    factor; FP64 Newton residual, cost, projection and line search, as on the device; OpenMP over
    worlds) from the same settled state as the timed GPU window, on `threads` threads over the
    first samples[0] worlds and on one thread over the first samples[1] worlds."""
    import numpy as np

    from oracle import oracle as O

    out = {}
    with O.use_library(O.library_path("fp32")):
        for n, nthreads in ((samples[0], threads), (samples[1], 1)):
            O.set_num_threads(nthreads)
            m = model_fn(n)
            om = O.OracleModel(m)
            cfg = O.make_config(tol=tol, dt_init=KW["dt_inner_init"], dt_max=dt_outer,
                                max_substeps=KW["max_substeps"], max_iterations=100,
                                opt_rel_tol=max(1e-3 * tol, 1e-8), warm_start=warm_start)
            s = O.OracleSolver(om, cfg)
            B = m.body_count
            q = np.ascontiguousarray(q0[:B].astype(np.float64))
            qd = np.ascontiguousarray(qd0[:B].astype(np.float64))
            f = np.zeros((B, 6))
            t0 = time.perf_counter()
            fail = 0
            for _ in range(frames):
                fail |= s.step(q, qd, f, dt_outer)
            el = time.perf_counter() - t0
            acc = int(s.a["cum_accepted"][0])
            out[nthreads] = (acc / el, acc, el, n, bool(fail == 0))
            log(f"[bench] cpu baseline ({nthreads} threads): {acc} accepted in {el:.1f} s ({acc / el:.0f}/s)")
    O.set_num_threads(threads)
    v, acc, el, n, ok = out[threads]
    v1, acc1, el1, n1, ok1 = out[1]
    sample = (f"CPU restatement (FP32 with FP64 residuals, dense Cholesky Newton, OpenMP over worlds), first {n} "
              f"worlds of the workload, {frames} frames of {dt_outer * 1e3:.0f} ms from the same settled state "
              f"({acc} accepted steps in {el:.1f} s)")
    return {"value": v, "unit": "accepted world-steps/s", "cores": int(threads), "kind": "port", "sample": sample,
            "solves_converged": ok,
            "single_thread": {"value": v1, "cores": 1, "solves_converged": ok1,
                              "sample": f"first {n1} worlds, {frames} frames ({acc1} accepted steps in {el1:.1f} s)"}}


def precision(na, model, q0, qd0, dt, tol, sample, log):
    """Per-result error of the device's step-doubling attempts against the FP64 restatement, on the
    first `sample` worlds from the timed window's start state at each world's working dt: the
    doubled pose, the error norm and the accept decision; on the contact sets the device solved with
    (every world), and end to end with the restatement's own collision where both passes found the
    same contact counts below the per-world capacity."""
    import numpy as np
    import torch

    from oracle import oracle as O

    GE = importlib.import_module("__graft_entry__")
    sub = na.scenes.uneven_pile(sample, seed=0, device="cuda:0") if model is None else model
    B = sub.body_count
    solver = na.SolverSAPAdaptive(sub, tol=tol, dt_inner_init=KW["dt_inner_init"])
    q64 = torch.as_tensor(q0[:B], dtype=torch.float64, device="cuda:0")
    qd = torch.as_tensor(qd0[:B], dtype=torch.float32, device="cuda:0")
    f = torch.zeros_like(qd)
    dtt = torch.as_tensor(dt[:sub.world_count], dtype=torch.float64, device="cuda:0")
    out = GE._attempt(solver, q64, qd, f, dtt, with_contacts=True)
    om = O.OracleModel(sub)
    cfg = O.make_config(tol=tol, max_iterations=100)
    qn, qdn, fn, dtn = q64.cpu().numpy(), qd.cpu().numpy().astype(np.float64), f.cpu().numpy(), dtt.cpu().numpy()
    given = O.attempt_all_given(om, cfg, qn, qdn, fn, dtn, out["contacts"], out["contact_counts"])
    own = O.attempt_all(om, cfg, qn, qdn, fn, dtn)
    n = out["contact_counts"]
    cap = int(cfg.max_contacts)
    comparable = (n[:, 0] == own["n_contacts"]) & (n[:, 1] == own["n_contacts_mid"]) & (n.max(axis=1) < cap) \
        & (np.maximum(own["n_contacts"], own["n_contacts_mid"]) < cap)

    def stats(ref, mask):
        dq = np.abs(out["q_double"] - ref["q_double"]).max(axis=1)
        wb = sub.host("world_body_start")
        dqw = np.array([dq[wb[w]:wb[w + 1]].max(initial=0.0) for w in range(sub.world_count)])[mask]
        de = np.abs(out["err"] - ref["err"])[mask]
        agree = ((out["err"] <= tol) == (ref["err"] <= tol))[mask]
        return {"worlds": int(mask.sum()), "q_double_abs_max": float(dqw.max(initial=0.0)),
                "q_double_abs_median": float(np.median(dqw)) if len(dqw) else 0.0,
                "err_abs_max": float(de.max(initial=0.0)), "err_abs_median": float(np.median(de)) if len(de) else 0.0,
                "accept_decisions_agree": float(agree.mean()) if len(agree) else 1.0}

    res = {"on_device_contacts": stats(given, np.ones(sub.world_count, bool)),
           "end_to_end": stats(own, comparable),
           "sample": f"first {sub.world_count} worlds at the timed window's start, each at its working dt",
           "fp64_on_device": "error norm, accept/reject decision, Newton residual / cost / line search, pose shadow",
           "reference": "FP64 CPU restatement (dense Cholesky Newton)"}
    log(f"[bench] precision vs FP64: |dq| max {res['on_device_contacts']['q_double_abs_max']:.2e}, "
        f"|d err| max {res['on_device_contacts']['err_abs_max']:.2e} over {sub.world_count} worlds")
    return res


def mujoco_form(na, args, log):
    """SolverMuJoCoAdaptive with MuJoCo's soft-constraint inner step (contact law 1) on the same
    uneven piles: its own 20 settling frames, then `args.steps` boundary calls timed (one
    persistent k_frame launch per frame, worlds ordered by their last frame's cost)."""
    import torch

    model = na.scenes.uneven_pile(args.worlds, seed=args.seed, device="cuda:0")
    sol = na.SolverMuJoCoAdaptive(model, **dict(KW, tol=args.tol, dt_inner_max=args.dt_outer))
    st = model.state()
    for _ in range(SETTLE_FRAMES):
        sol.step(st, st, model.control(), None, args.dt_outer)
    torch.cuda.synchronize()
    sol.reset_work_counters()
    before = sol.accepted_steps()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        sol.step(st, st, model.control(), None, args.dt_outer)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    acc = sol.accepted_steps() - before
    wk = sol._work.sum(dim=0).tolist()
    log(f"[bench] mujoco inner step: {acc} accepted in {wall:.3f} s ({acc / wall:.0f}/s)")
    return {"value": acc / wall, "unit": "accepted world-steps/s", "accepted": int(acc), "wall_s": wall,
            "device_ms": e0.elapsed_time(e1), "attempts": int(wk[0]), "newton_iterations": int(wk[1]),
            "contacts": int(wk[3]), "diverged": int(sol.diverged.sum().item()),
            "workload": f"uneven_pile_{args.worlds}, tol {args.tol:g}, {args.steps} frames of "
                        f"{args.dt_outer * 1e3:g} ms after {SETTLE_FRAMES} settling frames",
            "form": "SolverMuJoCoAdaptive boundary calls (k_frame per frame): MuJoCo soft-constraint rows, "
                    "pyramidal cone, MuJoCo controller"}


def articulated_form(na, args, log):
    """Articulated worlds (half 7-dof arms, half floating 12-dof quadrupeds on box terrain): served
    boundary calls (na_step_articulated, one persistent launch per frame) at tol `args.tol`, timed
    over `args.art_frames` frames after as many warm-up frames, with the MFMA work of the contact
    Hessians (2 (3 nc) nv^2 flops per Newton iteration, nc the contacts of the solve); the same
    frames with a fresh controller as boundary calls and as one served rollout launch; and the FP64
    CPU restatement timed on a sample of the same worlds from the same state."""
    import numpy as np
    import torch

    W = args.art_worlds
    model = na.scenes.articulated_mix(W, seed=args.seed, device="cuda:0")
    kw = dict(KW, tol=args.tol, dt_inner_max=args.dt_outer)
    sol = na.SolverSAPAdaptive(model, **kw)
    st = model.state()
    for _ in range(args.art_frames):
        sol.step(st, st, model.control(), None, args.dt_outer)
    torch.cuda.synchronize()
    q0, v0 = st.joint_q.clone(), st.joint_qd.clone()
    sol.reset_work_counters()
    sol.reset_compute_counter()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.art_frames + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for f in range(args.art_frames):
        sol.step(st, st, model.control(), None, args.dt_outer)
        ev[f + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    acc = sol.accepted_steps()
    launch_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.art_frames)]
    work = sol._work.cpu().numpy().astype(np.float64)
    nv = np.diff(model.host("world_qd_start")).astype(np.float64)
    collisions = np.maximum(2.0 * work[:, 0], 1.0)
    cbar = work[:, 3] / collisions
    flops = float((work[:, 1] * 2.0 * 3.0 * cbar * nv * nv).sum())
    tflops = flops / (sum(launch_ms) / 1e3) / 1e12
    log(f"[bench] articulated: {acc} accepted in {wall:.3f} s ({acc / wall:.0f}/s)")
    res = {"value": acc / wall, "unit": "accepted world-steps/s", "worlds": W, "frames": args.art_frames,
           "workload": f"articulated_mix_{W}: {W // 2} 7-dof arms + {W - W // 2} 12-dof floating quadrupeds "
                       f"(18 dofs) on random box terrain, tol {args.tol:g}, {args.art_frames} frames of "
                       f"{args.dt_outer * 1e3:g} ms after {args.art_frames} warm-up frames",
           "form": "one persistent launch per boundary frame (na_step_articulated, k_art_frame)",
           "accepted": acc, "wall_s": wall, "launch_ms": launch_ms, "attempts": int(work[:, 0].sum()),
           "newton_iterations": int(work[:, 1].sum()), "contacts": int(work[:, 3].sum()),
           "finite": bool(torch.isfinite(st.joint_q).all()),
           "mfma": {"kernel": "k_art_frame (contact Hessian J^T G J, v_mfma_f32_32x32x2f32)",
                    "algorithmic_flops": flops, "achieved": tflops, "peak": 157.3, "unit": "TFLOP/s",
                    "frac": tflops / 157.3, "note": "fp32 dense MFMA peak; the solve is latency-bound per wave"}}
    # the same frames from the same state with a fresh controller, as boundary calls and as one
    # served rollout launch (na_serve_mixed over an articulated pool, no synchronisation between
    # frames); the two must agree bit for bit
    import math

    fresh = na.SolverSAPAdaptive(model, **kw)
    sf = model.state()
    sf.joint_q.copy_(q0)
    sf.joint_qd.copy_(v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(args.art_frames):
        fresh.step(sf, sf, model.control(), None, args.dt_outer)
    torch.cuda.synchronize()
    fwall = time.perf_counter() - t0
    facc = fresh.accepted_steps()
    log(f"[bench] articulated boundary calls, fresh controller: {facc} accepted in {fwall:.3f} s ({facc / fwall:.0f}/s)")
    del fresh
    q_frames, v_frames = sf.joint_q.cpu().numpy(), sf.joint_qd.cpu().numpy()
    eng = na.ServingEngine(model, max_worlds=W, body_capacity=model.body_count,
                           shape_capacity=model.shape_count - int(model.global_shapes.numel()), budget_ms=0.0,
                           device="cuda:0", articulated=True, **kw)
    start = model.state()
    start.joint_q.copy_(q0)
    start.joint_qd.copy_(v0)
    reqs = na.requests_from_model(model, horizon_frames=args.art_frames, state=start)
    eng.submit(reqs)
    eng.admit(math.inf)
    if eng.pending:
        raise RuntimeError(f"articulated served rollout: {len(eng.pending)} worlds not admitted")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    eng.launch()
    e1.record(stream)
    torch.cuda.synchronize()
    swall = time.perf_counter() - t0
    eng.retire(swall)
    sacc = int(sum(r.accepted for r in reqs))
    qs, ds = model.host("joint_q_start"), model.host("joint_qd_start")
    wb = model.host("world_body_start")
    same = sacc == facc and all(
        np.array_equal(r.joint_q, q_frames[qs[wb[r.rid]]:qs[wb[r.rid + 1]]])
        and np.array_equal(r.joint_qd, v_frames[ds[wb[r.rid]]:ds[wb[r.rid + 1]]]) for r in reqs)
    log(f"[bench] articulated served: {sacc} accepted in {swall:.3f} s ({sacc / swall:.0f}/s), "
        f"identical to the boundary calls: {same}")
    res["served"] = {"value": sacc / swall, "unit": "accepted world-steps/s", "accepted": sacc, "wall_s": swall,
                     "launch_ms": e0.elapsed_time(e1), "boundary_calls_fresh_controller": facc / fwall,
                     "identical_to_boundary_calls": bool(same),
                     "form": "one persistent rollout launch over the articulated pool (na_serve_mixed), "
                             "largest worlds first"}
    del eng
    if not args.no_cpu:
        from oracle import oracle as O

        n = W  # the whole workload: the restatement runs it in about a second
        m = na.scenes.articulated_mix(n, seed=args.seed, device="cpu")
        om, oa = O.OracleModel(m), O.OracleArt(m)
        cfg = O.make_config(tol=args.tol, dt_init=KW["dt_inner_init"], dt_max=args.dt_outer,
                            max_substeps=KW["max_substeps"], max_iterations=100,
                            opt_rel_tol=max(1e-3 * args.tol, 1e-8))
        ref = O.OracleSolver(om, cfg)
        nq, nd = int(oa.world_q_start[n]), int(oa.world_qd_start[n])
        q = np.ascontiguousarray(q0.cpu().numpy()[:nq].astype(np.float64))
        v = np.ascontiguousarray(v0.cpu().numpy()[:nd].astype(np.float64))
        t0 = time.perf_counter()
        for _ in range(args.art_frames):
            ref.step_articulated(oa, q, v, args.dt_outer)
        el = time.perf_counter() - t0
        cacc = int(ref.a["cum_accepted"][0])
        res["cpu_baseline"] = {"value": cacc / el, "unit": "accepted world-steps/s",
                               "cores": int(O.lib().or_num_threads()), "kind": "port",
                               "sample": f"FP64 CPU restatement, first {n} worlds, {args.art_frames} frames from the "
                                         f"same state ({cacc} accepted steps in {el:.2f} s)"}
    del sol
    return res


def serving_tiles(na, args, dev, kw, log):
    """Autotuner of the serving schedule on the serving workload with every request waiting at
    t = 0: two-wave tiles for the free-body worlds with the most measured work left
    (``ServingEngine(tiles=k)``, concurrent Newton seeds) against one world per wave (extrapolated
    and concurrent seeds).  Tiled and untiled runs of the same seeds must leave bit-identical
    results; the fastest candidate is returned as ``best``."""
    import numpy as np
    import torch

    rows, ref = [], {}
    for ws, tiles in (("extrapolate", 0), ("concurrent", 0), ("concurrent", 64), ("concurrent", 128)):
        smodel, sreqs = na.scenes.serving_requests(args.serve_requests, seed=args.seed + 1, rate=None, device=dev,
                                                   distinct=args.serve_worlds, articulated_share=args.serve_articulated)
        seng = na.ServingEngine(smodel, max_worlds=2048, body_capacity=2048 * 40, shape_capacity=2048 * 40,
                                budget_ms=20.0, device=dev, articulated=args.serve_articulated > 0, tiles=tiles,
                                **dict(kw, warm_start=ws))
        r = seng.run(sreqs, realtime=True, max_wall_s=240.0)
        sig = np.concatenate([q.body_q.reshape(-1) for q in sreqs]
                             + [q.joint_q.reshape(-1) for q in sreqs if q.joint_q is not None])
        same = None
        if ws in ref:
            same = bool(np.array_equal(sig, ref[ws]))
            if not same:
                raise RuntimeError(f"serving tiles changed results under warm_start={ws}")
        ref.setdefault(ws, sig)
        rows.append({"warm_start": ws, "tiles": tiles, "accepted_steps_per_s": r["accepted_steps_per_s"],
                     "latency_p50_s": r["latency_p50_s"], "latency_p99_s": r["latency_p99_s"],
                     "completed": r["completed"], "identical_to_untiled": same})
        log(f"[bench] serving schedule, warm_start {ws}, tiles {tiles}: {r['accepted_steps_per_s']:.0f} accepted "
            f"steps/s, latency p50 {r['latency_p50_s']:.2f} s p99 {r['latency_p99_s']:.2f} s")
        del seng
        torch.cuda.synchronize()
    best = max(rows, key=lambda x: x["accepted_steps_per_s"])
    return {"candidates": rows, "best": best,
            "workload": f"the {args.serve_requests} requests of the serving workload, all waiting at t = 0, "
                        "20 ms rounds",
            "form": "tiles: the worlds with the most measured work left run on two wavefronts each (full step and "
                    "first half step solved at once)"}


def gap_form(na, args, kw, log):
    """The served rollout of the uneven-pile workload built with the reference's default contact
    gap (10 cm, ModelBuilder.ShapeConfig.gap) instead of the workload's 2 cm: far more speculative
    contacts, and worlds whose contacts exceed their 128-contact set keep the 128 closest.  Same
    window as the headline (20 settling frames, then args.steps frames in one launch)."""
    import numpy as np
    import torch

    dev = "cuda:0"
    model = na.scenes.uneven_pile(args.worlds, seed=args.seed, device=dev, gap=0.1)
    W = model.world_count
    st = model.state()
    warm = na.SolverSAPAdaptive(model, **kw)
    over = torch.zeros(W, dtype=torch.int32, device=dev)
    for _ in range(SETTLE_FRAMES):
        warm.step(st, st, model.control(), None, args.dt_outer)
        over |= warm._overflow
    torch.cuda.synchronize()
    q0, qd0 = st.body_q.clone(), st.body_qd.clone()
    del warm
    eng = na.ServingEngine(model, max_worlds=W, body_capacity=model.body_count,
                           shape_capacity=model.shape_count - int(model.global_shapes.numel()), budget_ms=0.0,
                           device=dev, **kw)
    reqs = na.requests_from_model(model, horizon_frames=args.steps, state=na.State(q0, qd0, None))
    eng.submit(reqs)
    eng.admit(math.inf)
    eng.solver.reset_work_counters()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    eng.launch()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    wk = eng.solver._work.sum(dim=0).tolist()
    ovf_window = int(eng.solver._overflow.sum().item())
    eng.retire(wall)
    acc = sum(r.accepted for r in reqs)
    log(f"[bench] gap 0.1 m: {acc} accepted in {wall:.3f} s ({acc / wall:.0f}/s), "
        f"{int((over > 0).sum())} worlds over 128 contacts while settling")
    return {"value": acc / wall, "unit": "accepted world-steps/s", "accepted": acc, "wall_s": wall,
            "launch_ms": e0.elapsed_time(e1), "attempts": int(wk[0]), "newton_iterations": int(wk[1]),
            "contacts_per_collision": wk[3] / max(2 * wk[0], 1),
            "worlds_over_128_contacts_settling": int((over > 0).sum()),
            "worlds_over_128_contacts_last_frame": ovf_window,
            "workload": f"uneven_pile_{W} with the reference's default 10 cm contact gap, tol {args.tol:g}, "
                        f"{args.steps} frames after {SETTLE_FRAMES} settling frames, one served launch"}


def collide_form(na, model, state, log, calls=20):
    """Contact generation through the public API (CollisionPipeline.collide, na_collide) on the
    settled piles: contacts per second and algorithmic HBM bytes per call (176 B per segment
    shape, 104 B per candidate pair, 80 B per contact), timed with HIP events over `calls`
    calls, plain and with the deterministic sort."""
    import numpy as np
    import torch

    stream = torch.cuda.current_stream()
    pipe = na.CollisionPipeline(model)
    _, _, _, pair_count = pipe.debug_pairs(state, max_pairs=1024)
    npairs = int(pair_count.sum().item())
    ns = float(model.shapes_per_world().sum())
    res = {}
    for name, det in (("plain", False), ("deterministic", True)):
        p = na.CollisionPipeline(model, deterministic=det)
        c = p.contacts()
        for _ in range(3):
            p.collide(state, c)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(calls):
            p.collide(state, c)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / calls
        n = min(int(c.rigid_contact_count.item()), p.rigid_contact_max)
        algo = 176.0 * ns + 104.0 * npairs + 80.0 * n
        res[name] = {"ms_per_call": ms, "contacts": n, "contacts_per_s": n / (ms / 1e3),
                     "algorithmic_bytes": algo, "achieved_GBps": algo / (ms / 1e3) / 1e9,
                     "frac_of_8000_GBps": algo / (ms / 1e3) / 8.0e12}
        log(f"[bench] collide ({name}): {ms:.3f} ms per call, {n} contacts, "
            f"{algo / (ms / 1e3) / 1e9:.1f} GB/s algorithmic")
    res["candidate_pairs"] = npairs
    res["form"] = "CollisionPipeline(model).collide(state, contacts) on the settled piles (sap broad phase)"
    del pipe
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2, help="untimed repeats of the timed launch")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--dt-outer", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=3600)
    ap.add_argument("--serve-rate", type=float, default=300.0)
    ap.add_argument("--serve-worlds", type=int, default=1200, help="distinct scenes the requests start from")
    ap.add_argument("--serve-tiles", type=int, default=64,
                    help="continuous serving without the autotuner: worlds per round run as two-wave tiles")
    ap.add_argument("--serve-warm-start", choices=("extrapolate", "concurrent"), default="concurrent",
                    help="Newton seeds of the serving legs without the autotuner (tiles need concurrent)")
    ap.add_argument("--serve-articulated", type=float, default=0.25,
                    help="share of requests that are articulated worlds (arm or quadruped on terrain)")
    ap.add_argument("--cpu-worlds", type=int, default=256, help="CPU baseline sample on all host cores")
    ap.add_argument("--cpu-worlds-1t", type=int, default=24, help="CPU baseline sample on one core")
    ap.add_argument("--cpu-frames", type=int, default=10)
    ap.add_argument("--precision-worlds", type=int, default=512)
    ap.add_argument("--no-articulated", action="store_true")
    ap.add_argument("--no-collide", action="store_true")
    ap.add_argument("--no-gap", action="store_true", help="skip the 10 cm contact-gap variant of the workload")
    ap.add_argument("--no-mujoco", action="store_true", help="skip the MuJoCo inner-step line")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default schedule (4 waves per CU)")
    ap.add_argument("--warm-start", choices=("extrapolate", "concurrent", "reference"), default="extrapolate",
                    help="Newton seeds of the attempts' solves (extrapolated; concurrent: the first half step's seed "
                         "independent of the full step, which two-wave tiles need; or the reference's)")
    ap.add_argument("--articulated-only", action="store_true",
                    help="run only the articulated workload (profiling of k_art_frame); prints its JSON object")
    ap.add_argument("--art-worlds", type=int, default=8192)
    ap.add_argument("--art-frames", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true",
                    help="finish with a 1 GiB read + 1 GiB write elementwise kernel (scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; worlds shard across GPUs by contiguous ranges "
                                   "with no per-step exchange, but only --gpus 1 is measured"}))
        sys.exit(2)

    import numpy as np
    import torch

    na = importlib.import_module("newton-adaptive_amd")
    dev = "cuda:0"
    kw = dict(KW, tol=args.tol, dt_inner_max=args.dt_outer, warm_start=args.warm_start)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.articulated_only:
        print(json.dumps({"articulated": articulated_form(na, args, log)}))
        return

    model = na.scenes.uneven_pile(args.worlds, seed=args.seed, device=dev)
    W = model.world_count
    log(f"[bench] {W} worlds, {model.body_count} bodies; {SETTLE_FRAMES} settling frames, "
        f"{args.warmup} untimed repeats of the timed launch")
    st = model.state()
    warm = na.SolverSAPAdaptive(model, **kw)
    for _ in range(SETTLE_FRAMES):
        warm.step(st, st, model.control(), None, args.dt_outer)
    torch.cuda.synchronize()
    q0, qd0 = st.body_q.clone(), st.body_qd.clone()
    dt0 = warm._ideal_dt.clamp(max=args.dt_outer).cpu().numpy().copy()  # each world's working step
    tune = None
    if not args.no_autotune:
        # schedule autotuner of the boundary call on the settled state (restores it)
        tune = warm.autotune(st, args.dt_outer, frames=2)
        log(f"[bench] autotune (boundary calls): best {tune['best']}")
    del warm

    # ---------------------------------------------------------------- served rollout (timed)
    stream = torch.cuda.current_stream()

    def served_run(grid=0, tiles=0, round_ms=0.0):
        """The served rollout of args.steps frames of every world from the settled state
        (``grid``: persistent grid in waves, 0 = one wave per resident slot).  Without tiles it is
        one launch; with ``tiles`` the rollout runs in launches of ``round_ms`` of device time, each
        running the ``tiles`` worlds with the most measured work left as two-wave tiles.  Returns
        the engine, the requests, the wall time and the summed device time of the launches."""
        eng = na.ServingEngine(model, max_worlds=W, body_capacity=model.body_count,
                               shape_capacity=model.shape_count - int(model.global_shapes.numel()),
                               budget_ms=round_ms, device=dev, tiles=tiles, **kw)
        if grid:
            eng.solver.set_schedule(grid)
        reqs = na.requests_from_model(model, horizon_frames=args.steps, state=na.State(q0, qd0, None))
        eng.submit(reqs)
        eng.admit(math.inf)
        eng.solver.reset_work_counters()
        torch.cuda.synchronize()
        if tiles or round_ms > 0:
            t0 = time.perf_counter()
            eng.run(realtime=False, pipelined=False)
            wall = time.perf_counter() - t0
            ms = eng.stats.device_s * 1e3
        else:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(stream)
            eng.launch()
            e1.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            eng.retire(wall)
            ms = e0.elapsed_time(e1)
        assert not eng.resident, "every rollout finishes"
        return eng, reqs, wall, ms

    served_sched, served_tune = (0, 0, 0.0), None
    k0 = 0
    if not args.no_autotune and args.warmup >= 3:
        # the first untimed repeats time the served rollout on each candidate schedule (same results)
        def tune_run(grid, tiles, round_ms):
            e, r, w_wall, ms = served_run(grid, tiles, round_ms)
            sig = torch.as_tensor(np.concatenate([x.body_q.reshape(-1) for x in r]))
            del e
            return w_wall * 1e3, sig
        cands = [{"waves_per_cu": w, "tiles": 0, "round_ms": 0.0} for w in (4, 3, 2)]
        if args.warm_start == "concurrent":
            cands += [{"waves_per_cu": 4, "tiles": t, "round_ms": r} for t, r in ((64, 250.0), (128, 250.0))]
        served_tune = na.serving.autotune_served(tune_run, cands)
        b = served_tune["best"]
        served_sched = (b["grid_waves"], b["tiles"], b["round_ms"])
        k0 = len(served_tune["candidates"])
        log(f"[bench] autotune (served rollout): {served_tune['candidates']} -> {served_tune['best']}")
    for k in range(k0, args.warmup):
        _, _, w_wall, _ = served_run(*served_sched)
        log(f"[bench] untimed repeat {k + 1}/{args.warmup} of the timed launch: {w_wall:.3f} s")
    eng, reqs, wall, launch_ms = served_run(*served_sched)
    accepted = sum(r.accepted for r in reqs)
    value = accepted / wall
    slots = np.array([r.slot for r in reqs])
    nb = np.zeros(W)
    ns = np.zeros(W)
    nb[slots] = [r.body_count for r in reqs]
    ns[slots] = [r.shape_count + eng.G for r in reqs]
    work = eng.solver._work.clone()
    wk = work.sum(dim=0).tolist()
    bytes_launch = algorithmic_bytes(nb, ns, work)
    achieved = bytes_launch / (launch_ms / 1e3) / 1e9
    wb = model.host("world_body_start")
    q_srv = np.zeros((model.body_count, 7), np.float32)
    qd_srv = np.zeros((model.body_count, 6), np.float32)
    for r in reqs:
        q_srv[wb[r.rid]:wb[r.rid + 1]] = r.body_q
        qd_srv[wb[r.rid]:wb[r.rid + 1]] = r.body_qd
    finite = bool(np.isfinite(q_srv).all())
    log(f"[bench] served: {accepted} accepted steps in {wall:.3f} s ({value:.0f}/s), launch {launch_ms:.1f} ms")
    del eng

    out = {
        "metric": "accepted world-steps/s at tol 1e-3 (served adaptive vs plain); p99 rollout latency",
        "value": value,
        "unit": "accepted world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 residual/controller/pose)",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": f"uneven_pile_{W}: {W} worlds x U{{1..32}} sphere/box/capsule in a 1 m^2 bin, "
                               f"tol {args.tol:g}, {args.steps} frames of {args.dt_outer * 1e3:g} ms after "
                               f"{SETTLE_FRAMES} settling frames",
                   "settle_frames": SETTLE_FRAMES, "warm_start": args.warm_start,
                   "worlds": W, "tol": args.tol, "dt_outer": args.dt_outer, "seed": args.seed},
        "served": {"form": ("persistent rollout launch (na_serve), no barrier between worlds or frames"
                            if not served_sched[1] else
                            f"persistent rollout launches (na_serve) of {served_sched[2]:g} ms of device time; in "
                            f"each, the {served_sched[1]} worlds with the most measured work left run as two-wave "
                            "tiles (full step and first half step solved at once), the rest one world per wave"),
                   "schedule": {"grid_waves": served_sched[0], "tiles": served_sched[1],
                                "round_ms": served_sched[2]},
                   "accepted": accepted, "wall_s": wall, "finite": finite, "attempts": int(wk[0]),
                   "newton_iterations": int(wk[1]), "cg_iterations": int(wk[2]), "contacts": int(wk[3]),
                   "pairs": int(wk[4])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_serve_tile" if served_sched[1] else "k_serve",
                     "algorithmic_bytes_per_launch": bytes_launch, "launch_ms": launch_ms},
    }
    if tune is not None:
        out["autotune"] = {"boundary_calls": dict(tune, form="2 boundary calls of the settled piles per candidate "
                                                               "(waves per CU x world order), from the same state")}
        if served_tune is not None:
            out["autotune"]["served"] = dict(served_tune, form="the timed served rollout on each candidate schedule "
                                                              "(the first untimed repeats; ms of wall time)")

    # ---------------------------------------------------------------- lock-step forms
    def frames_form(schedule):
        s2 = na.State(q0.clone(), qd0.clone(), torch.zeros_like(qd0))
        sol = na.SolverSAPAdaptive(model, schedule=schedule, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            sol.step(s2, s2, model.control(), None, args.dt_outer)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        acc = sol.accepted_steps()
        same = bool(np.array_equal(s2.body_q.cpu().numpy(), q_srv) and np.array_equal(s2.body_qd.cpu().numpy(), qd_srv))
        log(f"[bench] {schedule}: {acc} accepted in {el:.3f} s ({acc / el:.0f}/s), identical to served: {same}")
        return {"value": acc / el, "accepted": acc, "wall_s": el, "identical_to_served": same}

    if not args.no_plain:
        out["plain"] = frames_form("plain")
        out["plain"]["form"] = ("reference host loop: one launch per iteration over all worlds + 4-byte flag "
                                "readback; worlds at their boundary skip the iteration")
        out["served_over_plain"] = value / out["plain"]["value"]
        out["plain_batch"] = frames_form("plain_batch")
        out["plain_batch"]["form"] = ("the same host loop with the reference's fixed-batch work pattern: worlds at "
                                      "their boundary still run each iteration's two collision passes (batched "
                                      "CollisionPipeline.collide); their zero-length solves are not charged")
        out["served_over_plain_batch"] = value / out["plain_batch"]["value"]
    if not args.no_frames:
        out["frames"] = frames_form("served")
        out["frames"]["form"] = "one persistent launch per boundary frame (na_step)"

    # ---------------------------------------------------------------- serving (Poisson arrivals)
    if not args.no_serving:
        # the serving schedule (Newton seeds x two-wave tiles) is picked by the tile autotuner on the
        # saturated stream; without it the concurrent seeds with --serve-tiles
        if not args.no_autotune:
            tiles_tune = serving_tiles(na, args, dev, kw, log)
            serve_ws, serve_tiles = tiles_tune["best"]["warm_start"], tiles_tune["best"]["tiles"]
        else:
            tiles_tune, serve_ws, serve_tiles = None, args.serve_warm_start, args.serve_tiles
        skw = dict(kw, warm_start=serve_ws)
        res = {}
        for sched in ("continuous", "static"):
            smodel, sreqs = na.scenes.serving_requests(args.serve_requests, seed=args.seed + 1, rate=args.serve_rate,
                                                       device=dev, distinct=args.serve_worlds,
                                                       articulated_share=args.serve_articulated)
            seng = na.ServingEngine(smodel, max_worlds=2048, body_capacity=2048 * 40, shape_capacity=2048 * 40,
                                    budget_ms=20.0, schedule=sched, device=dev,
                                    articulated=args.serve_articulated > 0,
                                    tiles=serve_tiles if sched == "continuous" else 0, **skw)
            log(f"[bench] serving {len(sreqs)} requests at {args.serve_rate:g}/s, {sched} batching")
            res[sched] = seng.run(sreqs, realtime=True, max_wall_s=240.0)
            res[sched]["final_states"] = [(r.body_q, r.body_qd, r.joint_q, r.joint_qd) for r in sreqs]
            res[sched]["articulated_requests"] = int(sum(r.articulated for r in sreqs))
            log(f"[bench]   {res[sched]['accepted_steps_per_s']:.0f} accepted steps/s, latency p50 "
                f"{res[sched]['latency_p50_s']:.2f} s p99 {res[sched]['latency_p99_s']:.2f} s")
            del seng
        same = all(a[0] is not None and b[0] is not None and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                   and (a[2] is None or (np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])))
                   for a, b in zip(res["continuous"].pop("final_states"), res["static"].pop("final_states")))
        out["serving"] = dict(res["continuous"], requests=args.serve_requests, arrival_rate_per_s=args.serve_rate,
                              workload=f"1..64 primitives per world (uneven pile distribution, {args.serve_worlds} "
                                       f"distinct scenes), {args.serve_articulated:.0%} of the requests articulated "
                                       "(7-dof arm or 18-dof quadruped on box terrain), horizon U[200, 2000] "
                                       "accepted steps, tol in {1e-2, 1e-3, 1e-4}, Poisson arrivals; continuous "
                                       "batching in 20 ms device-time rounds",
                              schedule={"warm_start": serve_ws, "tiles": serve_tiles,
                                        "form": "each round the free-body worlds with the most measured work left "
                                                "run as two-wave tiles, the rest one world per wave"},
                              static=dict(res["static"], form="fixed batches in lock step: one launch per iteration "
                                                              "+ 4-byte readback, batch held to its longest rollout"),
                              served_over_static=res["continuous"]["accepted_steps_per_s"]
                              / res["static"]["accepted_steps_per_s"],
                              identical_results=same)
        out["p99_rollout_latency_s"] = res["continuous"]["latency_p99_s"]
        # the same stream with every request waiting at t = 0: the engine's capacity
        sat = {}
        for sched in ("continuous", "static"):
            smodel, sreqs = na.scenes.serving_requests(args.serve_requests, seed=args.seed + 1, rate=None,
                                                       device=dev, distinct=args.serve_worlds,
                                                       articulated_share=args.serve_articulated)
            seng = na.ServingEngine(smodel, max_worlds=2048, body_capacity=2048 * 40, shape_capacity=2048 * 40,
                                    budget_ms=20.0, schedule=sched, device=dev,
                                    articulated=args.serve_articulated > 0,
                                    tiles=serve_tiles if sched == "continuous" else 0, **skw)
            r = seng.run(sreqs, realtime=True, max_wall_s=240.0)
            sat[sched] = {k: r[k] for k in ("accepted_steps_per_s", "latency_p50_s", "latency_p99_s", "completed",
                                            "device_busy")}
            log(f"[bench] saturated serving, {sched}: {r['accepted_steps_per_s']:.0f} accepted steps/s, latency p50 "
                f"{r['latency_p50_s']:.2f} s p99 {r['latency_p99_s']:.2f} s")
            del seng
        out["serving"]["saturated"] = dict(sat, form="the same requests all waiting at t = 0 (capacity)")
        if tiles_tune is not None:
            out["serving"]["tiles"] = tiles_tune

    if not args.no_gap:
        out["gap_0.1"] = gap_form(na, args, kw, log)

    # ---------------------------------------------------------------- contact generation API
    if not args.no_collide:
        out["collide"] = collide_form(na, model, na.State(q0, qd0, None), log)

    # ---------------------------------------------------------------- articulated worlds
    if not args.no_articulated:
        out["articulated"] = articulated_form(na, args, log)
    if not args.no_mujoco:
        out["mujoco"] = mujoco_form(na, args, log)

    # ---------------------------------------------------------------- CPU baseline
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))

        def cpu_model(n):
            return na.scenes.uneven_pile(n, seed=args.seed, device="cpu")

        out["cpu_baseline"] = cpu_baseline(cpu_model, q0.cpu().numpy(), qd0.cpu().numpy(), args.tol, args.dt_outer,
                                           (args.cpu_worlds, args.cpu_worlds_1t), args.cpu_frames, threads, log,
                                           warm_start=args.warm_start)
        # the same CPU leg checks the device's attempts against the FP64 restatement
        out["precision"] = precision(na, na.scenes.uneven_pile(args.precision_worlds, seed=args.seed, device=dev),
                                     q0.cpu().numpy(), qd0.cpu().numpy(), dt0, args.tol, args.precision_worlds, log)
    traffic_file = os.path.join(REPO, "profiles", "k_serve_uneven_pile_4096_traffic.json")
    if os.path.exists(traffic_file) and W == 4096 and args.steps == 20 and args.seed == 0:
        with open(traffic_file) as fh:
            tr = json.load(fh)
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        out["roofline"]["traffic_source"] = os.path.relpath(traffic_file, REPO)
    print(json.dumps(out))
    if args.copy_probe:
        x = torch.ones(256 * 1024 * 1024, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
