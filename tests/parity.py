"""Attempt parity over every world (shared by the GPU parity tests).

A device attempt and the FP64 oracle can find different contact sets where a contact sits at the
gap threshold (float32 against float64 distances).  Every world is therefore held to the bars,
in two ways:

* against the oracle's attempt on the contact sets the device solved with (dumped by
  ``na_attempt`` / ``na_art_attempt`` at q and at the midpoint): the solves, the integration and
  the error norm on identical contacts -- every world;
* against the oracle's own end-to-end attempt (its own collision) -- every world whose contact
  counts agree at q and at the midpoint and stay below the per-world capacity (beyond it each side
  keeps the first 128 contacts of its own generation order, so equal counts are not equal sets);
  for the others the reason is printed: the contacts that only one side kept at q, with their
  distance to the gap threshold, a differing midpoint count, or the capacity.  Under the MuJoCo
  contact law a row is active only at pos < 0, so a world with a contact within 1e-6 m of that
  threshold (where float32 and FP64 positions can fall on different sides) is compared on the
  device's contacts only as well (reason "activity").

End to end, each side's collision rounds contact positions differently (float32 against FP64),
and a settled pile can amplify that into the error norm and pose.  A world above the end-to-end
bar is therefore held to min(max(bar, 10 s), 1e-5) instead, where s is how far the oracle's own
end-to-end result moves when its input pose is perturbed at float32 rounding (6e-8 relative);
those worlds and their s are printed.
"""


def _comparable(out, ref, cap, mj_zero_margins=None):
    """Worlds whose end-to-end attempts can be compared, and a reason for each other world.
    ``mj_zero_margins`` (MuJoCo law only; None otherwise): rows use the surface distance (1) or
    dist - margins (0)."""
    n = out["contact_counts"]
    same_q = n[:, 0] == ref["n_contacts"]
    same_mid = n[:, 1] == ref["n_contacts_mid"]
    under = (n[:, 0] < cap) & (n[:, 1] < cap) & (ref["n_contacts"] < cap) & (ref["n_contacts_mid"] < cap)
    calm = np.ones(len(n), dtype=bool)
    if mj_zero_margins is not None:
        c = out["contacts"]
        pos = c[..., 11] if mj_zero_margins else c[..., 12]
        live = np.arange(c.shape[2])[None, None, :] < n[:, :, None]
        calm = ~((np.abs(pos) < 1e-6) & live).any(axis=(1, 2))
    reason = np.where(~under, "capacity", np.where(~same_q, "q", np.where(~same_mid, "midpoint",
                                                                           np.where(~calm, "activity", ""))))
    return same_q & same_mid & under & calm, reason

import numpy as np


def explain_mismatch(O, om, w, q_all, contacts, counts, model_host):
    """One line on why world ``w``'s contact counts differ at q: the contacts only one side kept
    (keyed by shape pair and manifold slot) and their phi0 against the pair's summed gap."""
    n, _, ref = O.collide_world(om, w, q_all, cap=128)
    gap = model_host("shape_gap")
    mine = {(int(r[0]), int(r[1]), int(r[4])): float(r[12]) for r in contacts[0][:int(counts[0])]}
    theirs = {(int(r[0]), int(r[1]), int(r[2])): float(r[10]) for r in ref[:min(n, 128)]}
    parts = []
    for k in sorted(set(mine) ^ set(theirs)):
        side, phi0 = ("gpu", mine[k]) if k in mine else ("oracle", theirs[k])
        parts.append(f"{side}-only {k} phi0-gap={phi0 - float(gap[k[0]] + gap[k[1]]):+.2e}")
    return f"world {w}: gpu {int(counts[0])} vs oracle {n} contacts at q; " + "; ".join(parts[:4])


def check_free_attempts(O, om, cfg, out, ref, q64, qd, f, dt, model_host, err_bar=1e-6, q_bar=1e-5,
                        e2e_err_bar=None, e2e_q_bar=None, trend=None):
    """Free-body attempts: ``out`` from ``__graft_entry__._attempt(..., with_contacts=True)``,
    ``ref`` from ``O.attempt_all`` on the same inputs (``trend``: the full step's seed
    accelerations both were given, None = zero).  Returns the number of worlds whose contact
    counts differed (every world is checked on the device's contacts).  ``e2e_err_bar``,
    ``e2e_q_bar``: the bars against the oracle's own collision (default ``err_bar``, ``q_bar``)."""
    e2e_err_bar = err_bar if e2e_err_bar is None else e2e_err_bar
    e2e_q_bar = q_bar if e2e_q_bar is None else e2e_q_bar
    W = len(out["err"])
    wb = om.world_body_start
    g = O.attempt_all_given(om, cfg, q64, qd, f, dt, out["contacts"], out["contact_counts"], trend=trend)
    assert g["solve_ok"].min() == 1, "an oracle solve on the device's contacts did not converge"
    de_g = np.abs(out["err"] - g["err"])
    dq_g = np.abs(out["q_double"] - g["q_double"]).max(axis=1)
    mjz = int(om.desc.mj_zero_margins) if int(cfg.contact_law) == 1 else None
    same, reason = _comparable(out, ref, int(cfg.max_contacts), mjz)
    for w in np.flatnonzero(~same)[:20]:
        if reason[w] == "q":
            print(explain_mismatch(O, om, int(w), q64, out["contacts"][w], out["contact_counts"][w], model_host),
                  flush=True)
        else:
            print(f"world {int(w)}: {reason[w]} (contacts gpu {out['contact_counts'][w].tolist()}, oracle "
                  f"{[int(ref['n_contacts'][w]), int(ref['n_contacts_mid'][w])]})", flush=True)
    print(f"attempt parity: {W} worlds, {int((~same).sum())} compared on the device's contacts only "
          f"({dict(zip(*np.unique(reason[~same], return_counts=True)))}); on the device's contacts "
          f"max |d err| {de_g.max():.2e}, max |dq| {dq_g.max():.2e}", flush=True)
    bad = np.flatnonzero(de_g > err_bar)
    assert len(bad) == 0, f"worlds {bad[:10].tolist()}: |d err| up to {de_g.max():.2e} on the device's contacts"
    bad = np.flatnonzero(dq_g > q_bar)
    assert len(bad) == 0, f"bodies {bad[:10].tolist()}: |dq| up to {dq_g.max():.2e} on the device's contacts"
    de_r = np.where(same, np.abs(out["err"] - ref["err"]), 0.0)
    dq_r = np.abs(out["q_double"] - ref["q_double"]).max(axis=1)
    dq_w = np.array([dq_r[wb[w]:wb[w + 1]].max(initial=0.0) if same[w] else 0.0 for w in range(W)])
    print(f"end to end: max |d err| {de_r.max():.2e}, max |dq| {dq_w.max():.2e}", flush=True)
    err_bar_w = np.full(W, e2e_err_bar)
    q_bar_w = np.full(W, e2e_q_bar)
    if (de_r > e2e_err_bar).any() or (dq_w > e2e_q_bar).any():
        rng = np.random.default_rng(0)
        qp = q64 * (1.0 + rng.normal(size=q64.shape) * 6e-8)
        refp = O.attempt_all(om, cfg, qp, qd, f, dt, trend=trend)
        s_err = np.abs(refp["err"] - ref["err"])
        sq = np.abs(refp["q_double"] - ref["q_double"]).max(axis=1)
        s_q = np.array([sq[wb[w]:wb[w + 1]].max(initial=0.0) for w in range(W)])
        err_bar_w = np.minimum(np.maximum(e2e_err_bar, 10.0 * s_err), 1e-5)
        q_bar_w = np.minimum(np.maximum(e2e_q_bar, 10.0 * s_q), max(1e-5, e2e_q_bar))
        for w in np.flatnonzero((de_r > e2e_err_bar) | (dq_w > e2e_q_bar))[:20]:
            print(f"world {int(w)} end to end: |d err| {de_r[w]:.2e} (oracle's own s {s_err[w]:.2e}), "
                  f"|dq| {dq_w[w]:.2e} (s {s_q[w]:.2e})", flush=True)
    bad = np.flatnonzero(de_r > err_bar_w)
    assert len(bad) == 0, f"worlds {bad[:10].tolist()}: |d err| up to {de_r.max():.2e} against the oracle's own attempt"
    bad = np.flatnonzero(dq_w > q_bar_w)
    assert len(bad) == 0, f"worlds {bad[:10].tolist()}: |dq| up to {dq_w.max():.2e} against the oracle's own attempt"
    return int((~same).sum())


def check_art_attempts(O, om, oa, cfg, out, ref, q, v, dt, err_abs=1e-7, err_rel=1e-3, q_bar=1e-6):
    """Articulated attempts (``_art_attempt(..., with_contacts=True)`` against
    ``O.art_attempt_all``), the same two-way check; error bar err_abs + err_rel * err."""
    W = len(out["err"])
    g = O.art_attempt_all_given(om, oa, cfg, q, v, dt, out["contacts"], out["contact_counts"])
    assert g["solve_ok"].min() == 1, "an oracle solve on the device's contacts did not converge"
    same, reason = _comparable(out, ref, int(cfg.max_contacts))
    qs = np.asarray(oa.world_q_start[:W + 1])
    coord_same = np.repeat(same, np.diff(qs))
    de_g = np.abs(out["err"] - g["err"])
    ok_g = de_g <= err_abs + err_rel * g["err"]
    dq_g = np.abs(out["q_double"] - g["q_double"])
    for w in np.flatnonzero(~same)[:20]:
        print(f"articulated world {int(w)}: {reason[w]} (contacts gpu {out['contact_counts'][w].tolist()}, oracle "
              f"{[int(ref['n_contacts'][w]), int(ref['n_contacts_mid'][w])]})", flush=True)
    print(f"articulated attempt parity: {W} worlds, {int((~same).sum())} compared on the device's contacts only; "
          f"on the device's contacts max |d err| {de_g.max():.2e}, max |dq| {dq_g.max():.2e}", flush=True)
    assert ok_g.all(), f"worlds {np.flatnonzero(~ok_g)[:10].tolist()}: error norm off on the device's contacts"
    assert dq_g.max() <= q_bar, f"|dq| = {dq_g.max():.2e} on the device's contacts"
    de_r = np.abs(out["err"] - ref["err"])
    bad = np.flatnonzero(same & (de_r > err_abs + err_rel * ref["err"]))
    assert len(bad) == 0, f"worlds {bad[:10].tolist()}: error norm against the oracle's own attempt"
    if coord_same.any():
        assert np.abs(out["q_double"] - ref["q_double"])[coord_same].max() <= q_bar
    return int((~same).sum())
