"""GPU contact generation against the oracle.

* candidate pair sets: bit-exact as sorted sets against the brute-force n^2 oracle run on the
  GPU's own AABBs (sweep-and-prune and all-pairs), including the largest world segment
  (64 primitives + 5 static shapes);
* AABBs: against the oracle's FP64 closed forms (relative 1e-6, the 1e6 plane cube in f32);
* contacts: per world the same (shape_a, shape_b) multiset; within each shape pair a one-to-one
  assignment to the oracle's contacts with world core points within 1e-5 m, gap distances
  within 2e-6 m and normals within 1e-5 of the oracle's unit normal (largest seen: 7e-7 m,
  5e-8 m, 7e-7);
* determinism: two collide calls give bit-identical Contacts arrays.
"""

import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _scenes(na):
    return {
        "drop": na.scenes.box_drop(64, seed=0, device=DEV),
        "stack": na.scenes.box_stack(64, seed=1, device=DEV),
        "pile": na.scenes.uneven_pile(256, seed=2, device=DEV),
        "pile64": na.scenes.uneven_pile(16, seed=3, device=DEV, min_prims=64, max_prims=64),
    }


@pytest.fixture(scope="module")
def scenes(na):
    return _scenes(na)


@pytest.mark.parametrize("bp", ["sap", "nxn", "explicit"])
@pytest.mark.parametrize("name", ["drop", "stack", "pile", "pile64"])
def test_pair_sets_bit_exact(na, O, scenes, name, bp):
    m = scenes[name]
    state = m.state()
    pipe = na.CollisionPipeline(m, broad_phase=bp)
    lo, hi, pairs, cnt = pipe.debug_pairs(state, max_pairs=1024)
    torch.cuda.synchronize()
    om = O.OracleModel(m)
    lo_n, hi_n = lo.cpu().numpy().astype(np.float64), hi.cpu().numpy().astype(np.float64)
    rlo, rhi = O.shape_aabbs(om, m.host("body_q").astype(np.float32).astype(np.float64))
    assert np.allclose(lo_n, rlo, rtol=1e-6, atol=1e-6) and np.allclose(hi_n, rhi, rtol=1e-6, atol=1e-6)
    pairs, cnt = pairs.cpu().numpy(), cnt.cpu().numpy()
    assert cnt.max() <= 1024
    total = 0
    for w in range(m.world_count):
        got = pairs[w, :cnt[w]]
        assert len(set(map(tuple, got.tolist()))) == len(got)  # no duplicates
        ref = O.world_pairs(om, w, lo_n, hi_n)
        assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, ref.tolist())), f"world {w}"
        total += len(got)
    assert total > 0


@pytest.mark.parametrize("name", ["drop", "stack", "pile"])
def test_contacts_match_oracle(na, O, scenes, name):
    m = scenes[name]
    state = m.state()
    pipe = na.CollisionPipeline(m)
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    n = int(c.rigid_contact_count.item())
    wc = c.rigid_contact_world_count.cpu().numpy()
    assert n == int(np.minimum(wc, 128).sum())
    s0 = c.rigid_contact_shape0.cpu().numpy()[:n]
    s1 = c.rigid_contact_shape1.cpu().numpy()[:n]
    nrm = c.rigid_contact_normal.cpu().numpy()[:n]
    p0 = c.rigid_contact_point0.cpu().numpy()[:n]
    p1 = c.rigid_contact_point1.cpu().numpy()[:n]
    o0 = c.rigid_contact_offset0.cpu().numpy()[:n]
    o1 = c.rigid_contact_offset1.cpu().numpy()[:n]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    om = O.OracleModel(m)
    q = m.host("body_q").astype(np.float32).astype(np.float64)
    shape_body = m.host("shape_body")
    start = 0
    worst = {"point": 0.0, "normal": 0.0, "distance": 0.0}
    for w in range(m.world_count):
        k, _, rows = O.collide_world(om, w, q)
        assert k == wc[w], f"world {w}: {wc[w]} vs {k}"
        g = slice(start, start + k)
        start += k
        got = sorted(zip(s0[g], s1[g]))
        ref = sorted(zip(rows[:, 0].astype(int), rows[:, 1].astype(int)))
        assert got == ref
        # One-to-one: every GPU contact of a shape pair is assigned to a distinct oracle contact
        # of that pair (minimum total point distance), and the assigned rows agree on the
        # world-frame core points, the normal and the gap distance.
        def world(b, p):
            return p if b < 0 else _qmat(q[b][3:]) @ p + q[b][:3]

        def wvec(b, v):
            return v if b < 0 else _qmat(q[b][3:]) @ v

        by_pair = {}
        for i in range(g.start, g.stop):
            by_pair.setdefault((int(s0[i]), int(s1[i])), []).append(i)
        for (sa, sb), idx in by_pair.items():
            ba, bb = shape_body[sa], shape_body[sb]
            cand = rows[(rows[:, 0] == sa) & (rows[:, 1] == sb)]
            assert len(cand) == len(idx)
            # the oracle's core point on shape a: midpoint - n (dist / 2 + radius of a)
            pa_ref = cand[:, 6:9] - cand[:, 3:6] * (0.5 * cand[:, 9] + cand[:, 13])[:, None]
            got_pa = np.array([world(ba, p0[i].astype(np.float64)) for i in idx])
            got_pb = np.array([world(bb, p1[i].astype(np.float64)) for i in idx])
            cost = np.linalg.norm(got_pa[:, None, :] - pa_ref[None, :, :], axis=2)
            from scipy.optimize import linear_sum_assignment

            gi, ri = linear_sum_assignment(cost)
            for a_i, r_i in zip(gi, ri):
                i = idx[a_i]
                ref = cand[r_i]
                err_p = cost[a_i, r_i]
                worst["point"] = max(worst["point"], err_p)
                assert err_p < 1e-5, f"world {w} pair {(sa, sb)}: point {err_p:.2e}"
                err_n = np.linalg.norm(nrm[i].astype(np.float64) - ref[3:6])
                worst["normal"] = max(worst["normal"], err_n)
                assert err_n < 1e-5, f"world {w} pair {(sa, sb)}: normal {err_n:.2e}"
                pa = got_pa[a_i] + wvec(ba, o0[i].astype(np.float64))
                pb = got_pb[a_i] + wvec(bb, o1[i].astype(np.float64))
                err_d = abs(np.dot(pb - pa, nrm[i]) - ref[10])
                worst["distance"] = max(worst["distance"], err_d)
                assert err_d < 2e-6, f"world {w} pair {(sa, sb)}: distance {err_d:.2e}"
    print(f"{name}: worst point {worst['point']:.2e} m, normal {worst['normal']:.2e}, "
          f"distance {worst['distance']:.2e} m")


def _qmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_contacts_deterministic_and_capacity_clamped(na, scenes):
    m = scenes["pile"]
    state = m.state()
    pipe = na.CollisionPipeline(m, rigid_contact_max=500)
    a, b = pipe.contacts(), pipe.contacts()
    pipe.collide(state, a)
    pipe.collide(state, b)
    torch.cuda.synchronize()
    for name in ("rigid_contact_shape0", "rigid_contact_shape1", "rigid_contact_point0", "rigid_contact_normal"):
        assert torch.equal(getattr(a, name), getattr(b, name))
    total = int(a.rigid_contact_count.item())
    assert total > 500  # the count reports every contact; the arrays hold the first 500
    assert int((a.rigid_contact_shape0 >= 0).sum()) == 500


def _records(c, n):
    """Contacts as sortable records (shape pair, then the raw float words)."""
    cols = [c.rigid_contact_shape0.cpu().numpy()[:n, None].astype(np.float64),
            c.rigid_contact_shape1.cpu().numpy()[:n, None].astype(np.float64)]
    for a in (c.rigid_contact_point0, c.rigid_contact_point1, c.rigid_contact_normal, c.rigid_contact_offset0):
        cols.append(a.cpu().numpy()[:n].astype(np.float64))
    r = np.concatenate(cols, axis=1)
    return r[np.lexsort(r.T[::-1])]


@pytest.mark.parametrize("name", ["stack", "pile"])
def test_explicit_broad_phase_contacts_equal_all_pairs(na, scenes, name):
    """The explicit pair list (world / group / same-body / filter rules, model.contact_pairs) gives
    the same contacts as the all-pairs broad phase, bit for bit."""
    m = scenes[name]
    state = m.state()
    out = {}
    for bp in ("nxn", "explicit"):
        pipe = na.CollisionPipeline(m, broad_phase=bp)
        c = pipe.contacts()
        pipe.collide(state, c)
        torch.cuda.synchronize()
        n = int(c.rigid_contact_count.item())
        out[bp] = _records(c, n)
    assert len(out["nxn"]) > 0 and np.array_equal(out["nxn"], out["explicit"])


def _generated_order(pipe, c):
    """Rank of every generated (world-major) contact in the deterministic sorted buffer."""
    n = min(int(c.rigid_contact_count.item()), pipe.rigid_contact_max)
    return pipe._prev_rank.cpu().numpy()[:n]  # the rank of the frame just collided


def _match_inputs(pipe, m, c, body_q):
    """Per-world offsets / counts and world midpoints of a compacted Contacts buffer, in the
    generated (world-major) order the restatement takes."""
    cnt = np.minimum(c.rigid_contact_world_count.cpu().numpy(), pipe.per_world)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rank = _generated_order(pipe, c)
    s0, s1 = c.rigid_contact_shape0.cpu().numpy()[rank], c.rigid_contact_shape1.cpu().numpy()[rank]
    sb = m.host("shape_body")
    q = body_q.astype(np.float64)

    def world(s, p):
        b = sb[s]
        out = p.copy()
        for i in range(len(s)):
            if b[i] >= 0:
                x, y, z, w = q[b[i], 3:]
                R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                              [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                              [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
                out[i] = q[b[i], :3] + R @ p[i]
        return out

    p0 = c.rigid_contact_point0.cpu().numpy()[rank].astype(np.float64)
    p1 = c.rigid_contact_point1.cpu().numpy()[rank].astype(np.float64)
    mid = 0.5 * (world(s0, p0) + world(s1, p1))
    return {"off": off, "cnt": cnt, "s0": s0, "s1": s1, "mid": mid, "rank": rank,
            "n": c.rigid_contact_normal.cpu().numpy()[rank].astype(np.float64)}


def _settled_pile(na, worlds=32, seed=8):
    m = na.scenes.uneven_pile(worlds, seed=seed, device=DEV, max_prims=10)
    state = m.state()
    solver = na.SolverSAPAdaptive(m, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)
    for _ in range(40):
        solver.step(state, state, m.control(), None, 0.01)
    torch.cuda.synchronize()
    return m, state


def _moved(na, m, state):
    """One body of every other world moved by 1 mm along x (beyond the 0.5 mm threshold), a second
    body by 0.1 mm along y (within it)."""
    q = state.body_q.clone()
    wb = m.host("world_body_start")
    for w in range(0, m.world_count, 2):
        if wb[w + 1] - wb[w] >= 2:
            q[wb[w], 0] += 1e-3
            q[wb[w] + 1, 1] += 1e-4
    return na.State(q, state.body_qd.clone(), state.body_f.clone())


def test_contact_matching_against_restatement(na, O):
    m, state = _settled_pile(na)
    pipe = na.CollisionPipeline(m, contact_matching="latest", contact_report=True)
    assert pipe.deterministic  # matching implies the sorted order (collide.py:618-620)
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    n0 = int(c.rigid_contact_count.item())
    assert n0 > 0 and (c.rigid_contact_match_index.cpu().numpy()[:n0] == -1).all()
    assert int(c.rigid_contact_new_count.item()) == n0
    prev = _match_inputs(pipe, m, c, state.body_q.cpu().numpy())
    # the same state again: every contact matches itself, nothing new, nothing broken
    pipe.collide(state, c)
    torch.cuda.synchronize()
    assert np.array_equal(c.rigid_contact_match_index.cpu().numpy()[:n0], np.arange(n0))
    assert int(c.rigid_contact_new_count.item()) == 0 and int(c.rigid_contact_broken_count.item()) == 0
    moved = _moved(na, m, state)
    pipe.collide(moved, c)
    torch.cuda.synchronize()
    n1 = int(c.rigid_contact_count.item())
    cur = _match_inputs(pipe, m, c, moved.body_q.cpu().numpy())
    ref_gen = O.match_contacts(prev, cur, pipe.pos_threshold, pipe.normal_dot_threshold)
    # restatement indices are generated-order positions of both frames: map them to sorted rows
    ref = np.empty(n1, dtype=np.int64)
    ref[cur["rank"]] = np.where(ref_gen >= 0, prev["rank"][np.maximum(ref_gen, 0)], ref_gen)
    got = c.rigid_contact_match_index.cpu().numpy()[:n1]
    assert (got == -2).sum() > 0 and (got >= 0).sum() > 0
    assert np.array_equal(got, ref)
    new = c.rigid_contact_new_indices.cpu().numpy()[:int(c.rigid_contact_new_count.item())]
    assert np.array_equal(new, np.nonzero(got < 0)[0])
    broken = c.rigid_contact_broken_indices.cpu().numpy()[:int(c.rigid_contact_broken_count.item())]
    assert np.array_equal(broken, np.setdiff1d(np.arange(n0), got[got >= 0]))


_ROWS = ("rigid_contact_shape0", "rigid_contact_shape1", "rigid_contact_point0", "rigid_contact_point1",
         "rigid_contact_offset0", "rigid_contact_offset1", "rigid_contact_normal", "rigid_contact_margin0",
         "rigid_contact_margin1")


@pytest.mark.parametrize("name,cap", [("pile", None), ("pile", 500), ("pile64", None), ("stack", None)])
def test_deterministic_order_is_the_reference_key_order(na, O, scenes, name, cap):
    """deterministic=True gives the generated contacts, bit for bit, in ascending
    make_contact_sort_key order (sub key = rank within the shape pair), the sorted keys with the
    sentinel past the count, and the same rows on every call."""
    m = scenes[name]
    state = m.state()
    plain = na.CollisionPipeline(m, rigid_contact_max=cap)
    det = na.CollisionPipeline(m, rigid_contact_max=cap, deterministic=True)
    a, b, b2 = plain.contacts(), det.contacts(), det.contacts()
    plain.collide(state, a)
    det.collide(state, b)
    det.collide(state, b2)
    torch.cuda.synchronize()
    n = min(int(a.rigid_contact_count.item()), plain.rigid_contact_max)
    assert n > 0 and int(b.rigid_contact_count.item()) == int(a.rigid_contact_count.item())
    if cap is not None:
        assert int(a.rigid_contact_count.item()) > cap  # the clamped case sorts the stored rows
    keys, perm = O.contact_order(a.rigid_contact_shape0.cpu().numpy()[:n], a.rigid_contact_shape1.cpu().numpy()[:n])
    for name_ in _ROWS:
        ga, gb = getattr(a, name_).cpu().numpy(), getattr(b, name_).cpu().numpy()
        assert np.array_equal(gb[:n], ga[:n][perm]), name_
        assert np.array_equal(gb, getattr(b2, name_).cpu().numpy()), name_
    sk = det._sort_key_array.cpu().numpy()
    assert np.array_equal(sk[:n], keys[perm]) and np.all(np.diff(sk[:n]) > 0)
    assert np.all(sk[n:] == O.SORT_KEY_SENTINEL)
    assert np.array_equal(det._prev_rank.cpu().numpy()[perm], np.arange(n))


def test_sticky_matching_replays_previous_geometry(na, O):
    """contact_matching="sticky": matched rows carry the previous frame's body-frame points,
    offsets and normal (ContactMatcher.replay_matched); the match indices and the unmatched rows
    are those of "latest"."""
    m, state = _settled_pile(na)
    latest = na.CollisionPipeline(m, contact_matching="latest")
    sticky = na.CollisionPipeline(m, contact_matching="sticky")
    cl, cs = latest.contacts(), sticky.contacts()
    latest.collide(state, cl)
    sticky.collide(state, cs)
    torch.cuda.synchronize()
    prev = {k: getattr(cs, k).cpu().numpy().copy() for k in _ROWS}
    moved = _moved(na, m, state)
    latest.collide(moved, cl)
    sticky.collide(moved, cs)
    torch.cuda.synchronize()
    n = int(cs.rigid_contact_count.item())
    ml = cl.rigid_contact_match_index.cpu().numpy()[:n]
    ms = cs.rigid_contact_match_index.cpu().numpy()[:n]
    assert np.array_equal(ml, ms) and (ms >= 0).sum() > 0 and (ms < 0).sum() > 0
    hit = ms >= 0
    for k in _ROWS:
        vs, vl = getattr(cs, k).cpu().numpy()[:n], getattr(cl, k).cpu().numpy()[:n]
        if k in ("rigid_contact_point0", "rigid_contact_point1", "rigid_contact_offset0", "rigid_contact_offset1",
                 "rigid_contact_normal"):
            assert np.array_equal(vs[hit], prev[k][ms[hit]]), k
            assert np.array_equal(vs[~hit], vl[~hit]), k
        else:
            assert np.array_equal(vs, vl), k
    # the moved bodies' replayed points differ from the freshly generated ones
    assert not np.array_equal(cs.rigid_contact_point0.cpu().numpy()[:n][hit], cl.rigid_contact_point0.cpu().numpy()[:n][hit])


def test_pipeline_argument_checks(na, scenes):
    m = scenes["drop"]
    with pytest.raises(ValueError, match="contact_matching must be one of"):
        na.CollisionPipeline(m, contact_matching="always")
    with pytest.raises(ValueError, match="requires contact_matching"):
        na.CollisionPipeline(m, contact_report=True)
    with pytest.raises(ValueError, match="non-negative"):
        na.CollisionPipeline(m, contact_matching="latest", contact_matching_pos_threshold=-1.0)
    det = na.CollisionPipeline(m, deterministic=True)
    with pytest.raises(ValueError, match="sort buffer size"):
        det.collide(m.state(), na.Contacts(det.rigid_contact_max + 1, m.device))


def test_deterministic_pipelines_agree_over_a_rollout(na):
    """Two deterministic pipelines (broad phases nxn and sap) report bit-identical contacts and
    sort keys on every frame of a 100-frame pile rollout (test_collision_pipeline.py:1665-1800)."""
    m = na.scenes.uneven_pile(32, seed=12, device=DEV, max_prims=16)
    state = m.state()
    solver = na.SolverSAPAdaptive(m, tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256)
    pa = na.CollisionPipeline(m, broad_phase="nxn", deterministic=True)
    pb = na.CollisionPipeline(m, broad_phase="sap", deterministic=True)
    ca, cb = pa.contacts(), pb.contacts()
    seen = 0
    for frame in range(100):
        solver.step(state, state, m.control(), None, 0.01)
        pa.collide(state, ca)
        pb.collide(state, cb)
        torch.cuda.synchronize()
        n = int(ca.rigid_contact_count.item())
        assert n == int(cb.rigid_contact_count.item()), f"frame {frame}"
        seen = max(seen, n)
        for k in _ROWS:
            assert torch.equal(getattr(ca, k)[:n], getattr(cb, k)[:n]), f"frame {frame}: {k}"
        assert torch.equal(pa._sort_key_array, pb._sort_key_array), f"frame {frame}: sort keys"
    assert seen > 0


def test_contact_overflow_keeps_the_closest_like_the_oracle(na, O):
    """Worlds with far more contacts than their 128-contact set (27 boxes packed 4 mm apart at the
    builder's default 10 cm gap; under the 1024 candidate pairs a world holds): the device keeps
    the same 128 closest contacts as the oracle (sorted gap distances within 1e-5 m, so that a
    rounding-level swap at the threshold passes)."""
    G = importlib.import_module("test_oracle_geometry")
    m = G._box_grid(na, worlds=4, side=3, device=DEV)
    state = m.state()
    pipe = na.CollisionPipeline(m)
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    wc = c.rigid_contact_world_count.cpu().numpy()
    n = int(c.rigid_contact_count.item())
    assert n == int(np.minimum(wc, 128).sum())
    s0 = c.rigid_contact_shape0.cpu().numpy()[:n]
    s1 = c.rigid_contact_shape1.cpu().numpy()[:n]
    nrm = c.rigid_contact_normal.cpu().numpy()[:n].astype(np.float64)
    p0 = c.rigid_contact_point0.cpu().numpy()[:n].astype(np.float64)
    p1 = c.rigid_contact_point1.cpu().numpy()[:n].astype(np.float64)
    o0 = c.rigid_contact_offset0.cpu().numpy()[:n].astype(np.float64)
    o1 = c.rigid_contact_offset1.cpu().numpy()[:n].astype(np.float64)
    om = O.OracleModel(m)
    q = m.host("body_q").astype(np.float32).astype(np.float64)
    shape_body = m.host("shape_body")

    def world(b, p, off):
        if b < 0:
            return p + off
        R = _qmat(q[b][3:])
        return R @ p + q[b][:3] + R @ off

    start = 0
    for w in range(m.world_count):
        k, _, rows = O.collide_world(om, w, q)
        assert k > 128 and wc[w] == 128
        g = range(start, start + 128)
        start += 128
        d_dev = np.sort([np.dot(world(shape_body[s1[i]], p1[i], o1[i]) - world(shape_body[s0[i]], p0[i], o0[i]),
                                nrm[i]) for i in g])
        d_ref = np.sort(rows[:, 10])
        print(f"world {w}: {k} contacts, kept up to {d_ref.max():.4f} m, "
              f"max |d| difference {np.abs(d_dev - d_ref).max():.2e}")
        assert np.abs(d_dev - d_ref).max() < 1e-5
