"""Oracle geometry pinned against the reference's closed-form tables and invariants.

* Primitive distances: the known-answer tables of ``newton/tests/test_collision_primitives.py``
  (extracted into ``tests/golden/primitive_kats.json`` by ``tests/golden/extract_primitive_kats.py``).
* Box-box / capsule-box (this project's SAT + clipping / segment-SDF restatement): separating-axis
  depth, unit normals pointing A -> B, contact points between the two surfaces
  (``test_narrow_phase.py:635-727`` invariants).
* Broad phase: an independent NumPy brute force over random AABBs with world -1 and group rules
  (``test_broad_phase.py:90-148,905-1050`` semantics, ``broad_phase_common.py:132-180``).
* AABBs: closed forms per primitive (``collide.py:150-275``).
"""

import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "primitive_kats.json")))

PLANE, SPHERE, CAPSULE, BOX = 1, 3, 4, 7
IDQ = (0.0, 0.0, 0.0, 1.0)


def quat_z_to(axis):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.dot(z, a))
    if c > 1 - 1e-12:
        return IDQ
    if c < -1 + 1e-12:
        return (1.0, 0.0, 0.0, 0.0)
    ax = np.cross(z, a)
    ax /= np.linalg.norm(ax)
    s = math.sin(0.5 * math.acos(c))
    return (ax[0] * s, ax[1] * s, ax[2] * s, math.cos(0.5 * math.acos(c)))


def quat_from_mat(m):
    R = np.asarray(m, dtype=np.float64).reshape(3, 3)
    w = math.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    if w > 1e-6:
        return ((R[2, 1] - R[1, 2]) / (4 * w), (R[0, 2] - R[2, 0]) / (4 * w), (R[1, 0] - R[0, 1]) / (4 * w), w)
    raise ValueError("only near-identity rotations appear in the tables")


def xf(p, q=IDQ):
    return np.array([*p, *q], dtype=np.float64)


def cases(name):
    return KATS[name]["cases"]


def test_plane_sphere_table(O):
    for n, pp, sp, r, expected in cases("plane_sphere"):
        out = O.collide_shapes(PLANE, xf(pp, quat_z_to(n)), (0, 0, 0), SPHERE, xf(sp), (r, 0, 0))
        assert out[0, 0] == pytest.approx(expected, abs=1e-9)
        if expected < 0:  # contact point between sphere centre and plane
            c = out[0, 1:4]
            assert np.linalg.norm(c - np.asarray(sp)) < r + 0.01
            assert np.dot(c - np.asarray(pp), n) <= np.dot(np.asarray(sp) - np.asarray(pp), n) + 0.01


def test_sphere_sphere_table(O):
    for p1, r1, p2, r2, expected in cases("sphere_sphere"):
        out = O.collide_shapes(SPHERE, xf(p1), (r1, 0, 0), SPHERE, xf(p2), (r2, 0, 0))
        assert out[0, 0] == pytest.approx(expected, abs=1e-9)
        nrm = out[0, 4:7]
        d = np.asarray(p2) - np.asarray(p1)
        if np.linalg.norm(d) > 0:
            assert np.allclose(nrm, d / np.linalg.norm(d), atol=1e-12)


def test_sphere_capsule_table(O):
    for sp, sr, cp, ax, cr, hl, expected in cases("sphere_capsule"):
        out = O.collide_shapes(SPHERE, xf(sp), (sr, 0, 0), CAPSULE, xf(cp, quat_z_to(ax)), (cr, hl, 0))
        # closest_segment_point adds 1e-6 to |ab|^2 (collision_primitive.py:52)
        assert out[0, 0] == pytest.approx(expected, abs=1e-6)


def test_capsule_capsule_table(O):
    for p1, a1, r1, h1, p2, a2, r2, h2, expected in cases("capsule_capsule"):
        out = O.collide_shapes(CAPSULE, xf(p1, quat_z_to(a1)), (r1, h1, 0), CAPSULE, xf(p2, quat_z_to(a2)),
                               (r2, h2, 0))
        assert out[:, 0].min() == pytest.approx(expected, abs=1e-9)


def test_sphere_box_table(O):
    for sp, sr, bp, rot, size, expected in cases("sphere_box"):
        out = O.collide_shapes(SPHERE, xf(sp), (sr, 0, 0), BOX, xf(bp, quat_from_mat(rot)), size)
        assert out[0, 0] == pytest.approx(expected, abs=1e-9)


def test_plane_capsule_table(O):
    for n, pp, cp, ax, r, hl, expected in cases("plane_capsule"):
        out = O.collide_shapes(PLANE, xf(pp, quat_z_to(n)), (0, 0, 0), CAPSULE, xf(cp, quat_z_to(ax)), (r, hl, 0))
        assert len(out) == 2
        assert out[:, 0].min() == pytest.approx(expected, abs=1e-9)


def test_plane_box_table(O):
    for n, pp, bp, rot, size, count, expected in cases("plane_box"):
        out = O.collide_shapes(PLANE, xf(pp, quat_z_to(n)), (0, 0, 0), BOX, xf(bp, quat_from_mat(rot)), size,
                               gap_sum=0.0)
        assert len(out) == count
        if count:
            assert out[:, 0].min() == pytest.approx(expected, abs=1e-9)


def test_capsule_box_table(O):
    for cp, ax, r, hl, bp, rot, size, expected in cases("capsule_box"):
        out = O.collide_shapes(CAPSULE, xf(cp, quat_z_to(ax)), (r, hl, 0), BOX, xf(bp, quat_from_mat(rot)), size,
                               gap_sum=1.0)
        assert out[:, 0].min() == pytest.approx(expected, abs=1e-7)
        assert np.allclose(out[:, 4:7], [0, 0, -1], atol=1e-9)  # capsule above the box: normal A->B is -z


def test_box_box_table(O):
    sep, ovl = cases("box_box")
    out = O.collide_shapes(BOX, xf(sep[0]), sep[2], BOX, xf(sep[3]), sep[5], gap_sum=2.0)
    assert out[:, 0].min() == pytest.approx(1.0, abs=1e-12)
    out = O.collide_shapes(BOX, xf(ovl[0]), ovl[2], BOX, xf(ovl[3]), ovl[5], gap_sum=0.0)
    assert len(out) == 4 and np.allclose(out[:, 0], -0.5, atol=1e-12)
    assert np.allclose(out[:, 4:7], [1, 0, 0], atol=1e-12)
    # the four face-contact points lie on the overlap square x=0.75 (midway between faces)
    assert np.allclose(out[:, 1], 0.75, atol=1e-12)
    assert np.allclose(np.sort(np.abs(out[:, 2])), 1.0) and np.allclose(np.abs(out[:, 3]), 1.0)


def _box_sat_depth(pa, Ra, ha, pb, Rb, hb):
    """Brute-force separating-axis overlap (the minimum over the 15 axes)."""
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(Ra[:, i], Rb[:, j])
            if np.linalg.norm(c) > 1e-6:
                axes.append(c / np.linalg.norm(c))
    best = -np.inf
    for L in axes:
        ra = sum(ha[k] * abs(np.dot(Ra[:, k], L)) for k in range(3))
        rb = sum(hb[k] * abs(np.dot(Rb[:, k], L)) for k in range(3))
        best = max(best, abs(np.dot(pb - pa, L)) - ra - rb)
    return best


def _rand_quat(rng):
    u1, u2, u3 = rng.random(3)
    s1, s2 = math.sqrt(1 - u1), math.sqrt(u1)
    return (s1 * math.sin(2 * math.pi * u2), s1 * math.cos(2 * math.pi * u2), s2 * math.sin(2 * math.pi * u3),
            s2 * math.cos(2 * math.pi * u3))


def _qmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _box_sdf(p, h):
    q = np.abs(p) - h
    return np.linalg.norm(np.maximum(q, 0)) + min(q.max(), 0.0)


def test_box_box_invariants_random(O):
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(300):
        ha = rng.uniform(0.05, 0.2, 3)
        hb = rng.uniform(0.05, 0.2, 3)
        qa, qb = _rand_quat(rng), _rand_quat(rng)
        pa = np.zeros(3)
        pb = rng.normal(size=3)
        pb *= rng.uniform(0.05, 0.5) / np.linalg.norm(pb)
        out = O.collide_shapes(BOX, xf(pa, qa), ha, BOX, xf(pb, qb), hb, gap_sum=0.0)
        Ra, Rb = _qmat(qa), _qmat(qb)
        depth = _box_sat_depth(pa, Ra, ha, pb, Rb, hb)
        if depth > 0:
            assert len(out) == 0
            continue
        assert len(out) >= 1
        checked += 1
        n = out[0, 4:7]
        assert abs(np.linalg.norm(n) - 1) < 1e-9
        assert np.dot(n, pb - pa) > 0  # A -> B
        # the deepest reported contact equals the SAT overlap up to the edge-axis bias (1e-4 * size)
        assert out[:, 0].min() == pytest.approx(depth, abs=1e-4 * max(ha.max(), hb.max()) + 1e-9)
        for row in out:
            c, d = row[1:4], row[0]
            # the two surface points c -+ n d/2 sit on (or inside) their boxes' surfaces
            sa = _box_sdf(Ra.T @ (c - 0.5 * d * n - pa), ha)
            sb = _box_sdf(Rb.T @ (c + 0.5 * d * n - pb), hb)
            assert sa <= 1e-9 and sb <= 1e-9
    assert checked > 50


def test_capsule_box_invariants_random(O):
    rng = np.random.default_rng(11)
    for _ in range(200):
        h = rng.uniform(0.05, 0.2, 3)
        qb = _rand_quat(rng)
        r, hl = rng.uniform(0.02, 0.08), rng.uniform(0.02, 0.1)
        qc = _rand_quat(rng)
        cp = rng.normal(size=3)
        cp *= rng.uniform(0.0, 0.4) / np.linalg.norm(cp)
        out = O.collide_shapes(CAPSULE, xf(cp, qc), (r, hl, 0), BOX, xf((0, 0, 0), qb), h, gap_sum=1.0)
        Rb = _qmat(qb)
        ax = _qmat(qc)[:, 2]
        # brute-force min over the segment of sdf - r
        ts = np.linspace(-1, 1, 2001)
        brute = min(_box_sdf(Rb.T @ (cp + t * hl * ax), h) for t in ts) - r
        assert out[:, 0].min() == pytest.approx(brute, abs=2e-4)
        for row in out:
            assert abs(np.linalg.norm(row[4:7]) - 1) < 1e-9


def test_broad_phase_brute_force(O, na):
    """Oracle pair sets equal an independent NumPy n^2 check on random AABBs, with world -1
    shapes, mixed collision groups and same-body filter pairs."""
    rng = np.random.default_rng(3)
    b = na.ModelBuilder()
    for w in range(6):
        b.begin_world()
        for _ in range(int(rng.integers(2, 9))):
            body = b.add_body(xform=(*rng.uniform(-1, 1, 3), 0, 0, 0, 1))
            cfg = na.ModelBuilder.ShapeConfig(collision_group=int(rng.choice([1, 1, 2, -1, 0])))
            b.add_shape_sphere(body, radius=float(rng.uniform(0.05, 0.4)), cfg=cfg)
            if rng.random() < 0.3:
                b.add_shape_box(body, xform=(0.1, 0, 0, 0, 0, 0, 1), hx=0.1, hy=0.1, hz=0.1, cfg=cfg)
        b.end_world()
    b.add_ground_plane()
    b.add_shape_box(-1, xform=(0, 0, 0.5, 0, 0, 0, 1), hx=0.3, hy=0.3, hz=0.3)
    m = b.finalize("cpu")
    om = O.OracleModel(m)
    q = m.host("body_q").astype(np.float32).astype(np.float64)
    lo, hi = O.shape_aabbs(om, q)
    sw, sg = m.host("shape_world"), m.host("shape_collision_group")
    filt = set(map(tuple, m.host("shape_collision_filter_pairs").tolist()))

    def group_ok(a, b_):
        if a == 0 or b_ == 0:
            return False
        if a > 0:
            return a == b_ or b_ < 0
        return a != b_

    total = 0
    for w in range(m.world_count):
        seg = [s for s in range(m.shape_count) if sw[s] == w or sw[s] == -1]
        expect = set()
        for i in range(len(seg)):
            for j in range(i + 1, len(seg)):
                a, c = sorted((seg[i], seg[j]))
                if sw[a] == -1 and sw[c] == -1:
                    continue
                if not group_ok(sg[a], sg[c]) or (a, c) in filt:
                    continue
                if np.all(lo[a] <= hi[c]) and np.all(hi[a] >= lo[c]):
                    expect.add((a, c))
        got = set(map(tuple, O.world_pairs(om, w, lo, hi).tolist()))
        assert got == expect
        total += len(got)
    assert total > 10


def test_aabb_closed_forms(O, na):
    b = na.ModelBuilder()
    b.begin_world()
    q = (0.0, 0.0, math.sin(0.3), math.cos(0.3))  # yaw 0.6 rad
    s = b.add_body(xform=(1, 2, 3, *q))
    b.add_shape_box(s, hx=0.1, hy=0.2, hz=0.3, cfg=na.ModelBuilder.ShapeConfig(margin=0.01, gap=0.02))
    c = b.add_body(xform=(0, 0, 1, *IDQ))
    b.add_shape_capsule(c, radius=0.05, half_height=0.2, cfg=na.ModelBuilder.ShapeConfig(gap=0.0))
    b.end_world()
    b.add_ground_plane()
    m = b.finalize("cpu")
    om = O.OracleModel(m)
    lo, hi = O.shape_aabbs(om, m.host("body_q").astype(np.float32).astype(np.float64))
    R = _qmat(q)
    ext = np.abs(R) @ np.array([0.1, 0.2, 0.3]) + 0.03
    assert np.allclose(lo[0], np.array([1, 2, 3]) - ext, atol=1e-6)
    assert np.allclose(hi[0], np.array([1, 2, 3]) + ext, atol=1e-6)
    assert np.allclose(lo[1], [-0.05, -0.05, 1 - 0.25], atol=1e-7)
    assert np.allclose(hi[1], [0.05, 0.05, 1 + 0.25], atol=1e-7)
    assert np.allclose(hi[2], [1e6 + 0.1] * 3, rtol=1e-7)  # infinite plane: bounding-sphere cube + gap


def _box_grid(na, worlds=2, side=4, half=0.04, spacing=0.004, gap=0.1, device="cpu", seed=5):
    """Dense worlds of side^3 boxes with small jitter, far more candidate contacts than a world's
    128-contact set at a 10 cm gap (the builder's default)."""
    B = na.ModelBuilder
    builder = B()
    rng = np.random.default_rng(seed)
    cfg = B.ShapeConfig(density=1000.0, ke=1.0e4, mu=0.5, gap=gap)
    for _ in range(worlds):
        builder.begin_world()
        pitch = 2 * half + spacing
        for i in range(side):
            for j in range(side):
                for k in range(side):
                    p = np.array([i, j, k]) * pitch + rng.uniform(-0.3, 0.3, 3) * spacing + [0.0, 0.0, half + 0.01]
                    yaw = rng.uniform(-0.2, 0.2)
                    b = builder.add_body(xform=(*p, 0.0, 0.0, math.sin(0.5 * yaw), math.cos(0.5 * yaw)))
                    builder.add_shape_box(b, hx=half, hy=half, hz=half, cfg=cfg)
        builder.end_world()
    builder.add_ground_plane(cfg=B.ShapeConfig(ke=1.0e4, mu=0.8, gap=gap))
    return builder.finalize(device)


def test_contact_overflow_keeps_the_closest(O, na):
    """A world with more contacts than its set keeps the 128 with the smallest gap distance: the
    kept rows are rows of the full list, the count reports every contact, and no dropped contact
    is closer than a kept one (float32 distances, as the device compares them)."""
    m = _box_grid(na)
    om = O.OracleModel(m)
    q = m.host("body_q").astype(np.float32).astype(np.float64)
    for w in range(m.world_count):
        n_all, _, full = O.collide_world(om, w, q, cap=8192)
        n_kept, _, kept = O.collide_world(om, w, q, cap=128)
        assert n_all == n_kept > 128 and len(kept) == 128
        key = lambda r: (int(r[0]), int(r[1]), int(r[2]))
        index = {key(r): r for r in full}
        assert len(index) == len(full)
        assert all(key(r) in index and np.array_equal(r, index[key(r)]) for r in kept)
        kept_keys = {key(r) for r in kept}
        dropped = np.array([r[10] for r in full if key(r) not in kept_keys], dtype=np.float32)
        assert len(dropped) == n_all - 128
        assert np.float32(kept[:, 10]).max() <= dropped.min()
        # the plain clamp (the first 128 in pair order) would fail that check on these worlds
        first = np.float32(full[:128, 10])
        assert first.max() > np.float32(full[128:, 10]).min()
