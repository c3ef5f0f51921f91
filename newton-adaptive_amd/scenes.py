"""Seeded synthetic scenes for the step-doubling rigid-body workloads.

World ``w`` draws from ``numpy.random.default_rng(seed + w)`` and orientations are
Shoemake-uniform quaternions, as in the reference scene script
``scripts/scenes/contact_objects.py:136-185``.  The bin walls follow
``contact_objects.py:81-100`` (static boxes in world ``-1``).
"""

from __future__ import annotations

import math

import numpy as np

from .builder import ModelBuilder
from .model import GeoType


def random_unit_quaternion(rng) -> tuple:
    """Shoemake uniform rotation, ``(x, y, z, w)`` (``contact_objects.py:136``)."""
    u1, u2, u3 = rng.random(), rng.random(), rng.random()
    s1 = math.sqrt(1.0 - u1)
    s2 = math.sqrt(u1)
    a1 = 2.0 * math.pi * u2
    a2 = 2.0 * math.pi * u3
    return (s1 * math.sin(a1), s1 * math.cos(a1), s2 * math.sin(a2), s2 * math.cos(a2))


def _add_bin(builder: ModelBuilder, half_inner: float, cfg, wall_half_height: float = 0.5,
             thickness: float = 0.025) -> None:
    for px, py, hx, hy in [
        (-(half_inner + thickness), 0.0, thickness, half_inner + thickness),
        (half_inner + thickness, 0.0, thickness, half_inner + thickness),
        (0.0, -(half_inner + thickness), half_inner + thickness, thickness),
        (0.0, half_inner + thickness, half_inner + thickness, thickness),
    ]:
        builder.add_shape_box(-1, xform=(px, py, wall_half_height, 0.0, 0.0, 0.0, 1.0), hx=hx, hy=hy,
                              hz=wall_half_height, cfg=cfg)


def box_drop(world_count: int = 64, seed: int = 0, device="cuda:0", ke: float = 1.0e4, mu: float = 0.5):
    """One 0.1 m, 1 kg box per world dropped from z ~ U[0.3, 0.8] with a random orientation
    onto a ground plane (friction 0.5)."""
    builder = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig(density=1000.0, ke=ke, mu=mu)
    for w in range(world_count):
        rng = np.random.default_rng(seed + w)
        builder.begin_world()
        z = rng.uniform(0.3, 0.8)
        q = random_unit_quaternion(rng)
        b = builder.add_body(xform=(0.0, 0.0, z, *q))
        builder.add_shape_box(b, hx=0.05, hy=0.05, hz=0.05, cfg=cfg)
        builder.end_world()
    builder.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=ke, mu=mu))
    return builder.finalize(device)


def box_stack(world_count: int = 1024, seed: int = 0, device="cuda:0", ke: float = 1.0e6, layers: int = 4):
    """Stack of ``layers`` boxes per world: edge ~ U[0.05, 0.15] m, density 1000,
    stiffness ``ke``, one friction coefficient ~ U[0.3, 1.0] per world.  Boxes start
    1 mm apart, centred with a small seeded offset and yaw."""
    builder = ModelBuilder()
    for w in range(world_count):
        rng = np.random.default_rng(seed + w)
        mu = rng.uniform(0.3, 1.0)
        cfg = ModelBuilder.ShapeConfig(density=1000.0, ke=ke, mu=mu)
        builder.begin_world()
        z = 0.0
        for _ in range(layers):
            edge = rng.uniform(0.05, 0.15)
            h = 0.5 * edge
            z += h + 1.0e-3
            dx, dy = rng.uniform(-0.1, 0.1, size=2) * h
            yaw = rng.uniform(-0.3, 0.3)
            q = (0.0, 0.0, math.sin(0.5 * yaw), math.cos(0.5 * yaw))
            b = builder.add_body(xform=(dx, dy, z, *q))
            builder.add_shape_box(b, hx=h, hy=h, hz=h, cfg=cfg)
            z += h
        builder.end_world()
    builder.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=ke, mu=0.8))
    return builder.finalize(device)


SIZE_MEDIAN = 0.08
SIZE_SIGMA = 0.4


def _add_random_primitive(builder: ModelBuilder, rng, cfg, placed: list):
    """Add one random primitive; returns nothing.  Sizes are lognormal(ln 0.08, 0.4) clipped
    to +-2 sigma so that a drop position can always be found inside the bin."""
    kind = int(rng.integers(0, 3))
    s = float(np.clip(rng.lognormal(math.log(SIZE_MEDIAN), SIZE_SIGMA),
                      SIZE_MEDIAN * math.exp(-2 * SIZE_SIGMA), SIZE_MEDIAN * math.exp(2 * SIZE_SIGMA)))
    q = random_unit_quaternion(rng)
    if kind == 0:
        scale, bound = (0.5 * s,), 0.5 * s
    elif kind == 1:
        scale = tuple(0.5 * s * rng.uniform(0.6, 1.0, size=3))
        bound = float(np.linalg.norm(scale))
    else:
        scale, bound = (0.3 * s, 0.5 * s), 0.8 * s
    x, y = rng.uniform(-0.45 + bound, 0.45 - bound, size=2) if bound < 0.45 else (0.0, 0.0)
    # drop height: clear every bounding sphere already placed above this spot
    z = bound + 0.02
    for (px, py, pz, pr) in placed:
        if (px - x) ** 2 + (py - y) ** 2 < (pr + bound) ** 2:
            z = max(z, pz + pr + bound + 0.005)
    placed.append((x, y, z, bound))
    b = builder.add_body(xform=(x, y, z, *q))
    if kind == 0:
        builder.add_shape_sphere(b, radius=scale[0], cfg=cfg)
    elif kind == 1:
        builder.add_shape_box(b, hx=scale[0], hy=scale[1], hz=scale[2], cfg=cfg)
    else:
        builder.add_shape_capsule(b, radius=scale[0], half_height=scale[1], cfg=cfg)


PILE_GAP = 0.02


def uneven_pile(world_count: int = 4096, seed: int = 0, device="cuda:0", max_prims: int = 32,
                min_prims: int = 1, ke: float = 1.0e4, counts=None, gap: float = PILE_GAP):
    """Uneven worlds: N ~ U{min_prims..max_prims} primitives from {sphere, box, capsule},
    size ~ lognormal(ln 0.08 m, 0.4), dropped as a random pile into a 1 m^2 walled bin.
    Drop heights are chosen so that no two bounding spheres start overlapping.  Every shape
    uses a 2 cm contact gap (``ShapeConfig.gap``) by default, which keeps a settled pile within
    the 128-contact per-world set; with the builder default (0.1 m) a 32-primitive world
    produces up to ~300 speculative contacts, of which it keeps the 128 closest."""
    builder = ModelBuilder()
    for w in range(world_count):
        rng = np.random.default_rng(seed + w)
        n = int(counts[w]) if counts is not None else int(rng.integers(min_prims, max_prims + 1))
        mu = rng.uniform(0.3, 1.0)
        cfg = ModelBuilder.ShapeConfig(density=1000.0, ke=ke, mu=mu, gap=gap)
        builder.begin_world()
        placed: list = []
        for _ in range(n):
            _add_random_primitive(builder, rng, cfg, placed)
        builder.end_world()
    ground = ModelBuilder.ShapeConfig(ke=ke, mu=0.8, gap=gap)
    builder.add_ground_plane(cfg=ground)
    _add_bin(builder, 0.5, ModelBuilder.ShapeConfig(ke=ke, mu=0.5, gap=gap))
    return builder.finalize(device)


def serving_requests(count: int, seed: int = 0, rate: float | None = None, max_prims: int = 64,
                     horizons=(200, 2000), tols=(1.0e-2, 1.0e-3, 1.0e-4), device="cpu", distinct: int | None = None,
                     articulated_share: float = 0.0, return_articulated_model: bool = False):
    """Rollout requests of the serving workload: request ``i`` starts from world ``i % distinct``
    of ``uneven_pile(distinct, seed, max_prims=max_prims)`` (1..max_prims primitives), with its
    own horizon of U{horizons} accepted steps, a tolerance drawn from ``tols`` and Poisson
    arrivals at ``rate`` per second (all at t = 0 when ``rate`` is None).  ``distinct`` defaults
    to ``count``.  A seeded ``articulated_share`` of the requests are articulated worlds instead
    (an arm or a quadruped on box terrain, from ``articulated_mix``), with the same horizon,
    tolerance and arrival draws.  Returns ``(model, requests)``; the pile model's global shapes
    (the ground, with the same material as the articulated scenes' ground) are the shared
    environment.  Request ``i`` records its source in ``rid`` (= i) and, for articulated requests,
    ``source_world`` (the world of the articulated model it was taken from);
    ``return_articulated_model`` adds that model (on ``device``, None without articulated
    requests) as a third return value."""
    from .serving import RolloutRequest, requests_from_model, world_arrays

    distinct = count if distinct is None else min(int(distinct), count)
    model = uneven_pile(distinct, seed=seed, device=device, max_prims=max_prims)
    rng = np.random.default_rng(seed + 1_000_003)
    H = rng.integers(int(horizons[0]), int(horizons[1]) + 1, size=count)
    T = rng.choice(np.asarray(tols, dtype=np.float64), size=count)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count)) if rate else np.zeros(count)
    reqs = requests_from_model(model, horizon_accepted=[int(h) for h in H], tol=[float(t) for t in T],
                               arrivals=[float(a) for a in arrivals], worlds=[i % distinct for i in range(count)])
    amodel = None
    if articulated_share > 0.0:
        art = np.flatnonzero(np.random.default_rng(seed + 2_000_003).random(count) < float(articulated_share))
        if len(art):
            n_art = max(1, min(len(art), int(math.ceil(distinct * float(articulated_share)))))
            amodel = articulated_mix(n_art, seed=seed + 7, device=device)
            for k, i in enumerate(art):
                b, s, j = world_arrays(amodel, k % n_art)
                reqs[i] = RolloutRequest(b, s, j, horizon_accepted=int(H[i]), tol=float(T[i]),
                                         arrival=float(arrivals[i]), rid=int(i))
                reqs[i].source_world = k % n_art
    if return_articulated_model:
        return model, reqs, amodel
    return model, reqs


def shape_kinds(model) -> np.ndarray:
    """Per-world histogram of body shape types (diagnostic)."""
    st = model.host("shape_type")
    return np.bincount(st, minlength=GeoType.BOX + 1)


# ---------------------------------------------------------------- articulated worlds
ROBOT_GROUP = -1  # one negative group: links of a robot never collide with each other


def _terrain(builder, rng, cfg, count, half_extent, height=(0.02, 0.08)):
    """``count`` static boxes of the current world (random size, yaw and height) on the ground."""
    for _ in range(count):
        hx, hy = rng.uniform(0.05, 0.2, size=2)
        hz = rng.uniform(*height) * 0.5
        x, y = rng.uniform(-half_extent, half_extent, size=2)
        yaw = rng.uniform(-math.pi, math.pi)
        q = (0.0, 0.0, math.sin(0.5 * yaw), math.cos(0.5 * yaw))
        builder.add_shape_box(-1, xform=(x, y, hz, *q), hx=hx, hy=hy, hz=hz, cfg=cfg)


def _add_arm(builder, rng, ke=1.0e4):
    """7-dof revolute arm on a fixed base: capsule links of 0.16-0.26 m, joint axes alternating
    z / y, PD drives to seeded random targets."""
    link_cfg = ModelBuilder.ShapeConfig(density=1000.0, ke=ke, mu=0.6, gap=PILE_GAP, collision_group=ROBOT_GROUP)
    axes = [(0, 0, 1), (0, 1, 0), (0, 0, 1), (0, 1, 0), (0, 0, 1), (0, 1, 0), (0, 0, 1)]
    parent, offset = -1, (0.0, 0.0, 0.05, 0.0, 0.0, 0.0, 1.0)
    for k in range(7):
        half = float(rng.uniform(0.08, 0.13))
        link = builder.add_link()
        builder.add_shape_capsule(link, xform=(0.0, 0.0, half, 0.0, 0.0, 0.0, 1.0), radius=0.035, half_height=half,
                                  cfg=link_cfg)
        target = float(rng.uniform(-0.8, 0.8))
        builder.add_joint_revolute(parent, link, parent_xform=offset, axis=axes[k], target_pos=target,
                                   target_ke=60.0, target_kd=4.0, armature=0.01,
                                   q0=float(rng.uniform(-0.3, 0.3)))
        parent, offset = link, (0.0, 0.0, 2.0 * half, 0.0, 0.0, 0.0, 1.0)


def _add_quadruped(builder, rng, ke=1.0e4):
    """Floating-base quadruped: box trunk, four legs of hip (x axis), thigh and knee (y axis)
    joints, capsule segments and sphere feet; PD drives to a standing pose with seeded offsets."""
    link_cfg = ModelBuilder.ShapeConfig(density=500.0, ke=ke, mu=0.8, gap=PILE_GAP, collision_group=ROBOT_GROUP)
    z0 = 0.42
    trunk = builder.add_link(xform=(0.0, 0.0, z0, 0.0, 0.0, 0.0, 1.0))
    builder.add_shape_box(trunk, hx=0.25, hy=0.1, hz=0.05, cfg=link_cfg)
    builder.add_joint_free(trunk)
    for sx in (1.0, -1.0):
        for sy in (1.0, -1.0):
            hip = builder.add_link()
            builder.add_shape_sphere(hip, radius=0.035, cfg=link_cfg)
            builder.add_joint_revolute(trunk, hip, parent_xform=(0.2 * sx, 0.1 * sy, 0.0, 0.0, 0.0, 0.0, 1.0),
                                       axis=(1, 0, 0), target_pos=float(rng.uniform(-0.15, 0.15)), target_ke=40.0,
                                       target_kd=1.0, armature=0.01)
            thigh = builder.add_link()
            builder.add_shape_capsule(thigh, xform=(0.0, 0.0, -0.1, 0.0, 0.0, 0.0, 1.0), radius=0.025,
                                      half_height=0.08, cfg=link_cfg)
            t0 = 0.6 + float(rng.uniform(-0.2, 0.2))
            builder.add_joint_revolute(hip, thigh, parent_xform=(0.0, 0.05 * sy, 0.0, 0.0, 0.0, 0.0, 1.0),
                                       axis=(0, 1, 0), target_pos=t0, target_ke=40.0, target_kd=1.0,
                                       armature=0.01, q0=0.6)
            shank = builder.add_link()
            builder.add_shape_capsule(shank, xform=(0.0, 0.0, -0.1, 0.0, 0.0, 0.0, 1.0), radius=0.02,
                                      half_height=0.08, cfg=link_cfg)
            builder.add_shape_sphere(shank, xform=(0.0, 0.0, -0.2, 0.0, 0.0, 0.0, 1.0), radius=0.025, cfg=link_cfg)
            k0 = -1.2 + float(rng.uniform(-0.2, 0.2))
            builder.add_joint_revolute(thigh, shank, parent_xform=(0.0, 0.0, -0.2, 0.0, 0.0, 0.0, 1.0),
                                       axis=(0, 1, 0), target_pos=k0, target_ke=40.0, target_kd=1.0,
                                       armature=0.01, q0=-1.2)


def articulated_mix(world_count: int = 8192, seed: int = 0, device="cuda:0", kinds=None, terrain_boxes: int = 6):
    """Articulated worlds: world ``w`` is a 7-dof arm when ``kinds[w] == 0`` (default: even worlds)
    and a 12-dof floating quadruped (18 dofs) otherwise, with seeded random joint targets, over
    ``terrain_boxes`` random static boxes and the ground plane."""
    builder = ModelBuilder()
    terrain_cfg = ModelBuilder.ShapeConfig(ke=1.0e4, mu=0.8, gap=PILE_GAP)
    for w in range(world_count):
        rng = np.random.default_rng(seed + w)
        kind = (w % 2) if kinds is None else int(kinds[w])
        builder.begin_world()
        if kind == 0:
            _add_arm(builder, rng)
            _terrain(builder, rng, terrain_cfg, terrain_boxes, 0.8)
        else:
            _add_quadruped(builder, rng)
            _terrain(builder, rng, terrain_cfg, terrain_boxes, 0.6)
        builder.end_world()
    builder.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, mu=0.8, gap=PILE_GAP))
    return builder.finalize(device)
