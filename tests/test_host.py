"""Host-side checks that need no GPU: the C-ABI library loads and exports every symbol
``include/nadapt.h`` declares, the builder lays worlds out as the kernels assume, scenes are
seeded and deterministic, and invalid solver settings raise ``ValueError`` like the reference
(``solver_sap_adaptive.py:463-474``, ``solver_mujoco_adaptive.py:477-488``)."""

import math
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(REPO, "include", "nadapt.h")).read()
    return re.findall(r"^\s*(?:const char\*|int|int64_t)\s+(na_\w+)\(", text, re.M)


def test_library_exports_every_header_symbol(nalib):
    lib = nalib.lib()
    names = header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(nalib.SIGNATURES)


def test_builder_world_layout(na):
    b = na.ModelBuilder()
    for w in range(3):
        b.begin_world()
        for k in range(w + 1):
            body = b.add_body(xform=(k, 0, 1, 0, 0, 0, 1))
            b.add_shape_sphere(body, radius=0.1)
            b.add_shape_box(body, xform=(0.1, 0, 0, 0, 0, 0, 1), hx=0.05, hy=0.05, hz=0.05)
        b.end_world()
    b.add_ground_plane()
    m = b.finalize("cpu")
    assert m.world_count == 3 and m.body_count == 6 and m.shape_count == 13
    assert list(m.host("world_body_start")) == [0, 1, 3, 6]
    assert list(m.host("world_shape_start")) == [0, 2, 6, 12]
    assert list(m.host("global_shapes")) == [12]
    # same-body shape pairs are filtered
    assert len(m.host("shape_collision_filter_pairs")) == 6
    # composite mass: sphere + box about the COM
    r, h = 0.1, 0.05
    ms = 1000 * 4 / 3 * math.pi * r ** 3
    mb = 1000 * 8 * h ** 3
    assert m.host("body_mass")[0] == pytest.approx(ms + mb)
    assert m.host("body_com")[0][0] == pytest.approx(0.1 * mb / (ms + mb))
    assert m.joint_coord_count == 7 * 6 and m.joint_dof_count == 6 * 6


def test_builder_rejects_interleaved_worlds(na):
    b = na.ModelBuilder()
    b.begin_world()
    s0 = b.add_body()
    b.add_shape_sphere(s0, radius=0.1)
    b.end_world()
    b.begin_world()
    s1 = b.add_body()
    b.end_world()
    b.add_shape_sphere(s1, radius=0.1)
    b.add_shape_sphere(s0, radius=0.1)  # world 0 shape after world 1 shapes
    with pytest.raises(ValueError):
        b.finalize("cpu")


def test_scenes_seeded(na):
    a = na.scenes.uneven_pile(16, seed=3, device="cpu")
    b = na.scenes.uneven_pile(16, seed=3, device="cpu")
    assert np.array_equal(a.host("body_q"), b.host("body_q"))
    n = a.bodies_per_world()
    assert n.min() >= 1 and n.max() <= 32
    c = na.scenes.box_drop(64, seed=0, device="cpu")
    assert np.allclose(c.host("body_mass"), 1.0)
    z = c.host("body_q")[:, 2]
    assert z.min() >= 0.3 and z.max() <= 0.8
    q = c.host("body_q")[:, 3:]
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    s = na.scenes.box_stack(8, seed=0, device="cpu")
    edge = 2 * s.host("shape_scale")[:-1, 0]
    assert edge.min() >= 0.05 and edge.max() <= 0.15
    assert np.all(s.host("shape_material_ke")[:-1] == 1e6)


@pytest.mark.parametrize("kw", [dict(mode="bogus"), dict(tol=0.0), dict(dt_inner_init=-1.0), dict(dt_inner_min=0.0),
                                dict(dt_inner_max=0.0), dict(max_substeps=0), dict(max_rigid_contact=0),
                                dict(line_search_variant="nope"), dict(broad_phase="bvh")])
def test_invalid_settings_raise(na, kw):
    m = na.scenes.box_drop(2, device="cpu")
    with pytest.raises(ValueError):
        na.SolverSAPAdaptive(m, **kw)


@pytest.mark.parametrize("kw", [dict(dt_mode="global"), dict(tiling="even"), dict(max_substeps=0)])
def test_invalid_mujoco_settings_raise(na, kw):
    m = na.scenes.box_drop(2, device="cpu")
    with pytest.raises(ValueError):
        na.SolverMuJoCoAdaptive(m, **kw)


def test_cpu_model_is_refused_by_the_native_path(na):
    m = na.scenes.box_drop(2, device="cpu")
    with pytest.raises(RuntimeError):
        na.SolverSAPAdaptive(m)
