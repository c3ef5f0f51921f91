"""MuJoCo-path adapters: FREE-joint coordinate conversion (HIP kernels against the FP64
restatement of kernels.py:578-789) and contact-parameter mixing (kernels.py:127-207).

The conversion is pinned by closed forms: a body turned 90 degrees about z with its COM 0.1 m
off the origin, spinning at 2 rad/s about z while its COM moves at 1 m/s along x, has origin
velocity v_com - w x (R c) = (1.2, 0, 0) and body-frame angular velocity (0, 0, 2); the two
directions are inverse to each other.  convert_solref is pinned by its formula and the default
limit solref; the mixing rules by hand-computed pairs.
"""

import importlib
import math

import numpy as np
import pytest
import torch

na = importlib.import_module("newton-adaptive_amd")
mj = importlib.import_module("newton-adaptive_amd.mujoco_adapter")
DEV = "cuda:0"


def test_oracle_free_joint_closed_form(O):
    s = math.sqrt(0.5)
    jq = np.array([[0.3, -0.2, 0.5, 0.0, 0.0, s, s]])
    jqd = np.array([[1.0, 0.0, 0.0, 0.0, 0.0, 2.0]])
    com = np.array([[0.1, 0.0, 0.0]])
    qpos, qvel = O.coords_to_mujoco(com, jq, jqd)
    assert np.allclose(qpos, [[0.3, -0.2, 0.5, s, 0.0, 0.0, s]], atol=1e-15)
    assert np.allclose(qvel, [[1.2, 0.0, 0.0, 0.0, 0.0, 2.0]], atol=1e-15)
    back_q, back_qd = O.coords_from_mujoco(com, qpos, qvel)
    assert np.allclose(back_q, jq, atol=1e-15) and np.allclose(back_qd, jqd, atol=1e-15)


def test_oracle_round_trip_random(O):
    rng = np.random.default_rng(3)
    n = 50
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    jq = np.concatenate([rng.normal(size=(n, 3)), q], axis=1)
    jqd = rng.normal(size=(n, 6))
    com = rng.normal(size=(n, 3)) * 0.1
    qpos, qvel = O.coords_to_mujoco(com, jq, jqd)
    # the origin velocity differs from the COM velocity by w x (R c): never by a constant
    assert np.abs(qvel[:, :3] - jqd[:, :3]).max() > 1e-3
    back_q, back_qd = O.coords_from_mujoco(com, qpos, qvel)
    assert np.allclose(back_q, jq, atol=1e-13) and np.allclose(back_qd, jqd, atol=1e-13)


def test_oracle_articulated_joints_closed_form(O):
    """FREE root under a parent frame turned 90 degrees about z, followed by two revolute joints:
    the parent-frame COM velocity (1, 0, 0) is (0, 1, 0) in the world; with the COM 0.1 m along
    world x and spin 2 rad/s about z the origin moves at (0, 1, 0) - (0, 0, 2) x (0.1, 0, 0) =
    (0, 0.8, 0); revolute entries pass through."""
    s = math.sqrt(0.5)
    jt = [4, 1, 1]
    qs, ds = [0, 7, 8, 9], [0, 6, 7, 8]
    Xp = np.array([[0, 0, 0, 0, 0, s, s], [0, 0, 0.1, 0, 0, 0, 1], [0, 0, 0.1, 0, 0, 0, 1]], dtype=np.float64)
    com = np.array([[0.1, 0.0, 0.0], [0.0, 0.0, 0.05], [0.0, 0.0, 0.05]])
    jq = np.array([0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0, 0.3, -0.7])
    jqd = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 2.0, 1.5, -2.5])
    qpos, qvel = O.coords_to_mujoco_joints(jt, qs, ds, Xp, com, jq, jqd)
    assert np.allclose(qpos, [0.0, 0.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.3, -0.7], atol=1e-15)
    assert np.allclose(qvel, [0.0, 0.8, 0.0, 0.0, 0.0, 2.0, 1.5, -2.5], atol=1e-15)


def test_convert_solref_formula_and_default():
    r = mj.convert_solref([1.0e4, 0.0, 2.5e3], [100.0, 100.0, -1.0], 1.0, 1.0)
    assert np.allclose(r[0], [0.02, 0.5])  # timeconst 2/(kd w), dampratio kd/2 sqrt(d_r/ke)
    assert np.allclose(r[1], mj.DEFAULT_LIMIT_SOLREF) and np.allclose(r[2], mj.DEFAULT_LIMIT_SOLREF)
    r2 = mj.convert_solref(4.0e4, 200.0, d_width=0.5, d_r=0.9)
    assert np.allclose(r2, [2.0 / (200.0 * 0.5), 100.0 * math.sqrt(0.9 / 4.0e4)])


def test_contact_params_mixing_rules():
    S = 4
    kw = dict(priority=[0, 0, 1, 0], solmix=[1.0, 3.0, 1.0, 0.0],
              solref=[[0.02, 1.0], [0.04, 0.5], [0.01, 2.0], [-100.0, -1.0]],
              solimp=np.tile(np.arange(5.0), (S, 1)) + np.arange(S)[:, None],
              friction=[[0.5, 0.01, 0.001], [0.8, 0.0, 0.0], [0.2, 0.02, 0.0], [1e-9, 0.0, 0.0]],
              margin=[0.01, 0.02, 0.0, 0.0], gap=[0.1, 0.0, 0.05, 0.0], condim=[3, 1, 6, 4])
    p = mj.contact_params([[0, 1], [0, 2], [1, 3], [3, 3]], **kw)
    # equal priority: mix = 1 / (1 + 3), solref mixed, larger condim, elementwise max friction
    assert p["mix"][0] == 0.25 and np.allclose(p["solref"][0], 0.25 * np.array([0.02, 1.0]) + 0.75 * np.array([0.04, 0.5]))
    assert p["condim"][0] == 3 and np.allclose(p["friction"][0], [0.8, 0.8, 0.01, 0.001, 0.001])
    assert np.isclose(p["margin"][0], 0.03) and np.isclose(p["gap"][0], 0.1)
    # shape 2 has the higher priority: everything from shape 2
    assert p["mix"][1] == 0.0 and p["condim"][1] == 6 and np.allclose(p["friction"][1], [0.2, 0.2, 0.02, 1e-5, 1e-5])
    assert np.allclose(p["solref"][1], [0.01, 2.0]) and np.allclose(p["solimp"][1], kw["solimp"][2])
    # a non-positive time constant on one side takes the elementwise minimum; solmix 0 on one side
    assert p["mix"][2] == 1.0 and np.allclose(p["solref"][2], [-100.0, -1.0])
    # both solmix 0: mix 0.5; friction floored at MJ_MINMU
    assert p["mix"][3] == 0.5 and p["friction"][3].min() == mj.MJ_MINMU


def _offset_com_model(device, worlds=32, seed=5):
    """Bodies made of a box and a sphere off the body origin, so the COM is not the origin."""
    rng = np.random.default_rng(seed)
    b = na.ModelBuilder()
    for _ in range(worlds):
        b.begin_world()
        for _ in range(3):
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            body = b.add_body(xform=(*rng.uniform(-0.3, 0.3, 3), *q))
            b.add_shape_box(body, xform=(0.1, 0.02, 0.0, 0.0, 0.0, 0.0, 1.0), hx=0.05, hy=0.04, hz=0.03)
            b.add_shape_sphere(body, xform=(-0.03, 0.05, 0.04, 0.0, 0.0, 0.0, 1.0), radius=0.02)
        b.end_world()
    b.add_ground_plane()
    return b.finalize(device)


@pytest.mark.gpu
def test_gpu_conversion_matches_oracle_and_round_trips(O):
    model = _offset_com_model(DEV)
    assert np.abs(model.host("body_com")).max() > 0.01
    state = model.state()
    rng = np.random.default_rng(5)
    state.body_qd.copy_(torch.as_tensor(rng.normal(size=(model.body_count, 6)), dtype=torch.float32))
    qpos, qvel = mj.to_mujoco(model, state)
    torch.cuda.synchronize()
    com = model.body_com.cpu().numpy()
    rq, rv = O.coords_to_mujoco(com, state.body_q.cpu().numpy(), state.body_qd.cpu().numpy())
    assert np.abs(qpos.cpu().numpy() - rq).max() <= 1e-6
    assert np.abs(qvel.cpu().numpy() - rv).max() <= 2e-6 * max(1.0, np.abs(rv).max())
    s2 = model.state()
    mj.from_mujoco(model, qpos, qvel, s2)
    torch.cuda.synchronize()
    assert np.array_equal(s2.body_q.cpu().numpy(), state.body_q.cpu().numpy())
    assert np.abs(s2.body_qd.cpu().numpy() - state.body_qd.cpu().numpy()).max() <= 2e-6 * max(
        1.0, np.abs(state.body_qd.cpu().numpy()).max())
    with pytest.raises(ValueError):
        mj.to_mujoco(model, state, qpos=torch.zeros(3, device=DEV))


@pytest.mark.gpu
def test_gpu_articulated_conversion_matches_oracle_and_round_trips(O):
    """Arms and floating-base quadrupeds: FREE roots against the joint-wise restatement, revolute
    coordinates copied bit for bit, and the inverse conversion restores the state."""
    model = na.scenes.articulated_mix(32, seed=6, device=DEV)
    state = model.state()
    rng = np.random.default_rng(6)
    state.joint_qd.copy_(torch.as_tensor(rng.normal(size=model.joint_dof_count), dtype=torch.float32))
    qpos, qvel = mj.to_mujoco(model, state)
    torch.cuda.synchronize()
    assert qpos.shape == (model.joint_coord_count,) and qvel.shape == (model.joint_dof_count,)
    jq, jqd = state.joint_q.cpu().numpy(), state.joint_qd.cpu().numpy()
    jt = model.host("joint_type")
    rq, rv = O.coords_to_mujoco_joints(jt, model.host("joint_q_start"), model.host("joint_qd_start"),
                                       model.host("joint_X_p"), model.host("body_com"), jq, jqd)
    gq, gv = qpos.cpu().numpy(), qvel.cpu().numpy()
    assert (jt == 4).sum() > 0 and (jt == 1).sum() > 0
    assert np.abs(gq - rq).max() <= 1e-6
    assert np.abs(gv - rv).max() <= 2e-6 * max(1.0, np.abs(rv).max())
    rev_q = np.concatenate([np.arange(a, b) for a, b, t in zip(model.host("joint_q_start")[:-1],
                                                                model.host("joint_q_start")[1:], jt) if t == 1])
    rev_d = np.concatenate([np.arange(a, b) for a, b, t in zip(model.host("joint_qd_start")[:-1],
                                                                model.host("joint_qd_start")[1:], jt) if t == 1])
    assert np.array_equal(gq[rev_q], jq[rev_q]) and np.array_equal(gv[rev_d], jqd[rev_d])
    s2 = model.state()
    mj.from_mujoco(model, qpos, qvel, s2)
    torch.cuda.synchronize()
    assert np.array_equal(s2.joint_q.cpu().numpy(), jq)
    assert np.abs(s2.joint_qd.cpu().numpy() - jqd).max() <= 2e-6 * max(1.0, np.abs(jqd).max())
    with pytest.raises(ValueError):
        mj.to_mujoco(model, state, qpos=torch.zeros((model.body_count, 7), device=DEV))
