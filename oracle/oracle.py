"""ctypes binding of the CPU oracle (``oracle/oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, always as the checker, never as the product path.
``liboracle64.so`` is the FP64 restatement; timed on the host it is also the CPU baseline
("port").  The Python-side restatements here (contact matching, the contact sort key and order,
the joint-wise MuJoCo coordinate conversion) are pinned by hand-made cases and closed forms in
``tests/test_oracle_matching.py`` and ``tests/test_mujoco_adapter.py``.  Parity unpinned, for
want of the absent dependencies: the exact manifold points of the reference's box-box /
capsule-box MPR multicontact, the values of its ``sap_warp`` convex solve and its
``mujoco_warp`` inner step (checked by invariants only).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)
i64p = C.POINTER(C.c_int64)


class OrModel(C.Structure):
    _fields_ = [
        ("world_count", C.c_int), ("body_count", C.c_int), ("shape_count", C.c_int), ("global_count", C.c_int),
        ("world_body_start", i32p), ("world_shape_start", i32p), ("global_shapes", i32p),
        ("body_com", f64p), ("body_mass", f64p), ("body_inv_mass", f64p), ("body_inertia", f64p),
        ("body_inv_inertia", f64p),
        ("shape_type", i32p), ("shape_body", i32p), ("shape_group", i32p), ("shape_world", i32p),
        ("shape_xform", f64p), ("shape_scale", f64p), ("shape_margin", f64p), ("shape_gap", f64p),
        ("shape_ke", f64p), ("shape_mu", f64p), ("shape_radius", f64p),
        ("filter_count", C.c_int), ("filter_pairs", i32p),
        ("gravity", C.c_double * 3),
        ("mj_solref", f64p), ("mj_solimp", f64p), ("mj_solmix", f64p), ("mj_priority", i32p), ("mj_condim", i32p),
        ("mj_zero_margins", C.c_int),
    ]


class OrConfig(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("controller", C.c_int),
        ("tol", C.c_double), ("dt_init", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("divergence_threshold", C.c_double),
        ("max_substeps", C.c_int), ("max_iterations", C.c_int), ("max_contacts", C.c_int),
        ("opt_rel_tol", C.c_double), ("opt_abs_tol", C.c_double), ("tau_d", C.c_double), ("sigma", C.c_double),
        ("beta", C.c_double), ("ls_rel_tol", C.c_double), ("line_search", C.c_int),
        ("contact_law", C.c_int), ("impratio", C.c_double), ("warm_start", C.c_int),
    ]


class OrCtrl(C.Structure):
    _fields_ = [
        ("dt", f64p), ("dt_half", f64p), ("ideal_dt", f64p), ("sim_time", f64p), ("next_time", f64p),
        ("last_error", f64p), ("accepted_error", f64p),
        ("accepted", u8p), ("commit", u8p), ("diverged", u8p),
        ("substeps", i32p), ("iterations", i32p), ("solve_fail", i32p), ("cum_accepted", i64p),
        ("trend", f64p),
    ]


def build() -> None:
    """Compile both oracle builds with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _ptr(a, t):
    return a.ctypes.data_as(t)


_LIBS: dict = {}


def build_planted(plant: int) -> str:
    """Compile the oracle with planted mistake ``plant`` in the friction projection (-DOR_PLANT,
    see ``project`` in oracle.c) into ``_build/liboracle64_plant<k>.so``; tests use it to show that
    their checks catch such a mistake."""
    out = os.path.join(_BUILD, f"liboracle64_plant{int(plant)}.so")
    os.makedirs(_BUILD, exist_ok=True)
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-ffp-contract=off", "-fopenmp",
                    f"-DOR_PLANT={int(plant)}", "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"], check=True)
    return out


class use_library:
    """Context manager: route every oracle call to the library at ``path`` (a planted build)."""

    def __init__(self, path: str):
        self.path = path

    def __enter__(self):
        lib()
        self.saved = _LIBS["liboracle64.so"]
        _LIBS["liboracle64.so"] = _load(self.path)
        return self

    def __exit__(self, *exc):
        _LIBS["liboracle64.so"] = self.saved
        return False


def library_path(precision: str = "fp64") -> str:
    """Path of the FP64 checker ("fp64") or of the FP32 restatement with FP64 residuals ("fp32", the
    CPU baseline); both are built by the Makefile."""
    path = os.path.join(_BUILD, {"fp64": "liboracle64.so", "fp32": "liboracle32.so"}[precision])
    if not os.path.exists(path):
        build()
    return path


def set_num_threads(n: int) -> None:
    """OpenMP threads of the oracle's parallel loops from now on."""
    lib().or_set_num_threads(int(n))


def lib():
    name = "liboracle64.so"
    if name not in _LIBS:
        path = os.path.join(_BUILD, name)
        if not os.path.exists(path):
            build()
        _LIBS[name] = _load(path)
    return _LIBS[name]


def _load(path):
    L = C.CDLL(path)
    for name in ("or_collide_world_flat", "or_world_pairs_global", "or_collide_shapes", "or_substep_world",
                 "or_step_frame", "or_art_step_frame", "or_cubic_min", "or_num_threads"):
        getattr(L, name).restype = C.c_int
    L.or_cubic_min.argtypes = [C.c_double] * 5 + [f64p]
    L.or_mj_impedance.restype = C.c_double
    L.or_mj_impedance.argtypes = [f64p, C.c_double]
    L.or_mj_contact_params.restype = None
    L.or_mj_contact_params.argtypes = [C.c_void_p, C.c_int, C.c_int, f64p, f64p, f64p, i32p]
    return L


class OracleModel:
    """Host copy of a model's arrays in float64 (values rounded through float32, i.e. the
    numbers the device sees) plus the ctypes descriptor pointing at them."""

    def __init__(self, model):
        h = model._host
        f = lambda k: np.ascontiguousarray(np.asarray(h[k], dtype=np.float32).astype(np.float64))
        i = lambda k: np.ascontiguousarray(np.asarray(h[k], dtype=np.int32))
        self.keep = {
            "world_body_start": i("world_body_start"), "world_shape_start": i("world_shape_start"),
            "global_shapes": i("global_shapes") if len(h["global_shapes"]) else np.zeros(1, np.int32),
            "body_com": f("body_com"), "body_mass": f("body_mass"), "body_inv_mass": f("body_inv_mass"),
            "body_inertia": f("body_inertia"), "body_inv_inertia": f("body_inv_inertia"),
            "shape_type": i("shape_type"), "shape_body": i("shape_body"), "shape_group": i("shape_collision_group"),
            "shape_world": i("shape_world"), "shape_xform": f("shape_transform"), "shape_scale": f("shape_scale"),
            "shape_margin": f("shape_margin"), "shape_gap": f("shape_gap"), "shape_ke": f("shape_material_ke"),
            "shape_mu": f("shape_material_mu"), "shape_radius": f("shape_collision_radius"),
            "filter_pairs": i("shape_collision_filter_pairs").reshape(-1) if len(h["shape_collision_filter_pairs"])
            else np.zeros(2, np.int32),
        }
        k = self.keep
        m = OrModel()
        m.world_count = int(h["world_count"])
        m.body_count = int(h["body_count"])
        m.shape_count = int(h["shape_count"])
        m.global_count = int(len(h["global_shapes"]))
        for name in ("world_body_start", "world_shape_start", "global_shapes", "shape_type", "shape_body",
                     "shape_group", "shape_world", "filter_pairs"):
            setattr(m, name, _ptr(k[name], i32p))
        for name in ("body_com", "body_mass", "body_inv_mass", "body_inertia", "body_inv_inertia", "shape_xform",
                     "shape_scale", "shape_margin", "shape_gap", "shape_ke", "shape_mu", "shape_radius"):
            setattr(m, name, _ptr(k[name], f64p))
        m.filter_count = int(len(h["shape_collision_filter_pairs"]))
        g = np.asarray(h["gravity"], dtype=np.float32).astype(np.float64)
        m.gravity = (C.c_double * 3)(*g)
        # MuJoCo geom parameters (contact law 1), resolved as SolverMuJoCo does
        mj = getattr(model, "mujoco", None)
        if mj is not None and m.shape_count > 0:
            f32 = lambda t: np.ascontiguousarray(t.detach().cpu().numpy().astype(np.float32).astype(np.float64))
            sr = f32(mj.solref).reshape(-1, 2)
            authored = (sr[:, 0] != 0) | (sr[:, 1] != 0)
            conv = mj_convert_solref(k["shape_ke"], np.asarray(h["shape_material_kd"], np.float32).astype(np.float64))
            k["mj_solref"] = np.ascontiguousarray(np.where(authored[:, None], sr, conv).astype(np.float32)
                                                  .astype(np.float64)).reshape(-1)
            k["mj_solimp"] = f32(mj.geom_solimp).reshape(-1)
            k["mj_solmix"] = f32(mj.geom_solmix).reshape(-1)
            k["mj_priority"] = np.ascontiguousarray(mj.geom_priority.detach().cpu().numpy().astype(np.int32))
            k["mj_condim"] = np.ascontiguousarray(mj.condim.detach().cpu().numpy().astype(np.int32))
            for name in ("mj_solref", "mj_solimp", "mj_solmix"):
                setattr(m, name, _ptr(k[name], f64p))
            for name in ("mj_priority", "mj_condim"):
                setattr(m, name, _ptr(k[name], i32p))
            m.mj_zero_margins = int(np.any(k["shape_type"] == 7))
        self.desc = m
        self.world_count = m.world_count
        self.body_count = m.body_count
        self.shape_count = m.shape_count
        self.world_body_start = k["world_body_start"]


class OrArt(C.Structure):
    _fields_ = [
        ("joint_count", C.c_int), ("world_joint_start", i32p), ("joint_type", i32p), ("joint_parent", i32p),
        ("joint_X_p", f64p), ("joint_X_c", f64p), ("joint_axis", f64p), ("joint_q_start", i32p),
        ("joint_qd_start", i32p), ("target_pos", f64p), ("target_ke", f64p), ("target_kd", f64p),
        ("armature", f64p), ("target_vel", f64p), ("joint_f", f64p), ("joint_act", f64p), ("body_f", f64p),
    ]


class OracleArt:
    """Articulation arrays of an articulated model (float32-rounded, as the device sees them)."""

    def __init__(self, model):
        h = model._host
        f = lambda k: np.ascontiguousarray(np.asarray(h[k], dtype=np.float32).astype(np.float64))
        i = lambda k: np.ascontiguousarray(np.asarray(h[k], dtype=np.int32))
        self.keep = {k: i(k) for k in ("world_joint_start", "joint_type", "joint_parent", "joint_q_start",
                                       "joint_qd_start")}
        self.keep.update({k: f(k) for k in ("joint_X_p", "joint_X_c", "joint_axis")})
        self.keep.update({"target_pos": f("joint_target_pos"), "target_ke": f("joint_target_ke"),
                          "target_kd": f("joint_target_kd"), "armature": f("joint_armature")})
        a = OrArt()
        a.joint_count = int(len(h["joint_type"]))
        for k in ("world_joint_start", "joint_type", "joint_parent", "joint_q_start", "joint_qd_start"):
            setattr(a, k, _ptr(self.keep[k], i32p))
        for k in ("joint_X_p", "joint_X_c", "joint_axis", "target_pos", "target_ke", "target_kd", "armature"):
            setattr(a, k, _ptr(self.keep[k], f64p))
        self.desc = a
        self.q_start = self.keep["joint_q_start"]
        self.qd_start = self.keep["joint_qd_start"]
        self.world_q_start = np.asarray(h["world_q_start"]).tolist() + [int(h["joint_q_start"][-1])]
        self.world_qd_start = np.asarray(h["world_qd_start"]).tolist() + [int(h["joint_qd_start"][-1])]
        self._model_target = self.keep["target_pos"]

    def set_control(self, joint_f=None, joint_target_q=None, joint_target_qd=None, joint_act=None, body_f=None):
        """Control inputs of the following calls (float32-rounded, as the device reads them):
        per-dof joint forces, position / velocity targets, feedforward, and per-body COM wrenches
        [B, 6]; None = none (position targets: the model's)."""
        r = lambda x: None if x is None else np.ascontiguousarray(
            np.asarray(x, dtype=np.float32).astype(np.float64).reshape(-1))
        self.keep["target_pos"] = r(joint_target_q) if joint_target_q is not None else self._model_target
        self.desc.target_pos = _ptr(self.keep["target_pos"], f64p)
        for field, val in (("target_vel", joint_target_qd), ("joint_f", joint_f), ("joint_act", joint_act),
                           ("body_f", body_f)):
            arr = r(val)
            self.keep["ctl_" + field] = arr
            setattr(self.desc, field, _ptr(arr, f64p) if arr is not None else None)


def art_fk(om, oa, w, q):
    """Body poses [nb, 7] of articulated world ``w`` at its coordinates ``q`` [nq]."""
    nb = int(om.world_body_start[w + 1] - om.world_body_start[w])
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros((nb, 7))
    lib().or_art_fk(C.byref(om.desc), C.byref(oa.desc), w, _ptr(q, f64p), _ptr(out, f64p))
    return out


def art_dynamics(om, oa, w, q, v):
    """Mass matrix [nv, nv] and bias force [nv] (velocity products and gravity) of world ``w``."""
    nv = len(v)
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    M = np.zeros((nv, nv)); c = np.zeros(nv)
    lib().or_art_dynamics(C.byref(om.desc), C.byref(oa.desc), w, _ptr(q, f64p), _ptr(v, f64p), _ptr(M, f64p),
                          _ptr(c, f64p))
    return M, c


def art_substep(om, oa, cfg, w, q, v, h):
    """One inner step of articulated world ``w``: (q_out, v_out, stats, ok)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    qo, vo = np.zeros_like(q), np.zeros_like(v)
    stats = np.zeros(4, dtype=np.int32)
    ok = lib().or_art_substep(C.byref(om.desc), C.byref(oa.desc), C.byref(cfg), w, _ptr(q, f64p), _ptr(v, f64p),
                              C.c_double(h), _ptr(qo, f64p), _ptr(vo, f64p), _ptr(stats, i32p))
    return qo, vo, stats, ok


def art_attempt_all(om, oa, cfg, q, v, dt):
    """Step-doubling attempt of every articulated world (all coordinates of the model)."""
    q = np.ascontiguousarray(q, dtype=np.float64); v = np.ascontiguousarray(v, dtype=np.float64)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    W = om.world_count
    qf, qb = np.zeros_like(q), np.zeros_like(q)
    vf, vb = np.zeros_like(v), np.zeros_like(v)
    err = np.zeros(W); ok = np.zeros(W, np.int32); nc = np.zeros(W, np.int32); ncm = np.zeros(W, np.int32)
    lib().or_art_attempt_all(C.byref(om.desc), C.byref(oa.desc), C.byref(cfg), _ptr(q, f64p), _ptr(v, f64p),
                             _ptr(dt, f64p), _ptr(qf, f64p), _ptr(vf, f64p), _ptr(qb, f64p), _ptr(vb, f64p),
                             _ptr(err, f64p), _ptr(ok, i32p), _ptr(nc, i32p), _ptr(ncm, i32p))
    return dict(q_full=qf, v_full=vf, q_double=qb, v_double=vb, err=err, solve_ok=ok, n_contacts=nc,
                n_contacts_mid=ncm)


def mj_convert_solref(ke, kd):
    """MuJoCo (time constant, damping ratio) of contact stiffness ke and damping kd
    (convert_solref(ke, kd, 1, 1), solvers/mujoco/kernels.py:193-207): 2/kd and kd/2 sqrt(1/ke);
    (0.02, 1) when either is not positive."""
    ke = np.asarray(ke, dtype=np.float64)
    kd = np.asarray(kd, dtype=np.float64)
    ok = (ke > 0) & (kd > 0)
    tc = np.where(ok, 2.0 / np.where(ok, kd, 1.0), 0.02)
    dr = np.where(ok, kd / 2.0 * np.sqrt(1.0 / np.where(ok, ke, 1.0)), 1.0)
    return np.stack([tc, dr], axis=-1)


def make_config(mode="adaptive", controller="sap", tol=1e-3, dt_init=0.01, dt_min=1e-6, dt_max=float("inf"),
                max_substeps=16, max_iterations=30, max_contacts=128, opt_rel_tol=None, opt_abs_tol=1e-14,
                tau_d=0.01, sigma=1e-3, beta=1.0, ls_rel_tol=1e-10, line_search="exact_root",
                contact_law="sap", impratio=1.0, warm_start="extrapolate") -> OrConfig:
    """Oracle settings.  warm_start: the Newton seeds of an attempt's three solves, "reference"
    (v_t, (v_t + v_full) / 2, v_full: solver_sap_adaptive.py:756-771) or "extrapolate" (v_t + dt/2 a
    with a the body's mean acceleration over its last accepted step, (v_t + v_full) / 2,
    2 v_mid - v_t) or "concurrent" (as "extrapolate" with the first half step from v_t + dt/4 a)."""
    c = OrConfig()
    c.mode = 1 if mode == "adaptive" else 0
    c.controller = 0 if controller == "sap" else 1
    c.tol, c.dt_init, c.dt_min, c.dt_max = tol, dt_init, dt_min, dt_max
    c.divergence_threshold = 1.0e9
    c.max_substeps, c.max_iterations, c.max_contacts = max_substeps, max_iterations, max_contacts
    c.opt_rel_tol = opt_rel_tol if opt_rel_tol is not None else 1e-10
    c.opt_abs_tol, c.tau_d, c.sigma, c.beta, c.ls_rel_tol = opt_abs_tol, tau_d, sigma, beta, ls_rel_tol
    c.line_search = {"exact_root": 0, "armijo_decay": 1}[line_search]
    c.contact_law = {"sap": 0, "mujoco": 1}[contact_law]
    c.impratio = float(impratio)
    c.warm_start = {"reference": 0, "extrapolate": 1, "concurrent": 2}[warm_start]
    return c


def mj_impedance(solimp, pos: float) -> float:
    """MuJoCo impedance d(pos) of solimp (dmin, dmax, width, midpoint, power)."""
    si = np.ascontiguousarray(solimp, dtype=np.float64)
    return float(lib().or_mj_impedance(_ptr(si, f64p), float(pos)))


def mj_contact_params(om: "OracleModel", a: int, b: int) -> dict:
    """MuJoCo contact parameters of the shape pair (a, b): solref, solimp, friction, condim."""
    sr, si, mu = np.zeros(2), np.zeros(5), np.zeros(1)
    cd = np.zeros(1, np.int32)
    lib().or_mj_contact_params(C.byref(om.desc), int(a), int(b), _ptr(sr, f64p), _ptr(si, f64p), _ptr(mu, f64p),
                               _ptr(cd, i32p))
    return {"solref": sr, "solimp": si, "mu": float(mu[0]), "condim": int(cd[0])}


def collide_world(om: OracleModel, w: int, body_q: np.ndarray, cap: int = 128):
    """Contacts of world ``w`` at global pose ``body_q`` [B, 7]: (count, n_pairs, rows[count, 14])."""
    q = np.ascontiguousarray(body_q, dtype=np.float64)
    out = np.zeros((max(cap, 1), 14), dtype=np.float64)
    npairs = C.c_int(0)
    n = lib().or_collide_world_flat(C.byref(om.desc), w, _ptr(q, f64p), _ptr(out, f64p), cap,
                                             C.byref(npairs))
    return n, npairs.value, out[:min(n, cap)]


def world_pairs(om: OracleModel, w: int, lo: np.ndarray, hi: np.ndarray, cap: int = 1 << 16) -> np.ndarray:
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    out = np.zeros((cap, 2), dtype=np.int32)
    n = lib().or_world_pairs_global(C.byref(om.desc), w, _ptr(lo, f64p), _ptr(hi, f64p), _ptr(out, i32p), cap)
    return out[:min(n, cap)].copy()


def shape_aabbs(om: OracleModel, body_q: np.ndarray):
    L = lib()
    q = np.ascontiguousarray(body_q, dtype=np.float64)
    S = om.shape_count
    lo = np.zeros((S, 3)); hi = np.zeros((S, 3))
    for s in range(S):
        L.or_shape_aabb(C.byref(om.desc), _ptr(q, f64p), 0, s, _ptr(lo[s], f64p), _ptr(hi[s], f64p))
    return lo, hi


def collide_shapes(ta, xa, sa, tb, xb, sb, gap_sum=0.0):
    out = np.zeros((8, 8))
    xa = np.ascontiguousarray(xa, dtype=np.float64); xb = np.ascontiguousarray(xb, dtype=np.float64)
    sa = np.ascontiguousarray(sa, dtype=np.float64); sb = np.ascontiguousarray(sb, dtype=np.float64)
    n = lib().or_collide_shapes(int(ta), _ptr(xa, f64p), _ptr(sa, f64p), int(tb), _ptr(xb, f64p),
                                         _ptr(sb, f64p), C.c_double(gap_sum), _ptr(out, f64p))
    return out[:n]


def cubic_min(c0, d0, c1, d1, h):
    a = C.c_double(0.0)
    ok = lib().or_cubic_min(c0, d0, c1, d1, h, C.byref(a))
    return a.value if ok else None


def coords_to_mujoco(com, joint_q, joint_qd):
    """FP64 FREE-joint conversion Newton -> MuJoCo: (qpos [B,7] wxyz, qvel [B,6])."""
    com = np.ascontiguousarray(com, dtype=np.float64).reshape(-1, 3)
    jq = np.ascontiguousarray(joint_q, dtype=np.float64).reshape(-1, 7)
    jqd = np.ascontiguousarray(joint_qd, dtype=np.float64).reshape(-1, 6)
    qpos = np.zeros_like(jq); qvel = np.zeros_like(jqd)
    lib().or_coords_to_mujoco(len(com), _ptr(com, f64p), _ptr(jq, f64p), _ptr(jqd, f64p), _ptr(qpos, f64p),
                              _ptr(qvel, f64p))
    return qpos, qvel


def coords_from_mujoco(com, qpos, qvel):
    """FP64 FREE-joint conversion MuJoCo -> Newton: (joint_q [B,7] xyzw, joint_qd [B,6])."""
    com = np.ascontiguousarray(com, dtype=np.float64).reshape(-1, 3)
    qpos = np.ascontiguousarray(qpos, dtype=np.float64).reshape(-1, 7)
    qvel = np.ascontiguousarray(qvel, dtype=np.float64).reshape(-1, 6)
    jq = np.zeros_like(qpos); jqd = np.zeros_like(qvel)
    lib().or_coords_from_mujoco(len(com), _ptr(com, f64p), _ptr(qpos, f64p), _ptr(qvel, f64p), _ptr(jq, f64p),
                                _ptr(jqd, f64p))
    return jq, jqd


SORT_KEY_SENTINEL = 0x7FFFFFFFFFFFFFFF


def contact_sort_key(shape_a, shape_b, sub):
    """``make_contact_sort_key`` (geometry/contact_data.py:59-83): shape_a in bits 62..43,
    shape_b in bits 42..23, the sub key in bits 22..0, each field masked to its width."""
    a = np.asarray(shape_a, dtype=np.int64) & 0xFFFFF
    b = np.asarray(shape_b, dtype=np.int64) & 0xFFFFF
    c = np.asarray(sub, dtype=np.int64) & 0x7FFFFF
    return (a << 43) | (b << 23) | c


def contact_order(shape0: np.ndarray, shape1: np.ndarray):
    """Deterministic contact order (ContactSorter, geometry/contact_sort.py:194): the sub key of a
    contact is its rank among the contacts of its shape pair in generated order; contacts are
    ordered by ascending key.  Returns (keys in generated order, permutation: sorted slot ->
    generated slot)."""
    s0 = np.asarray(shape0, dtype=np.int64)
    s1 = np.asarray(shape1, dtype=np.int64)
    sub = np.zeros(len(s0), dtype=np.int64)
    seen = {}
    for i in range(len(s0)):
        k = (int(s0[i]), int(s1[i]))
        sub[i] = seen.get(k, 0)
        seen[k] = sub[i] + 1
    keys = contact_sort_key(s0, s1, sub)
    return keys, np.argsort(keys, kind="stable")


def match_contacts(prev: dict, cur: dict, pos_threshold: float, normal_dot_threshold: float) -> np.ndarray:
    """Frame-to-frame matching restated from geometry/contact_match.py (module docstring and
    ContactMatcher.match): per world, a new contact takes the closest previous contact of the same
    (shape0, shape1) pair whose midpoint moved at most ``pos_threshold`` and whose normal dot is at
    least ``normal_dot_threshold``; each previous contact goes to the closest claimant (ties: the
    lower new index within the world), the others become -2; a pair without previous contacts is
    -1.  ``prev``/``cur``: dicts of world ``off``/``cnt`` arrays and per-contact ``s0``, ``s1``,
    ``mid`` [n,3], ``n`` [n,3].  Returns the match index per current contact."""
    out = np.full(len(cur["s0"]), -1, dtype=np.int64)
    for w in range(len(cur["off"])):
        po, pc = int(prev["off"][w]), int(prev["cnt"][w])
        co, cc = int(cur["off"][w]), int(cur["cnt"][w])
        best, bestd = {}, {}
        for c in range(cc):
            i = co + c
            pair, bk, bd = False, -1, 0.0
            for k in range(pc):
                j = po + k
                if prev["s0"][j] != cur["s0"][i] or prev["s1"][j] != cur["s1"][i]:
                    continue
                pair = True
                d2 = float(np.sum((cur["mid"][i] - prev["mid"][j]) ** 2))
                dn = float(np.dot(cur["n"][i], prev["n"][j]))
                if d2 <= pos_threshold ** 2 and dn >= normal_dot_threshold and (bk < 0 or d2 < bd):
                    bk, bd = k, d2
            out[i] = -2 if pair else -1
            if bk >= 0:
                best[c], bestd[c] = bk, bd
        winner = {}
        for c in sorted(best):
            k = best[c]
            if k not in winner or (bestd[c], c) < (bestd[winner[k]], winner[k]):
                winner[k] = c
        for k, c in winner.items():
            out[co + c] = po + k
    return out


def substep_world(om: OracleModel, cfg: OrConfig, w: int, q, qd, f, h, guess=None):
    """One inner step of world ``w`` (contacts at q): returns (q_out, qd_out, stats, ok) with
    stats = (newton iterations, factorizations, converged, contacts)."""
    b0, b1 = int(om.world_body_start[w]), int(om.world_body_start[w + 1])
    q = np.ascontiguousarray(q[b0:b1], dtype=np.float64)
    qd = np.ascontiguousarray(qd[b0:b1], dtype=np.float64)
    f = np.ascontiguousarray(f[b0:b1], dtype=np.float64)
    g = None if guess is None else np.ascontiguousarray(np.asarray(guess, dtype=np.float64).reshape(-1))
    qo = np.zeros_like(q); qdo = np.zeros_like(qd)
    stats = np.zeros(4, dtype=np.int32)
    ok = lib().or_substep_world(C.byref(om.desc), C.byref(cfg), w, _ptr(q, f64p), _ptr(qd, f64p), _ptr(f, f64p),
                                C.c_double(h), _ptr(g, f64p) if g is not None else None, _ptr(qo, f64p),
                                _ptr(qdo, f64p), _ptr(stats, i32p))
    return qo, qdo, stats, ok


def substep_problem(om: OracleModel, cfg: OrConfig, w: int, q, qd, f, h):
    """The convex problem of one inner step of world ``w`` (contacts at q) and the oracle's solution:
    dict(A [N,N], vstar [N], J [3 nc, N] rows t1, t2, n per contact, Rt, Rn, vhat, mu [nc], v [N],
    converged), N = 6 bodies; under the MuJoCo law also D [nc] (the rows' 1/R, 0 inactive) and
    what [nc, 3] (the contact velocity the rows drive towards), mu 0 for condim 1."""
    b0, b1 = int(om.world_body_start[w]), int(om.world_body_start[w + 1])
    q = np.ascontiguousarray(q[b0:b1], dtype=np.float64)
    qd = np.ascontiguousarray(qd[b0:b1], dtype=np.float64)
    f = np.ascontiguousarray(f[b0:b1], dtype=np.float64)
    N, ncap = 6 * (b1 - b0), max(int(cfg.max_contacts), 1)
    A = np.zeros((N, N)); vs = np.zeros(N); J = np.zeros((3 * ncap, N)); v = np.zeros(N)
    Rt, Rn, vh, mu, D = (np.zeros(ncap) for _ in range(5))
    what = np.zeros((ncap, 3))
    conv = C.c_int(0)
    L = lib()
    L.or_substep_problem.restype = C.c_int
    nc = L.or_substep_problem(C.byref(om.desc), C.byref(cfg), w, _ptr(q, f64p), _ptr(qd, f64p), _ptr(f, f64p),
                              C.c_double(h), N, ncap, _ptr(A, f64p), _ptr(vs, f64p), _ptr(J, f64p), _ptr(Rt, f64p),
                              _ptr(Rn, f64p), _ptr(vh, f64p), _ptr(mu, f64p), _ptr(v, f64p), C.byref(conv),
                              _ptr(D, f64p), _ptr(what, f64p))
    assert nc >= 0
    return dict(A=A, vstar=vs, J=J[:3 * nc], Rt=Rt[:nc], Rn=Rn[:nc], vhat=vh[:nc], mu=mu[:nc], v=v,
                converged=bool(conv.value), D=D[:nc], what=what[:nc])


def _trend(trend, B):
    return None if trend is None else np.ascontiguousarray(np.asarray(trend, dtype=np.float64).reshape(B, 6))


def attempt_all(om: OracleModel, cfg: OrConfig, q, qd, f, dt, trend=None):
    """One step-doubling attempt of every world (no commit).  trend [B, 6] (None = zero): the
    bodies' accelerations that seed the full step under warm_start "extrapolate"."""
    W, B = om.world_count, om.body_count
    q = np.ascontiguousarray(q, dtype=np.float64); qd = np.ascontiguousarray(qd, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64); dt = np.ascontiguousarray(dt, dtype=np.float64)
    tr = _trend(trend, B)
    qf = np.zeros((B, 7)); qdf = np.zeros((B, 6)); qb = np.zeros((B, 7)); qdb = np.zeros((B, 6))
    err = np.zeros(W); ok = np.zeros(W, np.int32); nc = np.zeros(W, np.int32); ncm = np.zeros(W, np.int32)
    lib().or_attempt_all(C.byref(om.desc), C.byref(cfg), _ptr(q, f64p), _ptr(qd, f64p), _ptr(f, f64p),
                         _ptr(tr, f64p) if tr is not None else None, _ptr(dt, f64p), _ptr(qf, f64p),
                         _ptr(qdf, f64p), _ptr(qb, f64p), _ptr(qdb, f64p), _ptr(err, f64p), _ptr(ok, i32p),
                         _ptr(nc, i32p), _ptr(ncm, i32p))
    return dict(q_full=qf, qd_full=qdf, q_double=qb, qd_double=qdb, err=err, solve_ok=ok, n_contacts=nc,
                n_contacts_mid=ncm)


def _records(cfg, contacts, counts):
    """Device contact records [W, 2, 128, 16] and counts [W, 2] laid out for cfg.max_contacts."""
    cap = max(int(cfg.max_contacts), 1)
    contacts = np.asarray(contacts)
    W = contacts.shape[0]
    n = np.minimum(np.asarray(counts, dtype=np.int32).reshape(W, 2), cap)
    rec = np.zeros((W, 2, cap, 16))
    m = min(cap, contacts.shape[2])
    rec[:, :, :m] = contacts[:, :, :m]
    return np.ascontiguousarray(rec), np.ascontiguousarray(n)


def attempt_all_given(om: OracleModel, cfg: OrConfig, q, qd, f, dt, contacts, counts, trend=None):
    """Step-doubling attempt of every world that solves with the given contact sets (a device
    attempt's records [W, 2, 128, 16] at q and at its midpoint, counts [W, 2]) instead of colliding;
    the layout and the trend of :func:`attempt_all`."""
    W, B = om.world_count, om.body_count
    q = np.ascontiguousarray(q, dtype=np.float64); qd = np.ascontiguousarray(qd, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64); dt = np.ascontiguousarray(dt, dtype=np.float64)
    tr = _trend(trend, B)
    rec, n = _records(cfg, contacts, counts)
    qf = np.zeros((B, 7)); qdf = np.zeros((B, 6)); qb = np.zeros((B, 7)); qdb = np.zeros((B, 6))
    err = np.zeros(W); ok = np.zeros(W, np.int32)
    lib().or_attempt_all_given(C.byref(om.desc), C.byref(cfg), _ptr(q, f64p), _ptr(qd, f64p), _ptr(f, f64p),
                               _ptr(tr, f64p) if tr is not None else None, _ptr(dt, f64p), _ptr(rec, f64p),
                               _ptr(n, i32p), _ptr(qf, f64p), _ptr(qdf, f64p), _ptr(qb, f64p), _ptr(qdb, f64p),
                               _ptr(err, f64p), _ptr(ok, i32p))
    return dict(q_full=qf, qd_full=qdf, q_double=qb, qd_double=qdb, err=err, solve_ok=ok)


def art_attempt_all_given(om: OracleModel, oa: "OracleArt", cfg: OrConfig, q, v, dt, contacts, counts):
    """Articulated counterpart of :func:`attempt_all_given` (the layout of :func:`art_attempt_all`)."""
    q = np.ascontiguousarray(q, dtype=np.float64); v = np.ascontiguousarray(v, dtype=np.float64)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    W = om.world_count
    rec, n = _records(cfg, contacts, counts)
    qf, qb = np.zeros_like(q), np.zeros_like(q)
    vf, vb = np.zeros_like(v), np.zeros_like(v)
    err = np.zeros(W); ok = np.zeros(W, np.int32)
    lib().or_art_attempt_all_given(C.byref(om.desc), C.byref(oa.desc), C.byref(cfg), _ptr(q, f64p), _ptr(v, f64p),
                                   _ptr(dt, f64p), _ptr(rec, f64p), _ptr(n, i32p), _ptr(qf, f64p), _ptr(vf, f64p),
                                   _ptr(qb, f64p), _ptr(vb, f64p), _ptr(err, f64p), _ptr(ok, i32p))
    return dict(q_full=qf, v_full=vf, q_double=qb, v_double=vb, err=err, solve_ok=ok)


class OracleSolver:
    """Frame-level driver of the oracle with the reference solver's controller state."""

    def __init__(self, om: OracleModel, cfg: OrConfig):
        self.om, self.cfg = om, cfg
        W = om.world_count
        self.a = {
            "dt": np.full(W, cfg.dt_init), "dt_half": np.full(W, 0.5 * cfg.dt_init),
            "ideal_dt": np.full(W, cfg.dt_init), "sim_time": np.zeros(W), "next_time": np.zeros(W),
            "last_error": np.zeros(W), "accepted_error": np.zeros(W),
            "accepted": np.zeros(W, np.uint8), "commit": np.zeros(W, np.uint8), "diverged": np.zeros(W, np.uint8),
            "substeps": np.zeros(W, np.int32), "iterations": np.zeros(W, np.int32),
            "solve_fail": np.zeros(W, np.int32), "cum_accepted": np.zeros(1, np.int64),
            "trend": np.zeros((max(om.body_count, 1), 6)),
        }
        k = OrCtrl()
        for name, arr in self.a.items():
            t = {np.float64: f64p, np.uint8: u8p, np.int32: i32p, np.int64: i64p}[arr.dtype.type]
            setattr(k, name, _ptr(arr, t))
        self.ctrl = k

    def step(self, q: np.ndarray, qd: np.ndarray, f: np.ndarray, dt_outer: float) -> int:
        assert q.dtype == np.float64 and q.flags.c_contiguous and qd.dtype == np.float64 and qd.flags.c_contiguous
        f = np.ascontiguousarray(f, dtype=np.float64)
        return lib().or_step_frame(C.byref(self.om.desc), C.byref(self.cfg), _ptr(q, f64p),
                                                 _ptr(qd, f64p), _ptr(f, f64p), C.byref(self.ctrl),
                                                 C.c_double(dt_outer))

    def step_articulated(self, oa: "OracleArt", q: np.ndarray, v: np.ndarray, dt_outer: float) -> int:
        """Boundary call of an articulated model on joint coordinates q / velocities v (in place)."""
        assert q.dtype == np.float64 and q.flags.c_contiguous and v.dtype == np.float64 and v.flags.c_contiguous
        return lib().or_art_step_frame(C.byref(self.om.desc), C.byref(oa.desc), C.byref(self.cfg), _ptr(q, f64p),
                                       _ptr(v, f64p), C.byref(self.ctrl), C.c_double(dt_outer))




def _qrot(q, v, inverse=False):
    """Rotate v by the unit quaternion q (x, y, z, w); inverse rotates by its conjugate."""
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return (R.T if inverse else R) @ np.asarray(v, dtype=np.float64)


def coords_to_mujoco_joints(joint_type, q_start, qd_start, X_p, com, joint_q, joint_qd):
    """Joint-wise Newton -> MuJoCo conversion of an articulated model (convert_warp_coords_to_mj_kernel,
    solvers/mujoco/kernels.py:705-789; one joint per body, joint j moves body j): a FREE joint
    rotates its parent-frame twist into the world by joint_X_p's rotation and then converts as a
    free body; a revolute joint copies (no reference offset).  Returns flat (qpos, qvel)."""
    qpos = np.array(joint_q, dtype=np.float64)
    qvel = np.array(joint_qd, dtype=np.float64)
    for j in range(len(joint_type)):
        if int(joint_type[j]) != 4:
            continue
        q0, d0 = int(q_start[j]), int(qd_start[j])
        qp = np.asarray(X_p[j][3:7], dtype=np.float64)
        u = np.concatenate([_qrot(qp, joint_qd[d0:d0 + 3]), _qrot(qp, joint_qd[d0 + 3:d0 + 6])])
        p, v = coords_to_mujoco(np.asarray(com[j]).reshape(1, 3), np.asarray(joint_q[q0:q0 + 7]).reshape(1, 7),
                                u.reshape(1, 6))
        qpos[q0:q0 + 7], qvel[d0:d0 + 6] = p[0], v[0]
    return qpos, qvel
