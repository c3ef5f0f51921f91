"""GPU controller kernels (C ABI ``na_ctrl_*``) against the oracle, bit-exact in float32.

Inputs: the reference's known-answer cases (``test_adaptive_floor_nan_guard.py``,
``test_adaptive_time_rebase.py``, ``test_adaptive_reset_worlds.py``) plus seeded random
inputs that cover every branch of ``_calc_adjusted_step`` / ``_adapt_dt``.
"""

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV)


def _random_ctrl_inputs(n, seed):
    rng = np.random.default_rng(seed)
    err = (10.0 ** rng.uniform(-9, 1, n)).astype(np.float32)
    err[rng.random(n) < 0.1] = np.float32(1e10)
    err[rng.random(n) < 0.05] = np.float32(np.nan)
    err[rng.random(n) < 0.05] = np.float32(np.inf)
    dt = (10.0 ** rng.uniform(-6.2, -2, n)).astype(np.float32)
    dt[rng.random(n) < 0.05] = 0.0
    dt[rng.random(n) < 0.1] = np.float32(1e-6)
    return err, dt


@pytest.mark.parametrize("seed", [0, 1])
def test_calc_adjusted_step_bit_exact(nalib, O, seed):
    n = 4096
    err, dt = _random_ctrl_inputs(n, seed)
    err[:4] = [1e10, 1e-2, 5e-4, 1e10]
    dt[:4] = [1e-6, 1e-6, 1e-5, 1e-5]
    ideal0 = np.full(n, 0.0042, np.float32)
    # oracle
    oi = ideal0.copy(); oa = np.zeros(n, np.uint8); oc = np.zeros(n, np.uint8); od = np.zeros(n, np.uint8)
    O.lib().or_calc_adjusted_step_f32(n, _p(err, C.c_float), _p(dt, C.c_float), _p(oi, C.c_float), _p(oa, C.c_uint8),
                                      _p(oc, C.c_uint8), _p(od, C.c_uint8), C.c_float(1e-3), C.c_float(1e-6),
                                      C.c_float(1e9))
    # gpu
    ge, gd, gi = _t(err), _t(dt), _t(ideal0)
    ga = torch.zeros(n, dtype=torch.uint8, device=DEV); gc = torch.zeros_like(ga); gv = torch.zeros_like(ga)
    nalib.check(nalib.lib().na_ctrl_calc_adjusted_step_f32(n, ge.data_ptr(), gd.data_ptr(), gi.data_ptr(),
                                                           ga.data_ptr(), gc.data_ptr(), gv.data_ptr(), 1e-3, 1e-6,
                                                           1e9, None), "ctrl")
    torch.cuda.synchronize()
    assert np.array_equal(ga.cpu().numpy(), oa) and np.array_equal(gc.cpu().numpy(), oc)
    assert np.array_equal(gv.cpu().numpy(), od)
    assert np.array_equal(gi.cpu().numpy().view(np.uint32), oi.view(np.uint32))
    # the reference KATs (floor diverged held; floor finite committed; normal; above-floor reject)
    assert list(oa[:4]) == [1, 1, 1, 0] and list(oc[:4]) == [0, 1, 1, 0] and list(od[:4]) == [1, 0, 0, 0]


@pytest.mark.parametrize("mode", [0, 1])
def test_adapt_dt_bit_exact(nalib, O, mode):
    n = 4096
    err, dt = _random_ctrl_inputs(n, 10 + mode)
    rng = np.random.default_rng(20 + mode)
    nxt = np.full(n, np.float32(1.0 / 60.0))
    sim = (rng.random(n) * 0.02).astype(np.float32)
    arrs = dict(sim=sim, dt=dt, dth=(dt * np.float32(0.5)).astype(np.float32), ideal=dt.copy(),
                div=np.zeros(n, np.uint8), acc=np.zeros(n, np.uint8), ae=np.zeros(n, np.float32),
                sub=np.zeros(n, np.int32), cum=np.zeros(1, np.int32))
    o = {k: v.copy() for k, v in arrs.items()}
    O.lib().or_adapt_dt_f32(n, _p(err, C.c_float), _p(o["sim"], C.c_float), _p(nxt, C.c_float), _p(o["dt"], C.c_float),
                            _p(o["dth"], C.c_float), _p(o["ideal"], C.c_float), _p(o["div"], C.c_uint8),
                            _p(o["acc"], C.c_uint8), _p(o["ae"], C.c_float), _p(o["sub"], C.c_int32),
                            _p(o["cum"], C.c_int32), mode, C.c_float(1e-3), C.c_float(1e-6), C.c_float(0.01),
                            C.c_float(1e9))
    g = {k: _t(v.copy()) for k, v in arrs.items()}
    ge, gn = _t(err), _t(nxt)  # named: inputs must outlive the asynchronous launch
    nalib.check(nalib.lib().na_ctrl_adapt_dt_f32(n, ge.data_ptr(), g["sim"].data_ptr(), gn.data_ptr(),
                                                 g["dt"].data_ptr(), g["dth"].data_ptr(), g["ideal"].data_ptr(),
                                                 g["div"].data_ptr(), g["acc"].data_ptr(), g["ae"].data_ptr(),
                                                 g["sub"].data_ptr(), g["cum"].data_ptr(), mode, 1e-3, 1e-6, 0.01,
                                                 1e9, None), "adapt")
    torch.cuda.synchronize()
    for k in arrs:
        a, b = g[k].cpu().numpy(), o[k]
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k


def test_rebase_reset_clamp_infnorm_bit_exact(nalib, O):
    lib = nalib.lib()
    # rebase (Fix B cases)
    sim = np.array([1.0e6, 5.0 + 1.0e-4, np.float32(1048576.0) - np.float32(1e-3), 0.0, 10.0 + 2e-6], np.float32)
    nxt = np.array([1.0e6, 5.0, 1048576.0, 0.0, 10.0], np.float32)
    gs, gn = _t(sim), _t(nxt)
    nalib.check(lib.na_ctrl_rebase_time_f32(len(sim), gs.data_ptr(), gn.data_ptr(), None), "rebase")
    torch.cuda.synchronize()
    os_, on = sim.copy(), nxt.copy()
    O.lib().or_rebase_time_f32(len(sim), _p(os_, C.c_float), _p(on, C.c_float))
    torch.cuda.synchronize()
    assert np.array_equal(gs.cpu().numpy(), os_) and np.array_equal(gn.cpu().numpy(), on)
    assert np.all(on == 0.0)
    assert on[2] - os_[2] == np.float32(nxt[2]) - np.float32(sim[2])
    # masked reset (Fix C)
    n = 3
    mask = np.array([1, 0, 1], np.uint8)
    vals = {k: np.array([0.005, 0.007, 0.009], np.float32) for k in ("ideal", "dt", "dth", "sim", "nxt")}
    flags = {k: np.ones(n, np.uint8) for k in ("div", "acc")}
    g = {k: _t(v.copy()) for k, v in {**vals, **flags}.items()}
    gm = _t(mask)
    nalib.check(lib.na_ctrl_reset_worlds_f32(n, gm.data_ptr(), 0.01, g["ideal"].data_ptr(), g["dt"].data_ptr(),
                                             g["dth"].data_ptr(), g["sim"].data_ptr(), g["nxt"].data_ptr(),
                                             g["div"].data_ptr(), g["acc"].data_ptr(), None), "reset")
    torch.cuda.synchronize()
    assert g["dt"].cpu().numpy()[0] == np.float32(0.01) and g["dt"].cpu().numpy()[1] == np.float32(0.007)
    assert g["dth"].cpu().numpy()[2] == np.float32(0.01) * np.float32(0.5)
    assert list(g["div"].cpu().numpy()) == [0, 1, 0]
    # clamp to boundary
    dt = np.array([1e-3, 1e-3, 1e-3], np.float32)
    sim = np.array([0.0, 0.0166, 0.02], np.float32)
    nxt = np.full(3, np.float32(1 / 60), np.float32)
    gd, gh, gs, gn = _t(dt), _t(dt * np.float32(0.5)), _t(sim), _t(nxt)
    nalib.check(lib.na_ctrl_clamp_dt_to_boundary_f32(3, gd.data_ptr(), gh.data_ptr(), gs.data_ptr(), gn.data_ptr(),
                                                     None), "clamp")
    od, oh = dt.copy(), dt * np.float32(0.5)
    O.lib().or_clamp_dt_to_boundary_f32(3, _p(od, C.c_float), _p(oh, C.c_float), _p(sim, C.c_float),
                                        _p(nxt, C.c_float))
    torch.cuda.synchronize()
    assert np.array_equal(gd.cpu().numpy(), od) and np.array_equal(gh.cpu().numpy(), oh)
    assert od[2] == 0.0 and od[1] == np.float32(nxt[1]) - np.float32(sim[1])
    # inf-norm error with NaN / inf sentinel
    rng = np.random.default_rng(4)
    qf = rng.normal(size=(64, 7)).astype(np.float32)
    qd = (qf + rng.normal(size=(64, 7)) * 1e-4).astype(np.float32)
    qd[5, 3] = np.nan
    qd[9, 0] = np.inf
    sc = np.ones((64, 7), np.float32)
    ge = torch.zeros(64, dtype=torch.float32, device=DEV)
    gqf, gqd, gsc = _t(qf), _t(qd), _t(sc)  # keep the inputs alive until the kernel has run
    nalib.check(lib.na_ctrl_inf_norm_error_f32(64, 7, gqf.data_ptr(), gqd.data_ptr(), gsc.data_ptr(),
                                               ge.data_ptr(), None), "infnorm")
    oe = np.zeros(64, np.float32)
    O.lib().or_inf_norm_error_f32(64, 7, _p(qf, C.c_float), _p(qd, C.c_float), _p(sc, C.c_float), _p(oe, C.c_float))
    torch.cuda.synchronize()
    assert np.array_equal(ge.cpu().numpy(), oe) and oe[5] == np.float32(1e10) and oe[9] == np.float32(1e10)
