"""The schedule autotuner of the boundary call: every candidate schedule (waves per CU x world
order, and two-wave tiles under the concurrent seeds) leaves the same state bit for bit, the tuner
restores the state it started from, and the boundary calls after tuning equal those of an untuned
solver from the same state."""

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("seeds,n_candidates", [("extrapolate", 9), ("concurrent", 12)])
def test_autotune_restores_state_and_keeps_results(na, seeds, n_candidates):
    model = na.scenes.uneven_pile(256, seed=11, device=DEV)
    kw = dict(tol=1e-3, dt_inner_init=1e-3, dt_inner_max=0.01, max_substeps=256, warm_start=seeds)
    a, b = model.state(), model.state()
    sa, sb = na.SolverSAPAdaptive(model, **kw), na.SolverSAPAdaptive(model, **kw)
    for _ in range(5):
        sa.step(a, a, model.control(), None, 0.01)
        sb.step(b, b, model.control(), None, 0.01)
    torch.cuda.synchronize()
    q_before, dt_before = a.body_q.clone(), sa._dt.clone()
    table = sa.autotune(a, 0.01, frames=2)
    torch.cuda.synchronize()
    print(table["best"], [(r["waves_per_cu"], r["order"], r["tiles"], round(r["ms"], 2)) for r in table["candidates"]],
          flush=True)
    assert len(table["candidates"]) == n_candidates
    assert torch.equal(a.body_q, q_before) and torch.equal(sa._dt, dt_before)
    for _ in range(3):
        sa.step(a, a, model.control(), None, 0.01)
        sb.step(b, b, model.control(), None, 0.01)
    torch.cuda.synchronize()
    assert torch.equal(a.body_q, b.body_q) and torch.equal(a.body_qd, b.body_qd)
