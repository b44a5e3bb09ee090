"""RK4 steps replayed from a captured HIP graph (dgr_rk4_steps) give the same
bits as the same steps launched eagerly (dgr_rk4_step), replay after replay,
with unrelated eager GPU work between replays, and after a re-capture."""
import numpy as np
import pytest

import dgr
from dgr import workloads as W
from gpu_util import host_state, log, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def test_graph_steps_match_eager_steps_bitwise():
    import torch
    log("rk4 graph vs eager")
    t = W.uniform_13(24)
    t["bflag"][5] = 63  # all three fused launches take part
    st = host_state(t, seed=21)
    p = dgr.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
    plan_e = dgr.Plan(t, st.shape[1])
    plan_g = dgr.Plan(t, st.shape[1])
    ue, ug = to_dev(st), to_dev(st)

    for _ in range(3):
        plan_e.rk4_step(ue, 0.025, p)
    plan_g.rk4_steps(ug, 0.025, p, nsteps=3)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)

    # replays one at a time, with eager work of the program in between
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    for i in range(3):
        plan_g.rk4_steps(ug, 0.025, p, nsteps=1)
        scratch.add_(ug[0, : 1 << 20] if ug.shape[1] >= (1 << 20) else 1.0)
        plan_e.rk4_step(ue, 0.025, p)
        torch.cuda.synchronize()
        assert torch.equal(ue, ug), f"replay {i}"

    # different dt: the graph is captured again
    plan_g.rk4_steps(ug, 0.0125, p, nsteps=2)
    for _ in range(2):
        plan_e.rk4_step(ue, 0.0125, p)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)
    assert plan_g.status() == 0 and plan_e.status() == 0
    plan_e.close()
    plan_g.close()


def test_graph_steps_with_timing_enabled_run_eagerly():
    import torch
    t = W.uniform_13(4)
    st = host_state(t, seed=22)
    p = dgr.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
    plan = dgr.Plan(t, st.shape[1])
    ua, ub = to_dev(st), to_dev(st)
    plan.set_timing(True)
    plan.rk4_steps(ua, 0.025, p, nsteps=2)
    kt = plan.kernel_times()
    plan.set_timing(False)
    assert kt["fused_wide"][1] == 8 and kt["fused_thin"][1] == 8
    plan.rk4_steps(ub, 0.025, p, nsteps=2)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub)
    plan.close()
