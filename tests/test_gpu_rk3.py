"""The three-stage SSP Runge-Kutta of the reference's legacy solver
(dgr_rk3_step, rkBSSN.cpp:691-762) on the GPU against the oracle's
orc_rk3_step (pinned by the closed form of the linear gauge mode): uniform
13^3 blocks and uneven Brill-Lindquist blocks with frozen padding, and the
two-level lattice with the full padding refresh before every stage, in FP64;
interior increments within 1e-11 normwise per variable, padding untouched
without the refresh. Mixed precision within its error."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import host_state, interior_values, log, normwise, oracle_params, to_dev
from test_gpu_mixed import uneven_sample
from test_oracle_halo import smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


CASES = {
    "uniform": lambda: (W.uniform_13(64), host_state(W.uniform_13(64), seed=3, amp=1e-2),
                        dict(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0), 0.025),
    "puncture": lambda: (uneven_sample(), host_state(uneven_sample(), init="puncture", seed=5),
                         dict(ko_sigma=0.4), None),
}


@pytest.mark.parametrize("case", list(CASES))
def test_rk3_step_matches_oracle(case):
    import torch
    log(f"rk3: {case}")
    t, st, pkw, dt = CASES[case]()
    if dt is None:
        dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    p = dgr.default_params(**pkw)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk3_step(u, dt, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    O.rk3_step(ref, t, dt, oracle_params(p), nthreads=8)
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    inner = O.interior_mask(t, st.shape[1])
    assert np.array_equal(got[:, ~inner], st[:, ~inner])  # padding untouched
    # mixed precision: the same step within its error against the state's scale
    plan.set_precision("mixed")
    um = to_dev(st)
    plan.rk3_step(um, dt, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    gm = um.cpu().numpy()
    es = np.abs(gm - got).max() / np.abs(got).max()
    log(f"  mixed: state scale {es:.3e}")
    assert es < {"uniform": 3e-7, "puncture": 5e-10}[case]  # measured 1.0e-7 (amplitude 1e-2), 1.5e-10
    plan.close()


def test_rk3_with_full_refresh_matches_oracle():
    import torch
    log("rk3: two-level lattice, full refresh")
    t = W.two_level_lattice()
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_halo_refresh("full")
    u = to_dev(st)
    for _ in range(2):
        plan.rk3_step(u, 0.02, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    for _ in range(2):
        O.rk3_step(ref, t, 0.02, oracle_params(p), nbr=W.neighbors(t), general=True, nthreads=8)
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref - st).max()
    plan.close()
