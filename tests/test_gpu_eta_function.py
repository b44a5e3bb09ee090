"""The reference's BSSN_ETA_FUNCTION build (CMakeLists.txt:95-97, equations
bssneqs_eta_func_standard_gauge.cpp): B's damping eta = R0 |d chi| /
(1 - chi^p0)^p1 from chi's gradient in the conformal metric, selected per call
with dgr_bssn_params.eta_func. The oracle's eta function is pinned to the
reference's own generated equations (tests/test_oracle_golden.py); here the
GPU right-hand side (fused kernel, its own instantiations formed in chi's
phase; and the split path) matches the oracle to 1e-12 normwise per variable
on Brill-Lindquist blocks (chi < 1, so the eta is finite), with faces flagged,
and an RK4 step to 1e-11 or the oracle's own sensitivity to a 2e-16 input
perturbation where larger; mixed precision refuses it."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import gpu_rhs, host_state, interior_values, log, normwise, oracle_params, to_dev
from test_gpu_mixed import uneven_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def _blocks(case):
    if case == "uneven":
        return uneven_sample()
    t = W.single_block_19(63)  # all faces flagged, near the origin: chi well below 1
    return t


@pytest.mark.parametrize("kernel", ["fused", "split"])
@pytest.mark.parametrize("case", ["uneven", "faces"])
def test_eta_function_rhs_matches_oracle(monkeypatch, kernel, case):
    monkeypatch.setenv("DGR_KERNEL", kernel)
    t = _blocks(case)
    st = host_state(t, init="puncture", seed=5)
    p = dgr.default_params(ko_sigma=0.4, eta_func=1)
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    e = normwise(interior_values(got, t), interior_values(ref, t))
    # the RIT eta gives other B right-hand sides: the switch is live
    p0 = dgr.default_params(ko_sigma=0.4)
    ref0 = O.bssnrhs_batch(st, t, oracle_params(p0), nthreads=8)
    d = normwise(interior_values(ref0, t), interior_values(ref, t))
    log(f"eta function {kernel} {case}: normwise max {e.max():.3e}; B against the RIT eta {d[9:12].min():.2e}")
    assert np.isfinite(interior_values(got, t)).all()
    assert e.max() <= 1e-12
    assert d[9:12].min() > 1e-6  # measured 4e-5 (B is 1e-6 noise there: the Gt terms dominate B's right-hand side)


def test_eta_function_rk4_step_matches_oracle():
    import torch
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    p = dgr.default_params(ko_sigma=0.4, eta_func=1, eta_r0=0.9, eta_power=(2.0, 1.0))
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_steps(u, dt, p, nsteps=1)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    O.rk4_step(ref, t, dt, oracle_params(p), enforce_per_stage=True, nthreads=8)
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    # the oracle's own sensitivity: the same step from the state perturbed by 2e-16
    # relative (near the punctures grid-scale rounding noise is amplified by the
    # second derivatives and KO terms; tests/test_gpu_rk45.py measures the same)
    pert = st * (1 + 2e-16 * np.random.default_rng(1).standard_normal(st.shape))
    rp = pert.copy()
    O.rk4_step(rp, t, dt, oracle_params(p), enforce_per_stage=True, nthreads=8)
    sens = normwise(interior_values(rp - pert, t), interior_values(ref - st, t)).max()
    log(f"eta function rk4: interior increments normwise {e:.3e} (oracle sensitivity {sens:.3e})")
    assert e <= max(1e-11, sens)
    plan.set_precision("mixed")
    with pytest.raises(RuntimeError, match="eta function"):
        plan.rk4_step(to_dev(st), dt, p)
    plan.close()
