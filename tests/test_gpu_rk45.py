"""The embedded Runge-Kutta 4(5) with step-size control of the reference's
legacy solver (dgr_rk45_step, rkBSSN.cpp:861-1085) on the GPU against the
oracle's orc_rk45_step (pinned by the closed form of the linear gauge mode and
by its error estimate and step control): uniform 13^3 blocks and uneven
Brill-Lindquist blocks with frozen padding, accepted at once and after
retries, and the two-level lattice with the full padding refresh, in FP64.
Interior increments within 1e-11 normwise per variable, or within the
oracle's own sensitivity where that is larger: on the Brill-Lindquist blocks
the oracle's step moves by ~1.7e-9 normwise when its input is perturbed by
2e-16 relative (grid-scale rounding noise is amplified by the right-hand
side's second derivatives and KO terms near the punctures, through six stages
whose tableau reaches coefficients of 8), and the GPU's 2e-15 right-hand-side
rounding differences propagate the same way (measured 4.3e-10; 5e-13 at a
quarter of the step). The same number of tries and step sizes, the error
estimate within 1e-7 (a difference of two solutions that agree to ~1e-6, so
its last digits are rounding)."""
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


def _compare(got, ref, st, t, rg, rr, bound=1e-11):
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  attempts {rg['attempts']}/{rr['attempts']}, h_used {rg['h_used']:.9e}/{rr['h_used']:.9e}, "
        f"err {rg['err_norm']:.9e}/{rr['err_norm']:.9e}, interior increments normwise {e:.3e}")
    assert rg["accepted"] == rr["accepted"]
    assert rg["attempts"] == rr["attempts"]
    assert rg["h_used"] == pytest.approx(rr["h_used"], rel=1e-9)
    assert rg["err_norm"] == pytest.approx(rr["err_norm"], rel=1e-7)
    assert rg["h_next"] == pytest.approx(rr["h_next"], rel=1e-7)
    assert e <= bound


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("retry", [False, True])
def test_rk45_step_matches_oracle(case, retry):
    import torch
    log(f"rk45: {case}, {'with retries' if retry else 'accepted at once'}")
    t, st, pkw, dt = CASES[case]()
    if dt is None:
        dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    p = dgr.default_params(**pkw)
    tol = 1.0
    if retry:  # a tolerance 20x below the first try's estimate
        probe = st.copy()
        tol = O.rk45_step(probe, t, dt, 1.0, oracle_params(p), nthreads=8)["err_norm"] / 20
    ref = st.copy()
    rr = O.rk45_step(ref, t, dt, tol, oracle_params(p), nthreads=8)
    # the oracle's own sensitivity: the same step from the state perturbed by 2e-16 relative
    pert = st * (1 + 2e-16 * np.random.default_rng(1).standard_normal(st.shape))
    rp = pert.copy()
    O.rk45_step(rp, t, dt, tol, oracle_params(p), nthreads=8)
    sens = normwise(interior_values(rp - pert, t), interior_values(ref - st, t)).max()
    log(f"  oracle sensitivity to a 2e-16 input perturbation {sens:.3e}")
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    rg = plan.rk45_step(u, dt, tol, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    _compare(got, ref, st, t, rg, rr, bound=max(1e-11, sens))
    assert rg["accepted"] and (rg["attempts"] > 1) == retry
    inner = O.interior_mask(t, st.shape[1])
    assert np.array_equal(got[:, ~inner], st[:, ~inner])  # padding untouched
    plan.close()


def test_rk45_with_full_refresh_matches_oracle():
    import torch
    log("rk45: two-level lattice, full refresh")
    t = W.two_level_lattice()
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_halo_refresh("full")
    u = to_dev(st)
    rg = plan.rk45_step(u, 0.02, 1e-3, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    rr = O.rk45_step(ref, t, 0.02, 1e-3, oracle_params(p), nbr=W.neighbors(t), general=True, nthreads=8)
    _compare(got, ref, st, t, rg, rr)
    plan.close()


def test_rk45_reports_a_step_it_could_not_accept():
    """With max_attempts tries all above the tolerance the call returns the
    last try (not accepted) with the retry step size as the next one."""
    import torch
    t, st, pkw, dt = CASES["uniform"]()
    p = dgr.default_params(**pkw)
    # a tolerance 1e4 below the first try's estimate: the second try, at
    # 0.8 (1e-4)^(1/4) = 0.08 of the step, is still ~0.08 of it above
    tol = O.rk45_step(st.copy(), t, dt, 1.0, oracle_params(p), nthreads=8)["err_norm"] / 1e4
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    rg = plan.rk45_step(u, dt, tol, p, max_attempts=2)
    torch.cuda.synchronize()
    ref = st.copy()
    rr = O.rk45_step(ref, t, dt, tol, oracle_params(p), max_attempts=2, nthreads=8)
    assert not rg["accepted"] and not rr["accepted"] and rg["attempts"] == rr["attempts"] == 2
    _compare(u.cpu().numpy(), ref, st, t, rg, rr)
    plan.close()
