"""The reference's other gauge builds (rhs.cpp:219-265), selected per call:
the Rochester puncture gauge (USE_ROCHESTER_GAUGE, CodeGen/bssn.py:415-455;
dgr_bssn_params.rochester with BSSN_XI in xi) and the slow-start lapse
(BSSN_ENABLE_SSL_IN_RHS, CodeGen/bssn.py:90-99; params.ssl at the time t of
dgr_bssn_rhs, params.time in the RK steps), alone and together with the eta
function. The oracle's variants are pinned to the reference's own generated
equations (tests/test_oracle_golden.py); here the GPU right-hand side (the fused
kernel's gauge instantiations and the split path) matches the oracle to 1e-12
normwise per variable on Brill-Lindquist blocks, with faces flagged, and RK4
steps (the captured graph, and the eager steps the slow-start lapse takes, one
time per step) to 1e-11 or the oracle's own sensitivity to a 2e-16 input
perturbation where larger; mixed precision refuses them. The lapse starts at
1 (Brill-Lindquist's pre-collapsed lapse equals sqrt(chi), where the slow-start
term vanishes)."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import gpu_rhs, host_state, interior_values, log, normwise, oracle_params, to_dev
from test_gpu_mixed import uneven_sample

pytestmark = pytest.mark.gpu

# name -> (eta_func, rochester, ssl, time)
VARIANTS = {"rochester": (0, 1, 0, 0.0), "ssl": (0, 0, 1, 7.5), "rochester_ssl_etafunc": (1, 1, 1, 31.0)}


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def _params(name, **over):
    ef, ro, ssl, t = VARIANTS[name]
    kw = dict(ko_sigma=0.4, eta_func=ef, eta_r0=0.9, eta_power=(2.0, 1.0), rochester=ro, ssl=ssl,
              xi=(2.0, 3.0, 5.0), time=t)
    kw.update(over)
    return dgr.default_params(**kw)


def _state(t):
    """Brill-Lindquist with the lapse at 1 (the start the slow-start lapse is
    for: it drives alpha towards sqrt(chi), which the pre-collapsed lapse
    alpha = psi^-2 already is)."""
    st = host_state(t, init="puncture", seed=5)
    st[0] = 1.0
    return st


def _gpu_rhs_at(t, st, p):
    import torch
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    out = torch.full_like(u, np.nan)
    plan.rhs(out, u, p, t=p.time)
    torch.cuda.synchronize()
    assert plan.status() == 0
    res = out.cpu().numpy()
    plan.close()
    return res


@pytest.mark.parametrize("kernel,case,name", [
    ("fused", "uneven", "rochester"), ("fused", "uneven", "ssl"), ("fused", "uneven", "rochester_ssl_etafunc"),
    ("fused", "faces", "rochester"), ("fused", "faces", "ssl"), ("fused", "faces", "rochester_ssl_etafunc"),
    ("split", "uneven", "rochester_ssl_etafunc"), ("split", "faces", "ssl")])
def test_gauge_variant_rhs_matches_oracle(monkeypatch, kernel, case, name):
    monkeypatch.setenv("DGR_KERNEL", kernel)
    t = uneven_sample() if case == "uneven" else W.single_block_19(63)
    st = _state(t)
    p = _params(name)
    got = _gpu_rhs_at(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    e = normwise(interior_values(got, t), interior_values(ref, t))
    # the switches are live: the standard gauge without the lapse term differs
    # in alpha (slow-start lapse) and in beta, B (Rochester gauge)
    p0 = _params(name, rochester=0, ssl=0)
    ref0 = O.bssnrhs_batch(st, t, oracle_params(p0), nthreads=8)
    d = normwise(interior_values(ref0, t), interior_values(ref, t))
    log(f"gauge {name} {kernel} {case}: normwise max {e.max():.3e}; against the standard gauge "
        f"alpha {d[0]:.2e} beta {d[6:9].min():.2e} B {d[9:12].min():.2e}")
    assert np.isfinite(interior_values(got, t)).all()
    assert e.max() <= 1e-12
    if p.ssl:
        assert d[0] > 1e-6
    if p.rochester:
        assert d[6:9].min() > 1e-6 and d[9:12].min() > 1e-6


def _rk4_bound(t, st, dt, params_seq):
    """Oracle RK4 steps with the given parameters per step, and the oracle's
    own response to a 2e-16 relative perturbation of the input."""
    ref = st.copy()
    for q in params_seq:
        O.rk4_step(ref, t, dt, oracle_params(q), enforce_per_stage=True, nthreads=8)
    pert = st * (1 + 2e-16 * np.random.default_rng(1).standard_normal(st.shape))
    rp = pert.copy()
    for q in params_seq:
        O.rk4_step(rp, t, dt, oracle_params(q), enforce_per_stage=True, nthreads=8)
    sens = normwise(interior_values(rp - pert, t), interior_values(ref - st, t)).max()
    return ref, sens


@pytest.mark.parametrize("name,nsteps", [("rochester", 1), ("ssl", 2)])
def test_gauge_variant_rk4_steps_match_oracle(name, nsteps):
    """Rochester: one step from the captured graph. Slow-start lapse: two steps,
    the second at time + dt (rkBSSN.cpp:1087 moves the time on per step, and
    every stage reads the step's start time, rkBSSN.cpp:764-836); the same two
    steps at one time are measurably different."""
    import torch
    t = uneven_sample()
    st = _state(t)
    dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    p = _params(name, time=5.0)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_steps(u, dt, p, nsteps=nsteps)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    seq = [_params(name, time=5.0 + k * dt) for k in range(nsteps)]
    ref, sens = _rk4_bound(t, st, dt, seq)
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"gauge {name} rk4 x{nsteps}: interior increments normwise {e:.3e} (oracle sensitivity {sens:.3e})")
    assert e <= max(1e-11, sens)
    if nsteps > 1:
        one_time, _ = _rk4_bound(t, st, dt, [_params(name, time=5.0)] * nsteps)
        d = normwise(interior_values(one_time - st, t), interior_values(ref - st, t))
        log(f"  the same steps at one time: alpha increments differ by {d[0]:.2e}")
        assert d[0] > 100 * max(1e-11, sens)
    plan.set_precision("mixed")
    with pytest.raises(RuntimeError, match="Rochester gauge and the slow-start lapse"):
        plan.rk4_step(to_dev(st), dt, p)
    plan.close()


def test_time_does_not_matter_without_the_lapse_term():
    """Without the slow-start lapse params.time is not read: RK4 steps at two
    times give the same bits (and share one captured graph)."""
    import torch
    t = uneven_sample()
    st = _state(t)
    plan = dgr.Plan(t, st.shape[1])
    res = []
    for time in (0.0, 12.5):
        u = to_dev(st)
        plan.rk4_steps(u, 0.01, dgr.default_params(ko_sigma=0.4, time=time), nsteps=1)
        torch.cuda.synchronize()
        res.append(u.cpu().numpy())
    plan.close()
    np.testing.assert_array_equal(res[0], res[1])


def test_lapse_term_time_is_set_per_call_of_the_graph():
    """dgr_rk4_steps with the slow-start lapse replays one captured graph whose
    time lives in device memory: two calls of one step each, at time and
    time + dt, give the same bits as one call of two steps."""
    import torch
    t = uneven_sample()
    st = _state(t)
    dt = 0.005
    plan = dgr.Plan(t, st.shape[1])
    u2 = to_dev(st)
    plan.rk4_steps(u2, dt, _params("ssl", time=5.0), nsteps=2)
    u1 = to_dev(st)
    plan.rk4_steps(u1, dt, _params("ssl", time=5.0), nsteps=1)
    plan.rk4_steps(u1, dt, _params("ssl", time=5.0 + dt), nsteps=1)
    torch.cuda.synchronize()
    assert plan.status() == 0
    a, b = u1.cpu().numpy(), u2.cpu().numpy()
    plan.close()
    np.testing.assert_array_equal(interior_values(a, t), interior_values(b, t))
    # and a call at another time gives other bits
    plan = dgr.Plan(t, st.shape[1])
    u3 = to_dev(st)
    plan.rk4_steps(u3, dt, _params("ssl", time=9.0), nsteps=2)
    torch.cuda.synchronize()
    c = u3.cpu().numpy()
    plan.close()
    assert np.abs(interior_values(c, t) - interior_values(b, t)).max() > 0
