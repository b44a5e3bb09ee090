"""Failure detection on the GPU path (the reference's after-step check,
BSSNCtx::terminal_output, bssnCtx.cpp:1240-1262: the range of alpha, a NaN
stops the run): dgr_variable_range against numpy on the interior values, with
non-finite values planted from the host; a NaN planted in one block's input
stays in that block through RK4 steps (frozen padding: blocks are
independent), the other blocks' results are bit-identical to a clean run, and
the range check reports it."""
import numpy as np
import pytest

import oracle as O
import dgr
from gpu_util import host_state, interior_values, log, to_dev
from test_gpu_mixed import uneven_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def _interior_index(t, b, i, j, k):
    n = t["sz"][b]
    return int(t["offset"][b]) + i + int(n[0]) * (j + int(n[1]) * k)


def test_variable_range_matches_numpy_and_counts_non_finite():
    import torch
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    plan = dgr.Plan(t, st.shape[1])
    for v in (0, 1, 2, 19):
        lo, hi, bad = plan.variable_range(to_dev(st[v]))
        iv = interior_values(st, t)[v]
        assert (lo, hi, bad) == (iv.min(), iv.max(), 0)
    # planted from the host: two NaN and an Inf inside, a NaN in padding (not counted)
    a = st[0].copy()
    a[_interior_index(t, 0, 3, 3, 3)] = np.nan
    a[_interior_index(t, 5, 7, 4, 5)] = np.inf
    a[_interior_index(t, 11, 5, 9, 6)] = np.nan
    a[_interior_index(t, 2, 0, 0, 0)] = np.nan
    lo, hi, bad = plan.variable_range(to_dev(a))
    iv = a[O.interior_mask(t, st.shape[1])]
    fin = iv[np.isfinite(iv)]
    log(f"range with planted values: ({lo:.6e}, {hi:.6e}), {bad} non-finite")
    assert bad == 3 and (lo, hi) == (fin.min(), fin.max())
    lo, hi, bad = plan.variable_range(to_dev(np.full_like(a, np.nan)))
    assert np.isnan(lo) and np.isnan(hi) and bad == plan.interior_points
    torch.cuda.synchronize()
    plan.close()


def test_nan_in_one_block_is_reported_and_stays_in_that_block():
    import torch
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    p = dgr.default_params(ko_sigma=0.4)
    bad_block = 4
    planted = st.copy()
    planted[2, _interior_index(t, bad_block, 6, 6, 6)] = np.nan  # K
    res = {}
    for name, init in (("clean", st), ("planted", planted)):
        plan = dgr.Plan(t, st.shape[1])
        u = to_dev(init)
        plan.rk4_steps(u, dt, p, nsteps=2)
        torch.cuda.synchronize()
        res[name] = (u.cpu().numpy(), plan.variable_range(u[0]))
        plan.close()
    (clean, rc), (got, rg) = res["clean"], res["planted"]
    log(f"alpha range clean {rc}, planted {rg}")
    assert rc[2] == 0 and rg[2] > 0
    m = O.interior_mask(t, st.shape[1])
    inside = np.zeros(st.shape[1], bool)
    n = t["sz"][bad_block]
    inside[int(t["offset"][bad_block]):int(t["offset"][bad_block]) + int(np.prod(n.astype(np.int64)))] = True
    assert not np.isfinite(got[:, m & inside]).all()
    assert np.array_equal(got[:, m & ~inside], clean[:, m & ~inside])


def test_enforcement_keeps_nan_in_alpha_and_chi():
    """dgr_enforce_constraints floors alpha and chi as the reference does
    (chi < floor -> floor, std::max(alpha, floor)): NaN stays NaN, values below
    the floor are raised, the rest is the oracle's projection."""
    import torch
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    a, b = _interior_index(t, 1, 4, 4, 4), _interior_index(t, 7, 5, 6, 7)
    st[0, a], st[1, b] = np.nan, np.nan
    st[0, b], st[1, a] = 0.05, 0.02
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.enforce(u, 0.1)
    torch.cuda.synchronize()
    got = u.cpu().numpy()
    assert np.isnan(got[0, a]) and np.isnan(got[1, b])
    assert got[0, b] == 0.1 and got[1, a] == 0.1
    plan.close()
