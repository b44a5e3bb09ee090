"""Tilings of the fused kernel (dgr_plan_set_tiling / dgr_plan_autotune):
every tiling gives the oracle's FP64 result (bitwise the same result across
tilings), the mixed-precision result stays within its error bounds under
every tiling, and the autotuner's choice leaves results and graph replay
intact."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import host_state, interior_values, log, normwise, oracle_params, to_dev

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def mixed_sizes():
    """Two blocks of each octree size (one 55^3 with all faces flagged) and
    three 13^3 blocks: every tiling meets full, partial and one-deep columns."""
    full = W.puncture_octree(0.42)
    pick = []
    for n in (55, 31, 19, 13):
        idx = np.nonzero(full["sz"][:, 0] == n)[0]
        pick += list(idx[:3 if n == 13 else 2])
    t = full[np.array(sorted(pick))].copy()
    off = 0
    for b in range(len(t)):
        t["offset"][b] = off
        off += int(np.prod(t["sz"][b].astype(np.int64)))
    t["bflag"][np.nonzero(t["sz"][:, 0] == 55)[0][0]] = 63
    return t


SIZES = [(13, 13, 13), (19, 19, 19), (31, 31, 31), (55, 55, 55)]


def rhs(plan, u, p):
    import torch
    out = torch.full_like(u, np.nan)
    plan.rhs(out, u, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    return out.cpu().numpy()


def test_fp64_tilings_match_oracle_bitwise_alike():
    t = mixed_sizes()
    st = host_state(t, init="puncture", seed=7)
    p = dgr.default_params(ko_sigma=0.4)
    ref = interior_values(O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8), t)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    results = {}
    for tiling in ("wt", "w"):
        for size in SIZES:
            plan.set_tiling("fp64", size, tiling)
        assert set(plan.tiling("fp64").values()) == {tiling}
        got = interior_values(rhs(plan, u, p), t)
        e = normwise(got, ref).max()
        log(f"fp64 tiling {tiling}: normwise {e:.3e}")
        assert e <= TOL
        results[tiling] = got
    assert np.array_equal(results["wt"], results["w"])
    with pytest.raises(RuntimeError):
        plan.set_tiling("fp64", (13, 13, 13), "full")  # FP32 only
    plan.close()


def test_mixed_tilings_stay_within_error():
    t = mixed_sizes()
    st = host_state(t, init="puncture", seed=7)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    ref = interior_values(rhs(plan, u, p), t)
    plan.set_precision("mixed")
    for tiling in ("wt", "w", "full"):
        for size in SIZES:
            plan.set_tiling("mixed", size, tiling)
        got = interior_values(rhs(plan, u, p), t)
        e = normwise(got, ref).max()
        log(f"mixed tiling {tiling}: RHS normwise vs FP64 {e:.3e}")
        assert e < 3e-6  # measured 1.2e-6
    plan.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_autotune_keeps_results_and_graph(prec):
    import torch
    t = mixed_sizes()
    st = host_state(t, init="puncture", seed=3)
    p = dgr.default_params(ko_sigma=0.4)
    dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)).min())
    plan = dgr.Plan(t, st.shape[1])
    plan.set_precision(prec)
    u = to_dev(st)
    before = interior_values(rhs(plan, u, p), t)
    plan.rk4_steps(u, dt, p, nsteps=2)  # captures a graph with the default tiling
    torch.cuda.synchronize()
    scratch = torch.zeros_like(u)
    plan.autotune(scratch, to_dev(st), p)
    chosen = plan.tiling(prec)
    log(f"{prec} autotuned tiling: {chosen}")
    assert set(chosen) == set(SIZES)  # one 55^3 block is unflagged
    assert all(v in (("wt", "w") if prec == "fp64" else ("wt", "w", "full")) for v in chosen.values())
    after = interior_values(rhs(plan, to_dev(st), p), t)
    if prec == "fp64":
        assert np.array_equal(before, after)
    else:
        assert normwise(after, before).max() < 5e-6
    # same state pointers, dt and parameters as the captured graph: the tile
    # list changed, so the graph must be captured again (replay == eager steps)
    ug = u
    ug.copy_(to_dev(st))
    plan.rk4_steps(ug, dt, p, nsteps=2)
    ue = to_dev(st)
    for _ in range(2):
        plan.rk4_step(ue, dt, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    assert torch.equal(ug, ue)
    plan.close()
