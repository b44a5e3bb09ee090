"""dgr_constraint_l2_norm_noalloc (the entry point that allocates nothing)
against dgr_constraint_l2_norm through the C-ABI: two calls in a row, the
second on a smaller plan, each equal to the allocating norm (same kernels,
same partials: the same sums, so the same norm), with the device's free
memory unchanged across the calls; plus its error paths (no neighbour table,
null output)."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import log

pytestmark = pytest.mark.gpu


def _case(shape, holes, seed):
    t = W.lattice(shape=shape, n=13, h=0.05, holes=holes)
    n = W.total_padded(t)
    c = np.random.default_rng(seed).standard_normal(n) * 1e-3
    return t, n, c


def test_noalloc_norm_matches_allocating_norm_twice():
    import torch
    assert torch.cuda.is_available()
    dgr.load()
    bh, r = ((0.2, 0.1, 0.0), (0.0, 0.0, 0.3)), 0.12
    cases = [((4, 3, 3), ((1, 1, 1),), 1), ((2, 2, 2), ((0, 0, 0),), 2)]  # second: smaller
    plans, inputs = [], []
    for shape, holes, seed in cases:
        t, n, c = _case(shape, holes, seed)
        plan = dgr.Plan(t, n)
        plan.link_neighbors()
        dc = torch.from_numpy(c).cuda()
        # the allocating norm first: its partials buffer and the kernels' code are in place
        want = plan.constraint_norm(dc, bh=bh, r1=r, r2=r)
        ref = O.constraint_l2_norm(c, t, bh=bh, r1=r, r2=r, nbr=W.neighbors(t))
        assert want == pytest.approx(ref, rel=1e-14, abs=0)
        plans.append((plan, dc, want))
        inputs.append(c)
    out = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for k, (plan, dc, want) in enumerate(plans):
        log(f"noalloc norm call {k}: {plan.interior_points} interior points")
        plan.constraint_norm_noalloc(out[k:k + 1], dc, bh=bh, r1=r, r2=r)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    got = out.cpu().numpy()
    for k, (plan, dc, want) in enumerate(plans):
        assert got[k] == pytest.approx(want, rel=1e-15, abs=0), (k, got[k], want)
    assert free1 == free0, (free0, free1)
    # no qualifying point: -1, as the allocating norm returns
    plans[1][0].constraint_norm_noalloc(out[2:3], plans[1][1], r1=1e9, r2=1e9)
    torch.cuda.synchronize()
    assert float(out[2].item()) == -1.0 == plans[1][0].constraint_norm(plans[1][1], r1=1e9, r2=1e9)
    for plan, _, _ in plans:
        plan.close()


def test_noalloc_norm_needs_the_neighbour_table():
    import torch
    dgr.load()
    t, n, c = _case((2, 2, 1), (), 3)
    plan = dgr.Plan(t, n)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="neighbour table"):
        plan.constraint_norm_noalloc(out, torch.from_numpy(c).cuda())
    plan.close()


def test_noalloc_norm_rejects_plans_beyond_its_scratch():
    """A plan of more norm workgroups than the scratch holds (800000 blocks of
    13^3, 274 M interior points) is refused before any launch."""
    import torch
    dgr.load()
    t = W.uniform_13(800000)
    plan = dgr.Plan(t, W.total_padded(t))
    plan.link_neighbors()
    dc = torch.zeros(1, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="scratch holds"):
        plan.constraint_norm_noalloc(out, dc)
    plan.close()
