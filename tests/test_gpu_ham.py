"""Hamiltonian and momentum constraints, psi4 and the masked norm on the GPU
(through the C-ABI) against the oracle: per point within 1e-12 normwise
relative, the norm within 1e-12 relative (the reference sums in another
order)."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import host_state, log, padding_mask, to_dev

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def gpu_ham(t, st):
    import torch
    plan = dgr.Plan(t, st.shape[1])
    ham = torch.full((st.shape[1],), float("nan"), dtype=torch.float64, device="cuda")
    plan.ham(ham, to_dev(st))
    torch.cuda.synchronize()
    plan.close()
    return ham.cpu().numpy()


def uneven_sample():
    full = W.puncture_octree(0.42)
    pick = []
    for n in (55, 31, 19, 13):
        idx = np.nonzero(full["sz"][:, 0] == n)[0]
        pick += list(idx[:2])
    t = full[np.array(sorted(pick))].copy()
    off = 0
    for b in range(len(t)):
        t["offset"][b] = off
        off += int(np.prod(t["sz"][b].astype(np.int64)))
    return t


@pytest.mark.parametrize("bflag", [0, 63, 21])
def test_ham_block19_matches_oracle(bflag):
    log(f"ham block19 bflag={bflag}")
    t = W.single_block_19(bflag)
    st = host_state(t, seed=1, amp=1e-2)
    got = gpu_ham(t, st)
    ref = O.ham_batch(st, t, nthreads=1)
    m = ~padding_mask(t, st.shape[1])
    assert np.abs(got[m] - ref[m]).max() <= TOL * np.abs(ref[m]).max()
    assert np.isnan(got[~m]).all(), "padding of the output was written"


def test_ham_uneven_puncture_blocks_match_oracle():
    """Brill-Lindquist data satisfy the constraint: ham is a truncation
    residual (here ~1e-4) of terms of order one (chi, its derivatives, the
    Ricci scalar) that cancel, so agreement is measured against that term
    scale, max(|ham|, 1), rather than against the residual itself."""
    log("ham uneven puncture blocks")
    t = uneven_sample()
    t["bflag"][0] = 63
    st = host_state(t, init="puncture", seed=5)
    got = gpu_ham(t, st)
    ref = O.ham_batch(st, t, nthreads=8)
    m = ~padding_mask(t, st.shape[1])
    assert np.abs(got[m] - ref[m]).max() <= TOL * max(np.abs(ref[m]).max(), 1.0)


def test_constraint_norm_matches_oracle():
    import torch
    log("constraint norm")
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    ref_ham = O.ham_batch(st, t, nthreads=8)
    plan = dgr.Plan(t, st.shape[1])
    c = torch.from_numpy(ref_ham).cuda()
    for r in (0.0, 2.0, 5.0):
        got = plan.constraint_norm(c, r1=r, r2=r)
        want = O.constraint_l2_norm(ref_ham, t, r1=r, r2=r)
        assert got == pytest.approx(want, rel=1e-12, abs=0), r
    assert plan.constraint_norm(c, r1=1e9, r2=1e9) == -1.0
    plan.close()


def test_constraint_norm_counts_shared_nodes_once():
    """On lattices whose neighbours share their boundary nodes (full, and with
    holes so that some edge nodes are shared by diagonal neighbours only), the
    norm counts each node once and equals the oracle's with the neighbour
    table (which is pinned to a brute-force norm over the distinct nodes)."""
    import torch
    log("constraint norm on lattices (shared nodes)")
    for shape, holes in (((2, 2, 2), ()), ((3, 2, 2), ((1, 0, 0), (0, 1, 1)))):
        t = W.lattice(shape=shape, n=13, h=0.05, holes=holes)
        nbr = W.neighbors(t)
        n = W.total_padded(t)
        c = np.cos(np.arange(n) * 0.37) + 0.2
        plan = dgr.Plan(t, n)
        dc = torch.from_numpy(c).cuda()
        for bh, r in ((((30.0, 0, 0), (-30.0, 0, 0)), 5.0), (((0.2, 0.1, 0.0), (0.0, 0.0, 0.3)), 0.12)):
            got = plan.constraint_norm(dc, bh=bh, r1=r, r2=r)
            want = O.constraint_l2_norm(c, t, bh=bh, r1=r, r2=r, nbr=nbr)
            assert got == pytest.approx(want, rel=1e-14, abs=0), (shape, holes, r)
        plan.close()


def test_ham_norm_unchanged_over_an_rk4_step():
    """The constraint norm of the state after one GPU RK4 step equals the
    oracle's (same step on the CPU) to far better than 3 digits."""
    import torch
    log("ham norm over rk4")
    t = uneven_sample()
    st = host_state(t, init="puncture", seed=5)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    hmin = float(np.min((t["pmax"][:, 0] - t["pmin"][:, 0]) / (t["sz"][:, 0] - 1)))
    dt = 0.25 * hmin
    plan.rk4_step(u, dt, p, enforce_per_stage=True)
    ham = torch.zeros(st.shape[1], dtype=torch.float64, device="cuda")
    plan.ham(ham, u)
    got = plan.constraint_norm(ham, r1=1.0, r2=1.0)
    assert plan.status() == 0
    ref = st.copy()
    O.rk4_step(ref, t, dt, O.default_params(ko_sigma=0.4), enforce_per_stage=True, nthreads=8)
    want = O.constraint_l2_norm(O.ham_batch(ref, t, nthreads=8), t, r1=1.0, r2=1.0)
    assert want > 0.0
    assert got == pytest.approx(want, rel=1e-9)
    plan.close()


def gpu_constraints(t, st, which=(True,) * 6):
    """dgr_constraints into six NaN-filled arrays: ham, mom0..2, psi4 real and
    imaginary (an output left out of `which` is passed as NULL)."""
    import torch
    plan = dgr.Plan(t, st.shape[1])
    out = torch.full((6, st.shape[1]), float("nan"), dtype=torch.float64, device="cuda")
    sel = [out[i] if w else None for i, w in enumerate(which)]
    plan.constraints(sel[0], sel[1:4], to_dev(st), psi4=sel[4:6])
    torch.cuda.synchronize()
    plan.close()
    return out.cpu().numpy()


def check_constraints(got, t, st, scale_floor=0.0):
    ham, mom, psi4 = O.constraints_batch(st, t, nthreads=8, psi4=True)
    m = ~padding_mask(t, st.shape[1])
    for i, ref in enumerate([ham] + list(mom) + list(psi4)):
        err = np.abs(got[i][m] - ref[m]).max()
        scale = max(np.abs(ref[m]).max(), scale_floor)
        log(f"  output {i}: max err {err:.3e} scale {scale:.3e}")
        assert scale > 0.0
        assert err <= TOL * scale, i
    assert np.isnan(got[:, ~m]).all(), "padding of the output was written"


@pytest.mark.parametrize("bflag", [0, 63])
def test_momentum_block19_matches_oracle(bflag):
    """dgr_constraints (ham and mom0..2) on a perturbed flat block against the
    oracle's physical_constraints restatement."""
    log(f"momentum block19 bflag={bflag}")
    t = W.single_block_19(bflag)
    st = host_state(t, seed=2, amp=1e-2)
    check_constraints(gpu_constraints(t, st), t, st)


def test_momentum_uneven_puncture_blocks_match_oracle():
    """Brill-Lindquist data with 1e-3 noise on every variable, so that At, K
    and the metric carry momentum-constraint violations of their own; ham is
    measured against max(|ham|, 1) as in the Hamiltonian-only test."""
    log("momentum uneven puncture blocks")
    t = uneven_sample()
    t["bflag"][0] = 63
    st = host_state(t, init="puncture", seed=6, noise=1e-3)
    got = gpu_constraints(t, st)
    ham, mom, psi4 = O.constraints_batch(st, t, nthreads=8, psi4=True)
    m = ~padding_mask(t, st.shape[1])
    assert np.abs(got[0][m] - ham[m]).max() <= TOL * max(np.abs(ham[m]).max(), 1.0)
    for i in range(3):
        assert np.abs(got[1 + i][m] - mom[i][m]).max() <= TOL * np.abs(mom[i][m]).max(), i
    for i in range(2):
        assert np.abs(got[4 + i][m] - psi4[i][m]).max() <= TOL * np.abs(psi4[i][m]).max(), i


def test_constraints_skip_null_outputs():
    import torch
    log("constraints null outputs")
    t = W.single_block_19(21)
    st = host_state(t, seed=3, amp=1e-2)
    got = gpu_constraints(t, st, which=(False, True, False, True, False, False))
    assert np.isnan(got[0]).all() and np.isnan(got[2]).all() and np.isnan(got[4:]).all()
    full = gpu_constraints(t, st)
    m = ~padding_mask(t, st.shape[1])
    for i in (1, 3):
        assert np.abs(got[i][m] - full[i][m]).max() <= TOL * np.abs(full[i][m]).max()
    ham_only = gpu_ham(t, st)
    assert np.abs(ham_only[m] - full[0][m]).max() <= TOL * np.abs(full[0][m]).max()
    plan = dgr.Plan(t, st.shape[1])
    with pytest.raises(RuntimeError):
        plan.constraints(None, None, to_dev(st))
    with pytest.raises(RuntimeError):
        plan.constraints(None, [None, None, None], to_dev(st), psi4=[None, None])
    plan.close()
    only_psi4 = gpu_constraints(t, st, which=(False, False, False, False, False, True))
    assert np.isnan(only_psi4[:5]).all()
    assert np.abs(only_psi4[5][m] - full[5][m]).max() <= TOL * np.abs(full[5][m]).max()


def test_psi4_on_a_block_through_the_z_axis():
    """A 13^3 block whose grid contains x = y = 0: psi4 is exactly zero on the
    axis (physcon.cpp:82-85) and matches the oracle everywhere else, where the
    tetrad degenerates as the axis is approached."""
    log("psi4 block through the z axis")
    t = W._table([13], [(-0.3, -0.3, 0.5)], [0.05], [0])
    st = host_state(t, seed=4, amp=1e-2)
    got = gpu_constraints(t, st)
    check_constraints(got, t, st)
    b = t[0]
    h = (b["pmax"] - b["pmin"]) / (b["sz"] - 1)
    x = b["pmin"][0] + np.arange(13) * h[0]
    i0 = int(np.argmin(np.abs(x)))
    assert abs(x[i0]) <= 1e-7
    sl = got[4:, b["offset"]:b["offset"] + 13 ** 3].reshape(2, 13, 13, 13)  # (c, k, j, i)
    assert (sl[:, 3:10, i0, i0] == 0.0).all()
    assert (sl[:, 3:10, i0, i0 + 1] != 0.0).all()
