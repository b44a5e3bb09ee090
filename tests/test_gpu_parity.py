"""Parity of the HIP path (through the C-ABI) with the oracle and with the
reference fixtures, FP64, normwise relative error per variable <= 1e-12
(max |gpu - ref| / max |ref| over the interior points of a variable)."""
import os

import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import (gpu_rhs, host_state, interior_values, log, normwise, oracle_params,
                      padding_mask, to_dev)

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()  # the HIP library must be present: no fallback
    yield


def block_params():
    return dgr.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)


@pytest.mark.parametrize("bflag", [0, 63, 21, 42, 1, 8, 32])
def test_block19_matches_oracle(bflag):
    log(f"block19 bflag={bflag}")
    t = W.single_block_19(bflag)
    st = host_state(t, seed=1)
    p = block_params()
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=1)
    err = normwise(interior_values(got, t), interior_values(ref, t))
    assert err.max() <= TOL, err
    pad = padding_mask(t, st.shape[1])
    assert np.isnan(got[:, pad]).all(), "padding of the output was written"


@pytest.mark.parametrize("bflag", [0, 63])
def test_block19_matches_reference_fixture(bflag):
    g = np.load(os.path.join(GOLD, "block19.npz"))
    t = W.single_block_19(bflag)
    st = host_state(t, seed=1)
    got = interior_values(gpu_rhs(t, st, block_params()), t)
    ref = g[f"rhs_bflag{bflag}"].reshape(24, -1)
    assert normwise(got, ref).max() <= TOL


def test_uniform13_batch_matches_oracle_and_fixture():
    log("uniform13 x 64")
    t = W.uniform_13(64)
    st = host_state(t, seed=2)
    p = block_params()
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL
    g = np.load(os.path.join(GOLD, "uniform13_8.npz"))
    fix = np.moveaxis(g["rhs"], 1, 0).reshape(24, -1)
    assert normwise(interior_values(got, t[:8]), fix).max() <= TOL


@pytest.mark.parametrize("params", [
    dict(),  # reference defaults: eta varies with r (RIT eta), sigma 0.01
    dict(lambda_=(2, 3, 5, 7), lambda_f=(0.7, 0.3), ko_sigma=0.3),
    dict(cako_enabled=1),
])
def test_parameters_match_oracle(params):
    log(f"params {params}")
    t = W.uniform_13(6)
    t["bflag"][1] = 63
    t["bflag"][4] = 5
    st = host_state(t, seed=11, amp=1e-2)
    p = dgr.default_params(**params)
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=6)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL


def test_uneven_puncture_blocks_match_oracle():
    """Every size class (55^3 -> 343 tiles, 31^3, 19^3 -> 8 ragged tiles, 13^3)
    with Brill-Lindquist data near the punctures."""
    log("uneven puncture blocks")
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
    t["bflag"][0] = 63
    st = host_state(t, init="puncture", seed=5)
    p = dgr.default_params(ko_sigma=0.4)
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL


def test_chunking_does_not_change_results(monkeypatch):
    """Forcing many small chunks (one block each) gives bit-identical output."""
    t = W.uniform_13(40)
    st = host_state(t, seed=3)
    p = block_params()
    a = gpu_rhs(t, st, p)
    monkeypatch.setenv("DGR_CHUNK_POINTS", "4096")
    b = gpu_rhs(t, st, p)
    assert np.array_equal(np.nan_to_num(a, nan=7.0), np.nan_to_num(b, nan=7.0))


def test_full_uniform_batch_properties():
    """4096 x 13^3 (the benchmark batch): flat space gives exactly zero, and a
    sample of blocks matches the oracle run on those blocks alone."""
    import torch
    log("uniform13 x 4096")
    t = W.uniform_13(4096)
    n = W.total_padded(t)
    plan = dgr.Plan(t, n)
    u = torch.zeros((24, n), dtype=torch.float64, device="cuda")
    for v, val in ((0, 1.0), (1, 1.0), (12, 1.0), (15, 1.0), (17, 1.0)):
        u[v] = val
    out = torch.full_like(u, float("nan"))
    plan.rhs(out, u, block_params())
    torch.cuda.synchronize()
    inter = ~torch.from_numpy(padding_mask(t, n)).cuda()
    assert torch.count_nonzero(out[:, inter]).item() == 0
    st = host_state(t, seed=4)
    plan.rhs(out, to_dev(st), block_params())
    got = out.cpu().numpy()
    sample = np.array([0, 1, 777, 2048, 4095])
    sub = t[sample].copy()
    ref = O.bssnrhs_batch(st, sub, oracle_params(block_params()), nthreads=5)
    assert normwise(interior_values(got, sub), interior_values(ref, sub)).max() <= TOL
    plan.close()


@pytest.mark.parametrize("enforce", [True, False])
def test_rk4_step_matches_oracle(enforce):
    import torch
    log(f"rk4 enforce={enforce}")
    t = W.uniform_13(12)
    t["bflag"][3] = 63
    st = host_state(t, seed=9)
    p = block_params()
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_step(u, 0.025, p, enforce_per_stage=enforce)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    O.rk4_step(ref, t, 0.025, oracle_params(p), enforce_per_stage=enforce, nthreads=8)
    # padding untouched, interior increments match
    pad = padding_mask(t, st.shape[1])
    assert np.array_equal(got[:, pad], st[:, pad])
    du_g = interior_values(got - st, t)
    du_r = interior_values(ref - st, t)
    assert normwise(du_g, du_r).max() <= 1e-11
    plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_rk4_after_padding_change_matches_oracle(graph):
    """The RK4 stages alternate between u and a plan-owned work state whose
    padding is copied from u at every call: a caller that refills u's padding
    between calls (as the solver's unzip does) must see the new padding used."""
    import torch
    log(f"rk4 after padding change, graph={graph}")
    t = W.uniform_13(8)
    t["bflag"][5] = 21
    st = host_state(t, seed=13)
    p = block_params()
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    step = (lambda x: plan.rk4_steps(x, 0.025, p, nsteps=1)) if graph else (lambda x: plan.rk4_step(x, 0.025, p))
    step(u)
    torch.cuda.synchronize()
    # new padding: the state of another seed; the interior keeps the first step's result
    pad = padding_mask(t, st.shape[1])
    other = host_state(t, seed=14)
    mid = u.cpu().numpy()
    mid[:, pad] = other[:, pad]
    u.copy_(to_dev(mid))
    step(u)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = mid.copy()
    O.rk4_step(ref, t, 0.025, oracle_params(p), nthreads=8)
    assert np.array_equal(got[:, pad], mid[:, pad])
    assert normwise(interior_values(got - mid, t), interior_values(ref - mid, t)).max() <= 1e-11
    plan.close()


def test_rk4_matches_reference_fixture():
    import torch
    g = np.load(os.path.join(GOLD, "block19.npz"))
    t = W.single_block_19(0)
    st = host_state(t, seed=1)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_step(u, 0.025, block_params(), enforce_per_stage=True)
    torch.cuda.synchronize()
    got = interior_values(u.cpu().numpy(), t)
    u0 = interior_values(st, t)
    ref = g["rk4_dt0.025"].reshape(24, -1)
    assert normwise(got - u0, ref - u0).max() <= 1e-11
    plan.close()


def test_enforce_constraints_matches_oracle():
    import torch
    t = W.uniform_13(4)
    rng = np.random.default_rng(0)
    st = host_state(t, seed=6, amp=0.05)
    st[18:24] += 0.01 * rng.uniform(-1, 1, st[18:24].shape)
    st[1] = rng.uniform(0.02, 0.5, st.shape[1])
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.enforce(u, 0.1)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    ref = st.copy()
    m = ~padding_mask(t, st.shape[1])
    for node in np.nonzero(m)[0]:
        assert O.enforce_point(ref, int(node), 0.1) == 0
    assert np.abs(got - ref).max() <= 1e-14
    plan.close()


def test_negative_metric_determinant_is_reported():
    import torch
    t = W.uniform_13(1)
    st = host_state(t, seed=6, amp=0.0)
    n = int(t["sz"][0][0])
    i = 3 + 3 * n + 3 * n * n + 2  # an interior point
    st[17, i] = -1.0               # gt_zz < 0 -> det < 0
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.enforce(u, 0.1)
    assert plan.status() == dgr.api.C.c_int(-4).value
    assert plan.status() == 0  # cleared
    plan.close()


@pytest.mark.parametrize("bflag", [0, 63, 21])
def test_fused_path_matches_oracle(monkeypatch, bflag):
    """The single-kernel path (no derivative workspace) gives the same RHS."""
    monkeypatch.setenv("DGR_KERNEL", "fused")
    log(f"fused block19 bflag={bflag}")
    t = W.single_block_19(bflag)
    st = host_state(t, seed=1)
    p = block_params()
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=1)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL
    pad = padding_mask(t, st.shape[1])
    assert np.isnan(got[:, pad]).all(), "padding of the output was written"


def test_fused_path_uneven_blocks_match_oracle(monkeypatch):
    monkeypatch.setenv("DGR_KERNEL", "fused")
    log("fused uneven puncture blocks")
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
    t["bflag"][0] = 63
    t["bflag"][3] = 42
    st = host_state(t, init="puncture", seed=5)
    p = dgr.default_params(ko_sigma=0.4, cako_enabled=1)
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL


@pytest.mark.parametrize("enforce", [True, False])
def test_fused_path_rk4_matches_oracle(monkeypatch, enforce):
    import torch
    monkeypatch.setenv("DGR_KERNEL", "fused")
    log(f"fused rk4 enforce={enforce}")
    t = W.uniform_13(6)
    t["bflag"][2] = 63
    st = host_state(t, seed=9)
    p = block_params()
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_step(u, 0.025, p, enforce_per_stage=enforce)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    O.rk4_step(ref, t, 0.025, oracle_params(p), enforce_per_stage=enforce, nthreads=6)
    got = u.cpu().numpy()
    pad = padding_mask(t, st.shape[1])
    assert np.array_equal(got[:, pad], st[:, pad])
    assert normwise(interior_values(got - st, t), interior_values(ref - st, t)).max() <= 1e-11
    plan.close()


@pytest.mark.parametrize("bflag", [0, 63])
def test_split_path_matches_oracle(monkeypatch, bflag):
    """The two-kernel path (derivative workspace) stays selectable and correct."""
    monkeypatch.setenv("DGR_KERNEL", "split")
    t = W.single_block_19(bflag)
    st = host_state(t, seed=1)
    p = block_params()
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=1)
    assert normwise(interior_values(got, t), interior_values(ref, t)).max() <= TOL


def test_split_path_rk4_matches_oracle(monkeypatch):
    import torch
    monkeypatch.setenv("DGR_KERNEL", "split")
    t = W.uniform_13(6)
    t["bflag"][2] = 63
    st = host_state(t, seed=9)
    p = block_params()
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_step(u, 0.025, p, enforce_per_stage=True)
    torch.cuda.synchronize()
    ref = st.copy()
    O.rk4_step(ref, t, 0.025, oracle_params(p), enforce_per_stage=True, nthreads=6)
    got = u.cpu().numpy()
    assert normwise(interior_values(got - st, t), interior_values(ref - st, t)).max() <= 1e-11
    plan.close()
