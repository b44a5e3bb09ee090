"""Padding refresh from same-level neighbours on the GPU (dgr_halo_fill,
dgr_plan_set_halo_refresh, through the C-ABI) against the oracle: the
library's neighbour table equals the brute-force one, the fill is bit-exact
(it only copies), and RK4 with the refresh between stages matches the
oracle's stage-by-stage step, eagerly and from the captured graph."""
import numpy as np
import pytest

import oracle as O
import dgr
from dgr import workloads as W
from gpu_util import interior_values, log, normwise, oracle_params, to_dev
from test_oracle_halo import smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


def lattice():
    t = W.lattice(shape=(3, 3, 2), holes=[(1, 1, 0)], extra_level=True)
    t["bflag"][0] = 63  # a face-flagged block keeps its links
    return t


def test_neighbour_table_matches_brute_force():
    log("halo: neighbour table")
    t = lattice()
    plan = dgr.Plan(t, W.total_padded(t))
    got = plan.link_neighbors()
    assert np.array_equal(got, W.neighbors(t))
    plan.close()


def test_fill_is_bit_exact():
    import torch
    log("halo: fill")
    t = lattice()
    st = np.random.default_rng(3).standard_normal((24, W.total_padded(t)))
    nbr = W.neighbors(t)
    plan = dgr.Plan(t, st.shape[1])
    u = to_dev(st)
    plan.halo_fill(u)
    torch.cuda.synchronize()
    ref = st.copy()
    O.halo_fill(ref, t, nbr)
    assert np.array_equal(u.cpu().numpy(), ref)
    assert not np.array_equal(ref, st)
    plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_rk4_with_refresh_matches_oracle(graph):
    import torch
    log(f"halo: rk4 graph={graph}")
    t = lattice()
    st = smooth_state(t, amp=1e-2)
    nbr = W.neighbors(t)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_halo_refresh(True)
    u = to_dev(st)
    dt = 0.02
    if graph:
        plan.rk4_steps(u, dt, p, nsteps=2)
    else:
        for _ in range(2):
            plan.rk4_step(u, dt, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    for _ in range(2):
        O.rk4_step_halo(ref, t, nbr, dt, oracle_params(p), enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    # padding: refreshed from the new interiors where linked, untouched elsewhere
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref - st).max() + 0.0
    plan.close()


def test_refresh_off_keeps_padding_frozen():
    import torch
    log("halo: refresh off")
    t = lattice()
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_halo_refresh(True)
    plan.set_halo_refresh(False)
    u = to_dev(st)
    plan.rk4_steps(u, 0.02, p, nsteps=1)
    torch.cuda.synchronize()
    ref = st.copy()
    O.rk4_step(ref, t, 0.02, oracle_params(p), enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    inner = np.zeros(st.shape[1], bool)
    for b in range(len(t)):
        n = int(t["sz"][b][0])
        m = np.zeros((n, n, n), bool)
        m[3:n - 3, 3:n - 3, 3:n - 3] = True
        inner[t["offset"][b]:t["offset"][b] + n ** 3] = m.ravel()
    assert np.array_equal(got[:, ~inner], st[:, ~inner])
    assert normwise(interior_values(got - st, t), interior_values(ref - st, t)).max() <= 1e-11
    plan.close()


def test_mixed_precision_rk4_with_refresh_stays_within_error():
    """Mixed precision with the in-kernel refresh (the FP32 stage kernels read
    same-level padding from the neighbours as the FP64 ones do). Two steps
    against the oracle's FP64 stage-by-stage step: the increments within 2e-5
    normwise (measured 7.1e-6; FP32 derivatives and equations; the FP64 path
    meets 1e-11)."""
    import torch
    log("halo: mixed rk4")
    t = lattice()
    st = smooth_state(t, amp=1e-2)
    nbr = W.neighbors(t)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_precision("mixed")
    plan.set_halo_refresh(True)
    u = to_dev(st)
    plan.rk4_steps(u, 0.02, p, nsteps=2)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    for _ in range(2):
        O.rk4_step_halo(ref, t, nbr, 0.02, oracle_params(p), enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  mixed interior increments normwise {e:.3e}")
    assert e <= 2e-5
    plan.close()


def test_refresh_on_a_lattice_of_19_cubed_blocks():
    """Neighbours of 19^3 blocks start 12 nodes apart: the fill stays bit-exact
    and RK4 with the in-kernel refresh matches the oracle (two tiles per axis
    per block, so tiles see both own and neighbour padding)."""
    import torch
    log("halo: 19^3 lattice")
    t = W.lattice(shape=(2, 2, 2), n=19, h=0.04, origin=(-0.5, -0.4, -0.3))
    nbr = W.neighbors(t)
    plan = dgr.Plan(t, W.total_padded(t))
    assert np.array_equal(plan.link_neighbors(), nbr)
    rnd = np.random.default_rng(5).standard_normal((24, W.total_padded(t)))
    u = to_dev(rnd)
    plan.halo_fill(u)
    torch.cuda.synchronize()
    ref = rnd.copy()
    O.halo_fill(ref, t, nbr)
    assert np.array_equal(u.cpu().numpy(), ref)
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan.set_halo_refresh(True)
    u = to_dev(st)
    plan.rk4_steps(u, 0.01, p, nsteps=1)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    O.rk4_step_halo(ref, t, nbr, 0.01, oracle_params(p), enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    plan.close()


# --- the rest of the refresh: blocks of other sizes and across level jumps ---

def fill_cases():
    """A two-level lattice (one coarse cell refined: copies, injections from
    the finer blocks, interpolations from the coarser ones), a lattice of one
    spacing and two block sizes (copies only), and the lattice with a hole
    and a face-flagged block (diagonal neighbours fill padding next to the
    hole)."""
    return {"two_level": W.two_level_lattice(), "mixed_size": W.mixed_size_lattice(), "holes": lattice()}


@pytest.mark.parametrize("case", ["two_level", "mixed_size", "holes"])
def test_padding_fill_is_bit_exact(case):
    """dgr_padding_fill equals the oracle's orc_padding_fill(general) bit for
    bit at every point (copies, and interpolations summed in the oracle's
    order with contraction off)."""
    import torch
    log(f"padding fill: {case}")
    t = fill_cases()[case]
    st = np.random.default_rng(5).standard_normal((24, W.total_padded(t)))
    plan = dgr.Plan(t, st.shape[1])
    ncopy, ninterp = plan.fill_sources()
    log(f"  {ncopy} copied, {ninterp} interpolated padding points")
    if case == "two_level":
        assert ncopy > 0 and ninterp > 0
    if case == "mixed_size":
        assert ncopy > 0 and ninterp == 0
    u = to_dev(st)
    plan.padding_fill(u)
    torch.cuda.synchronize()
    ref = st.copy()
    O.padding_fill(ref, t, W.neighbors(t), general=True)
    got = u.cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_rk4_with_full_refresh_matches_oracle(graph):
    """RK4 on the two-level lattice with the padding of every stage refreshed
    across the level jump (DGR_REFRESH_FULL) against the oracle's
    orc_rk4_step_fill(general): interior increments within 1e-11 normwise,
    the refreshed padding within 1e-11 of the increments' scale."""
    import torch
    log(f"full refresh: rk4 graph={graph}")
    t = W.two_level_lattice()
    st = smooth_state(t, amp=1e-2)
    nbr = W.neighbors(t)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_halo_refresh("full")
    u = to_dev(st)
    dt = 0.02
    if graph:
        plan.rk4_steps(u, dt, p, nsteps=2)
    else:
        for _ in range(2):
            plan.rk4_step(u, dt, p)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    for _ in range(2):
        O.rk4_step_fill(ref, t, nbr, dt, oracle_params(p), general=True, enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref - st).max()
    # the level-jump refresh changes the answer: without it the steps differ
    ref2 = st.copy()
    for _ in range(2):
        O.rk4_step_halo(ref2, t, nbr, dt, oracle_params(p), enforce_per_stage=True, nthreads=8)
    assert np.abs(ref2 - ref).max() > 1e-9
    plan.close()


def test_refined_octree_fill_and_rk4_match_oracle():
    """A three-level 2:1-balanced octree of 13^3 blocks (288 blocks): the
    padding fill is bit-exact, and one graph-replayed RK4 step with the full
    refresh between stages matches the oracle's orc_rk4_step_fill(general)."""
    import torch
    log("full refresh: refined octree (288 blocks)")
    t = W.refined_octree(2, 4, 1.0)
    nbr = W.neighbors(t)
    plan = dgr.Plan(t, W.total_padded(t))
    ncopy, ninterp = plan.fill_sources()
    log(f"  {ncopy} copied, {ninterp} interpolated padding points")
    assert ncopy > 0 and ninterp > 0
    rnd = np.random.default_rng(9).standard_normal((24, W.total_padded(t)))
    u = to_dev(rnd)
    plan.padding_fill(u)
    torch.cuda.synchronize()
    ref = rnd.copy()
    O.padding_fill(ref, t, nbr, general=True)
    assert np.array_equal(u.cpu().numpy(), ref)
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan.set_halo_refresh("full")
    u = to_dev(st)
    dt = 0.25 * float(((t["pmax"][:, 0] - t["pmin"][:, 0]) / 12).min())
    plan.rk4_steps(u, dt, p, nsteps=1)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = st.copy()
    O.rk4_step_fill(ref, t, nbr, dt, oracle_params(p), general=True, enforce_per_stage=True, nthreads=8)
    got = u.cpu().numpy()
    e = normwise(interior_values(got - st, t), interior_values(ref - st, t)).max()
    log(f"  interior increments normwise {e:.3e}")
    assert e <= 1e-11
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref - st).max()
    plan.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_copy_path_equals_in_kernel_refresh(prec):
    """The refresh copied into the stage input before each stage (the path of
    plans of 2^29 points or more, forced here with DGR_HALO_IN_KERNEL=0 in a
    child process) gives the same bits as the in-kernel refresh, on the
    two-level lattice with the full refresh, in both precisions (in mixed
    precision the staging offset of a tile whose first point is linked
    padding is read from the neighbour, where the copy would come from)."""
    import os
    import subprocess
    import sys
    import torch
    log(f"refresh: copy path vs in-kernel, {prec}")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    t = W.two_level_lattice()
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    plan = dgr.Plan(t, st.shape[1])
    plan.set_precision(prec)
    plan.set_halo_refresh("full")
    u = to_dev(st)
    plan.rk4_steps(u, 0.02, p, nsteps=2)
    torch.cuda.synchronize()
    assert plan.status() == 0
    got = u.cpu().numpy()
    plan.close()
    code = ("import sys, numpy as np, torch; sys.path[:0] = {paths!r}\n"
            "import dgr\nfrom dgr import workloads as W\nfrom test_oracle_halo import smooth_state\n"
            "t = W.two_level_lattice(); st = smooth_state(t, amp=1e-2)\n"
            "plan = dgr.Plan(t, st.shape[1]); plan.set_precision({prec!r}); plan.set_halo_refresh('full')\n"
            "u = torch.from_numpy(st).cuda(); plan.rk4_steps(u, 0.02, dgr.default_params(ko_sigma=0.4), nsteps=2)\n"
            "torch.cuda.synchronize(); np.save(sys.argv[1], u.cpu().numpy())\n").format(prec=prec, paths=[here, root, os.path.join(root, "oracle"), os.path.join(root, "dendro-gr_amd")])
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "copy.npy")
        env = dict(os.environ, DGR_HALO_IN_KERNEL="0")
        subprocess.run([sys.executable, "-c", code, out], check=True, env=env, timeout=300)
        ref = np.load(out)
    assert np.array_equal(got, ref)
