"""The oracle's Hamiltonian constraint (C_HAM of physical_constraints,
BSSN_GR/src/physcon.cpp:12-89) and its masked L2 norm
(BSSN_GR/include/grUtils.tcc:37-134), against the reference's generated
constraint equations and against what the mathematics fixes."""
import os

import numpy as np
import pytest

import oracle as O
from dgr import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def interior_flat(c, blocks):
    out = []
    for b in blocks:
        n = int(b["sz"][0])
        off = int(b["offset"])
        out.append(c[off:off + n ** 3].reshape(n, n, n)[3:n - 3, 3:n - 3, 3:n - 3].ravel())
    return np.concatenate(out)


def test_ham_matches_reference_constraint_equations():
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    h = np.load(os.path.join(GOLD, "ham_points.npz"))
    inp = g["inputs"]
    assert np.allclose(h["inputs_sum"], [inp.sum(), np.abs(inp).sum()], rtol=0, atol=1e-9)
    ours = O.ham_points(inp)
    ref = h["ham"]
    assert np.abs(ours - ref).max() / np.abs(ref).max() < 1e-13


def test_ham_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    inp = np.load(os.path.join(GOLD, "algebra_points.npz"))["inputs"]
    h = np.load(os.path.join(GOLD, "ham_points.npz"))["ham"]
    assert np.array_equal(O.ref_ham_points(inp), h)


def test_ham_of_minkowski_is_exactly_zero():
    t = W.uniform_13(2)
    st = np.zeros((24, W.total_padded(t)))
    for v in (0, 1, 12, 15, 17):
        st[v] = 1.0
    ham = O.ham_batch(st, t, nthreads=1)
    assert np.count_nonzero(ham) == 0


def _bl_block(center, n, h):
    t = W.single_block_19(0)
    t["sz"][0] = (n, n, n)
    half = (n - 1) * h / 2.0
    t["pmin"][0] = np.asarray(center) - half
    t["pmax"][0] = np.asarray(center) + half
    return t


def test_ham_of_brill_lindquist_data_converges_to_zero():
    """Brill-Lindquist data (conformally flat, psi harmonic, time symmetric)
    solve the Hamiltonian constraint exactly; the discrete constraint is pure
    truncation error and must shrink at the stencils' 6th order (a dropped or
    wrong term would leave an O(1) residual)."""
    center = (0.4, 1.3, -0.7)
    errs = []
    for n, h in ((19, 0.1), (37, 0.05)):
        t = _bl_block(center, n, h)
        st = np.zeros((24, n ** 3))
        O.init_brill_lindquist(st, t, seed=0, noise=0.0, chi_floor=1e-12)
        ham = O.ham_batch(st, t, nthreads=1).reshape(n, n, n)
        # the same physical points: the coarse interior, every other fine point
        c = ham[3:n - 3, 3:n - 3, 3:n - 3] if n == 19 else ham[6:n - 6:2, 6:n - 6:2, 6:n - 6:2]
        errs.append(np.abs(c).max())
    # scale of the individual terms, for reference: |chi R| ~ |grad^2 chi| ~ 1e-1 here
    assert errs[0] < 1e-5
    assert errs[0] / errs[1] > 40.0, errs


def test_ham_is_invariant_under_axis_swap():
    """A scalar: relabelling x<->y (fields, tensor components, coordinates)
    permutes the points only."""
    from test_oracle_golden import SWAP_XY_VARS
    n = 13
    t = W.single_block_19(0)
    t["sz"][0] = (n, n, n)
    t["pmax"][0] = t["pmin"][0] + (n - 1) * 0.1
    st = np.zeros((24, n ** 3))
    O.init_perturbed_minkowski(st, t, seed=5, amp=1e-2)
    ham = O.ham_batch(st, t, nthreads=1).reshape(n, n, n)
    sw = np.ascontiguousarray(np.swapaxes(st.reshape(24, n, n, n)[SWAP_XY_VARS], 2, 3)).reshape(24, -1)
    t2 = t.copy()
    t2["pmin"][0] = t["pmin"][0][[1, 0, 2]]
    t2["pmax"][0] = t["pmax"][0][[1, 0, 2]]
    ham2 = O.ham_batch(sw, t2, nthreads=1).reshape(n, n, n)
    a = np.swapaxes(ham, 1, 2)[3:-3, 3:-3, 3:-3]
    b = ham2[3:-3, 3:-3, 3:-3]
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_constraint_norm_masks_and_averages_like_the_reference():
    t = W.puncture_octree(0.002)
    npad = W.total_padded(t)
    rng = np.random.default_rng(3)
    c = rng.normal(size=npad)
    R = 2.0  # the reference's default is 5.0; this small octree lies within 5 of a puncture
    got = O.constraint_l2_norm(c, t, bh=((3.0, 0, 0), (-3.0, 0, 0)), r1=R, r2=R)
    vals, skipped = [], 0
    for b in t:
        n = int(b["sz"][0])
        h = (b["pmax"] - b["pmin"]) / (n - 1)
        off = int(b["offset"])
        blk = c[off:off + n ** 3].reshape(n, n, n)
        for k in range(3, n - 3):
            for j in range(3, n - 3):
                for i in range(3, n - 3):
                    x, y, z = b["pmin"] + np.array([i, j, k]) * h
                    d1 = np.sqrt((x - 3) ** 2 + y ** 2 + z ** 2)
                    d2 = np.sqrt((x + 3) ** 2 + y ** 2 + z ** 2)
                    if d1 < R or d2 < R:
                        skipped += 1
                        continue
                    vals.append(blk[k, j, i])
    assert len(vals) > 0 and skipped > 0
    assert got == pytest.approx(np.sqrt(np.mean(np.square(vals))), rel=1e-12)
    assert O.constraint_l2_norm(c, t, r1=1e9, r2=1e9) == -1.0


# --- momentum constraints (CodeGen/constraints.py:256-282) -----------------

def test_mom_matches_reference_constraint_equations():
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    h = np.load(os.path.join(GOLD, "ham_points.npz"))
    _, mom = O.constraint_points(g["inputs"])
    ref = h["mom"]
    assert ref.shape == (g["inputs"].shape[0], 3)
    assert (np.abs(mom - ref).max(axis=0) / np.abs(ref).max(axis=0)).max() < 1e-13


def test_mom_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    inp = np.load(os.path.join(GOLD, "algebra_points.npz"))["inputs"]
    assert np.array_equal(O.ref_constraint_points(inp)[1], np.load(os.path.join(GOLD, "ham_points.npz"))["mom"])


def _flat_block_state(fill):
    """One 13^3 block of flat data (alpha = chi = 1, gt = delta) plus fill(st, x, y, z)."""
    t = W.uniform_13(1)
    n = 13
    st = np.zeros((24, W.total_padded(t)))
    for v in (0, 1, 12, 15, 17):
        st[v] = 1.0
    h = (t["pmax"][0] - t["pmin"][0]) / (n - 1)
    ax = [t["pmin"][0][a] + h[a] * np.arange(n) for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    fill(st, x.ravel(), y.ravel(), z.ravel())
    return t, st


def test_mom_of_minkowski_is_exactly_zero():
    t, st = _flat_block_state(lambda st, x, y, z: None)
    ham, mom = O.constraints_batch(st, t, nthreads=1)
    assert np.count_nonzero(mom) == 0 and np.count_nonzero(ham) == 0


def test_mom_of_linear_K_is_minus_two_thirds_grad_K():
    """Flat metric, At = 0, K = c.x: mom_i = -2/3 c_i (the stencils are exact on
    linear data; a wrong sign or coefficient of the d_i K term fails)."""
    c = np.array([0.3, -0.7, 1.1])

    def fill(st, x, y, z):
        st[2] = c[0] * x + c[1] * y + c[2] * z
    t, st = _flat_block_state(fill)
    _, mom = O.constraints_batch(st, t, nthreads=1)
    for i in range(3):
        got = interior_flat(mom[i], t)
        assert np.abs(got - (-2.0 / 3.0) * c[i]).max() < 1e-12


def test_mom_divergence_term_contracts_the_right_index():
    """Flat metric, At_xy = a y (everything else flat): mom_x = d_y At_xy = a and
    mom_y = d_x At_yx = 0, mom_z = 0. Differentiating along the wrong index of
    At_ij gives mom_x = 0."""
    a = 0.37

    def fill(st, x, y, z):
        st[19] = a * y  # At01
    t, st = _flat_block_state(fill)
    _, mom = O.constraints_batch(st, t, nthreads=1)
    assert np.abs(interior_flat(mom[0], t) - a).max() < 1e-12
    assert np.abs(interior_flat(mom[1], t)).max() < 1e-12
    assert np.abs(interior_flat(mom[2], t)).max() < 1e-12
