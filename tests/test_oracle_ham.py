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


# --- psi4 (CodeGen/constraints.py:42-244) -----------------------------------

def test_psi4_matches_reference_constraint_equations():
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    h = np.load(os.path.join(GOLD, "ham_points.npz"))
    ham, mom, psi4 = O.constraint_points_xyz(g["inputs"], h["xyz"])
    ref = h["psi4"]
    assert ref.shape == (g["inputs"].shape[0], 2)
    assert (np.abs(psi4 - ref).max(axis=0) / np.abs(ref).max(axis=0)).max() < 1e-13
    # ham and mom do not depend on the coordinates
    assert np.array_equal(ham, O.constraint_points(g["inputs"])[0])
    assert np.array_equal(mom, O.constraint_points(g["inputs"])[1])


def test_psi4_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    inp = np.load(os.path.join(GOLD, "algebra_points.npz"))["inputs"]
    h = np.load(os.path.join(GOLD, "ham_points.npz"))
    assert np.array_equal(O.ref_constraints_xyz(inp, h["xyz"])[2], h["psi4"])


def _flat_point():
    """Point input of flat data: alpha = chi = 1, gt = delta, nothing else."""
    inp = np.zeros((1, 163))
    inp[0, 0] = inp[0, 1] = 1.0
    inp[0, 12] = inp[0, 15] = inp[0, 17] = 1.0
    return inp


ON_X = np.array([[1.0, 0.0, 0.0]])  # r = e_x, theta = -e_z, phi = e_y: m = (-e_z + i e_y)/sqrt(2)


def test_psi4_of_curvature_in_the_transverse_plane():
    """Flat metric with only chi's second derivatives set: R_ij = 1/2 chi_ij +
    1/2 delta_ij lap(chi), so at (x,y,z) = (1,0,0) (and anywhere on the x axis,
    where the tetrad is the same) psi4 = -1/4 (chi_zz - chi_yy) - i/2 chi_yz.
    A wrong tetrad orientation or a dropped factor 1/2 of m fails."""
    a, b, c = 0.3, -0.7, 0.45
    inp = _flat_point()
    inp[0, 96 + 6 + 0] = 0.2  # chi_xx (no psi4 contribution)
    inp[0, 96 + 6 + 3], inp[0, 96 + 6 + 4], inp[0, 96 + 6 + 5] = a, c, b
    for x in (1.0, 2.5):
        _, _, p = O.constraint_points_xyz(inp, ON_X * x)
        assert p[0] == pytest.approx([-0.25 * (b - a), -0.5 * c], rel=1e-14)


def test_psi4_of_constant_At():
    """Flat metric, constant At, K = 0, at (1,0,0): the A_temp and quadratic-A
    terms give psi4 = -1/2 At_xx (At_zz - At_yy) + 1/2 (At_xz^2 - At_xy^2)
    + i (At_xz At_xy - At_xx At_yz) (signs of psi4_2 and psi4_4)."""
    A = dict(xx=0.3, xy=-0.2, xz=0.5, yy=0.7, yz=-0.4, zz=0.15)
    inp = _flat_point()
    inp[0, 18:24] = [A["xx"], A["xy"], A["xz"], A["yy"], A["yz"], A["zz"]]
    _, _, p = O.constraint_points_xyz(inp, ON_X)
    want = [-0.5 * A["xx"] * (A["zz"] - A["yy"]) + 0.5 * (A["xz"] ** 2 - A["xy"] ** 2),
            A["xz"] * A["xy"] - A["xx"] * A["yz"]]
    assert p[0] == pytest.approx(want, rel=1e-13)


def test_psi4_of_At_gradients():
    """Flat metric, At = 0 with d_x At_zz = e and d_x At_yz = f at (1,0,0): only
    the magnetic term psi4_3 contributes, psi4 = e/2 + i f (which index of
    d_j At_ki is differentiated and the sign of psi4_3)."""
    e, f = 0.8, -0.35
    inp = _flat_point()
    inp[0, 24 + 3 * 23 + 0] = e  # d_x At_zz
    inp[0, 24 + 3 * 22 + 0] = f  # d_x At_yz
    _, _, p = O.constraint_points_xyz(inp, ON_X)
    assert p[0] == pytest.approx([e / 2, f], rel=1e-14)


def test_psi4_of_minkowski_is_zero_and_axis_is_zeroed():
    t, st = _flat_block_state(lambda st, x, y, z: None)
    _, _, psi4 = O.constraints_batch(st, t, nthreads=1, psi4=True)
    assert np.count_nonzero(psi4) == 0
    # a block through the z axis with data that does not vanish: psi4 is set
    # to zero where |x|, |y| <= 1e-7 (physcon.cpp:82-85) and nowhere else
    t = W._table([13], [(-0.3, -0.3, 0.5)], [0.05], [0])
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=4, amp=1e-2)
    _, _, psi4 = O.constraints_batch(st, t, nthreads=1, psi4=True)
    p = psi4[:, :13 ** 3].reshape(2, 13, 13, 13)[:, 3:10, 3:10, 3:10]
    assert (p[:, :, 3, 3] == 0.0).all()
    assert np.count_nonzero(p) == p.size - 2 * 7
