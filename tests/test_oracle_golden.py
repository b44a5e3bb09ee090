"""The oracle against the reference: golden vectors written from the
reference's generated equations (tests/golden/make_golden.py), plus closed
forms and invariants for the parts no fixture covers."""
import os

import numpy as np
import pytest

import oracle as O
from dgr import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def normwise(a, b, axis=None):
    """max |a-b| / max |b| per variable (variables on axis 0 after moveaxis)."""
    return np.abs(a - b).max(axis=axis) / np.maximum(np.abs(b).max(axis=axis), 1e-300)


def block_params():
    return O.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)


def interior(state, blocks):
    out = []
    for b in blocks:
        n = int(b["sz"][0])
        off = int(b["offset"])
        out.append(state[:, off:off + n ** 3].reshape(24, n, n, n)[:, 3:n - 3, 3:n - 3, 3:n - 3])
    return np.stack(out)


@pytest.mark.parametrize("pset", ["A", "B"])
def test_algebra_matches_reference_vectors(pset):
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    lam, lf = {"A": ((1, 1, 1, 1), (1.0, 0.0)), "B": ((2, 3, 5, 7), (0.7, 0.3))}[pset]
    got = O.bssn_points(g["inputs"], O.default_params(lambda_=lam, lambda_f=lf))
    err = normwise(got, g[f"out_{pset}"], axis=0)
    assert err.max() < 1e-13, err


ETAFUNC_SETS = {"A": ((1, 1, 1, 1), (1.0, 0.0), 1.31, (2.0, 2.0)),
                "B": ((2, 3, 5, 7), (0.7, 0.3), 0.9, (4.0, 1.0))}


@pytest.mark.parametrize("pset", ["A", "B"])
def test_eta_function_algebra_matches_reference_vectors(pset):
    """The reference's BSSN_ETA_FUNCTION build (bssneqs_eta_func_standard_gauge.cpp,
    compiled in place; tests/golden/algebra_points_etafunc.npz): the oracle's
    eta of chi's gradient (orc_eta_func) fed to its point algebra gives the
    reference's 24 outputs to 1e-13; only B's three depend on it, and the same
    eta with the first power raised by one misses them."""
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    ge = np.load(os.path.join(GOLD, "algebra_points_etafunc.npz"))
    lam, lf, r0, pw = ETAFUNC_SETS[pset]
    p = O.default_params(lambda_=lam, lambda_f=lf, eta_func=1, eta_r0=r0, eta_power=pw)
    inp = g["inputs"].copy()
    inp[:, 162] = O.eta_func_points(inp, p)
    got = O.bssn_points(inp, p)
    err = normwise(got, ge[f"out_{pset}"], axis=0)
    assert err.max() < 1e-13, err
    pq = O.default_params(lambda_=lam, lambda_f=lf, eta_func=1, eta_r0=r0, eta_power=(pw[0] + 1, pw[1]))
    inp[:, 162] = O.eta_func_points(inp, pq)
    bad = normwise(O.bssn_points(inp, pq), ge[f"out_{pset}"], axis=0)
    assert bad[9:12].min() > 1e-6 and np.delete(bad, [9, 10, 11]).max() < 1e-13


def test_eta_function_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("reference checkout not present")
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    ge = np.load(os.path.join(GOLD, "algebra_points_etafunc.npz"))
    for k, (lam, lf, r0, pw) in ETAFUNC_SETS.items():
        np.testing.assert_array_equal(O.ref_bssn_points_etafunc(g["inputs"], lam, lf, r0, *pw), ge[f"out_{k}"])


# The reference's other builds (rhs.cpp:219-265) in tests/golden/algebra_points_gauge.npz:
# name -> (eta_func, rochester, ssl, time) with lambda (2, 3, 5, 7), lambda_f (0.7, 0.3),
# BSSN_ETA_R0 0.9, BSSN_ETA_POWER (4, 1), BSSN_XI (2, 3, 5)
GAUGE_SETS = {"rochester": (0, 1, 0, 0.0), "rochester_etafunc": (1, 1, 0, 0.0),
              "ssl": (0, 0, 1, 7.5), "ssl_etafunc": (1, 0, 1, 7.5),
              "rochester_ssl": (0, 1, 1, 31.0), "rochester_ssl_etafunc": (1, 1, 1, 31.0)}


def gauge_params(name, **over):
    ef, ro, ssl, t = GAUGE_SETS[name]
    kw = dict(lambda_=(2, 3, 5, 7), lambda_f=(0.7, 0.3), eta_func=ef, eta_r0=0.9, eta_power=(4.0, 1.0),
              rochester=ro, ssl=ssl, xi=(2, 3, 5), time=t)
    kw.update(over)
    return O.default_params(**kw)


def gauge_outputs(p, inp):
    inp = inp.copy()
    if p.eta_func:
        inp[:, 162] = O.eta_func_points(inp, p)
    return O.bssn_points(inp, p)


@pytest.mark.parametrize("name", list(GAUGE_SETS))
def test_gauge_variant_algebra_matches_reference_vectors(name):
    """The Rochester gauge (bssneqs_*_rochester_gauge*.cpp) and the slow-start
    lapse (bssneqs_*_SSL.cpp) of the reference, compiled in place: the oracle's
    24 outputs agree to 1e-13 (B's exactly 0 under the Rochester gauge). Each
    planted change misses exactly the outputs it touches: the lapse term left
    out or read at another time (alpha), xi[1] and xi[2] swapped or the
    standard gauge's shift (beta), B not frozen (B)."""
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    gg = np.load(os.path.join(GOLD, "algebra_points_gauge.npz"))
    ref = gg[f"out_{name}"]
    p = gauge_params(name)
    err = normwise(gauge_outputs(p, g["inputs"]), ref, axis=0)
    live = np.abs(ref).max(axis=0) > 0
    assert err[live].max() < 1e-13, err
    if p.rochester:
        assert not live[9:12].any()
        np.testing.assert_array_equal(gauge_outputs(p, g["inputs"])[:, 9:12], 0.0)

    def misses(q, idx):
        got = gauge_outputs(q, g["inputs"])
        bad = np.abs(got - ref).max(axis=0) / np.maximum(np.abs(ref).max(axis=0), 1e-300)
        others = np.delete(bad, idx)
        return bad[idx].min() > 1e-6 and others.max() < 1e-13

    ef, ro, ssl, t = GAUGE_SETS[name]
    if ssl:
        assert misses(gauge_params(name, ssl=0), [0])
        assert misses(gauge_params(name, time=t + 1.0), [0])
    if ro:
        assert misses(gauge_params(name, xi=(2, 5, 3)), [6, 7, 8])
        q = gauge_params(name, rochester=0)
        got = gauge_outputs(q, g["inputs"])
        assert np.abs(got[:, 9:12]).max() > 1e-6 and np.abs(got[:, 6:9] - ref[:, 6:9]).max() > 1e-6


def test_gauge_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("reference checkout not present")
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    gg = np.load(os.path.join(GOLD, "algebra_points_gauge.npz"))
    for k in GAUGE_SETS:
        np.testing.assert_array_equal(O.ref_bssn_points_variant(g["inputs"], gauge_params(k)), gg[f"out_{k}"])
    # the default build through the same entry point is the plain one
    p0 = O.default_params(lambda_=(2, 3, 5, 7), lambda_f=(0.7, 0.3))
    np.testing.assert_array_equal(O.ref_bssn_points_variant(g["inputs"], p0),
                                  O.ref_bssn_points(g["inputs"], (2, 3, 5, 7), (0.7, 0.3)))


def test_algebra_fixture_still_matches_reference_build():
    if O.ref_lib() is None:
        pytest.skip("reference checkout not present")
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    out = O.ref_bssn_points(g["inputs"], (1, 1, 1, 1), (1.0, 0.0))
    assert np.array_equal(out, g["out_A"])


@pytest.mark.parametrize("bflag", [0, 63])
def test_block19_matches_reference(bflag):
    g = np.load(os.path.join(GOLD, "block19.npz"))
    t = W.single_block_19(bflag)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    if bflag == 0:
        np.testing.assert_allclose([st.sum(), np.abs(st).sum()], g["input_sum"], rtol=1e-14)
    rhs = O.bssnrhs_batch(st, t, block_params(), nthreads=1)
    got = interior(rhs, t)[0]
    err = normwise(got.reshape(24, -1), g[f"rhs_bflag{bflag}"].reshape(24, -1), axis=1)
    assert err.max() < 1e-12, err


def test_block19_rk4_matches_reference():
    g = np.load(os.path.join(GOLD, "block19.npz"))
    t = W.single_block_19(0)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    u0 = interior(st, t)[0].copy()
    O.rk4_step(st, t, 0.025, block_params(), enforce_per_stage=True, nthreads=1)
    got = interior(st, t)[0]
    ref = g["rk4_dt0.025"]
    # compare the increments u1 - u0 normwise per variable
    err = normwise((got - u0).reshape(24, -1), (ref - u0).reshape(24, -1), axis=1)
    assert err.max() < 1e-10, err


def test_uniform13_matches_reference():
    g = np.load(os.path.join(GOLD, "uniform13_8.npz"))
    t = W.uniform_13(8)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=2, amp=1e-3)
    np.testing.assert_allclose([st.sum(), np.abs(st).sum()], g["input_sum"], rtol=1e-14)
    rhs = O.bssnrhs_batch(st, t, block_params(), nthreads=2)
    got = interior(rhs, t)
    err = normwise(np.moveaxis(got, 1, 0).reshape(24, -1), np.moveaxis(g["rhs"], 1, 0).reshape(24, -1), axis=1)
    assert err.max() < 1e-12, err


@pytest.mark.parametrize("bflag", [0, 63])
def test_minkowski_rhs_is_exactly_zero(bflag):
    t = W.single_block_19(bflag)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=3, amp=0.0)
    rhs = O.bssnrhs_batch(st, t, block_params(), nthreads=1)
    assert np.count_nonzero(interior(rhs, t)) == 0


SWAP_XY_VARS = list(range(24))
SWAP_XY_VARS[3], SWAP_XY_VARS[4] = 4, 3      # Gt^0 <-> Gt^1
SWAP_XY_VARS[6], SWAP_XY_VARS[7] = 7, 6      # beta
SWAP_XY_VARS[9], SWAP_XY_VARS[10] = 10, 9    # B
for base in (12, 18):                         # gt, At: 00<->11, 02<->12
    SWAP_XY_VARS[base + 0], SWAP_XY_VARS[base + 3] = base + 3, base + 0
    SWAP_XY_VARS[base + 2], SWAP_XY_VARS[base + 4] = base + 4, base + 2


@pytest.mark.parametrize("bflag", [0, 1 | 8 | 16])
def test_axis_swap_covariance(bflag):
    """Relabelling x<->y (fields, tensor components, coordinates, face bits)
    relabels the RHS the same way."""
    n = 13
    t = W.single_block_19(bflag)
    t["sz"][0] = (n, n, n)
    t["pmax"][0] = t["pmin"][0] + (n - 1) * 0.1
    st = np.zeros((24, n ** 3))
    O.init_perturbed_minkowski(st, t, seed=5, amp=1e-2)
    p = block_params()
    rhs = O.bssnrhs_batch(st, t, p, nthreads=1)

    def swap(a):
        a = a.reshape(24, n, n, n)[SWAP_XY_VARS]
        return np.ascontiguousarray(np.swapaxes(a, 2, 3)).reshape(24, -1)

    t2 = t.copy()
    t2["pmin"][0] = t["pmin"][0][[1, 0, 2]]
    t2["pmax"][0] = t["pmax"][0][[1, 0, 2]]
    fb = int(bflag)
    t2["bflag"][0] = (fb & 0b110000) | ((fb & 0b11) << 2) | ((fb >> 2) & 0b11)
    rhs2 = O.bssnrhs_batch(swap(st), t2, p, nthreads=1)
    a = interior(swap(rhs), t)[0].reshape(24, -1)
    b = interior(rhs2, t2)[0].reshape(24, -1)
    assert normwise(a, b, axis=1).max() < 1e-12


def test_rk4_closed_form_linear_gauge_mode():
    """Flat space with a uniform B: dbeta/dt = 3/4 B, dB/dt = -eta B, all other
    right-hand sides vanish, so one RK4 step multiplies B by the RK4 stability
    polynomial R(z), z = -eta dt, and adds 3/4 B0 dt (1 + z/2 + z^2/6 + z^3/24)
    to beta (include/rkBSSN.h:50-53). The padding stays frozen at the initial
    state, and each stage's stencils carry that jump at most 3 points inward, so
    only points more than 12 points from the padding are checked."""
    n = 31
    t = W.single_block_19(0)
    t["sz"][0] = (n, n, n)
    t["pmax"][0] = t["pmin"][0] + (n - 1) * 0.1
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=0, amp=0.0)
    B0 = np.array([0.3, -0.2, 0.1])
    for i in range(3):
        st[9 + i] = B0[i]
    eta, dt = 2.0, 0.05
    p = O.default_params(ko_sigma=0.4, eta_central=eta, eta_outer=eta)
    O.rk4_step(st, t, dt, p, enforce_per_stage=True, nthreads=1)
    z = -eta * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    S = 1 + z / 2 + z ** 2 / 6 + z ** 3 / 24
    I = st.reshape(24, n, n, n)[:, 15:16, 15:16, 15:16]
    far = st.reshape(24, n, n, n)[:, 12:n - 12, 12:n - 12, 12:n - 12]
    for i in range(3):
        np.testing.assert_allclose(far[9 + i], B0[i] * R, rtol=1e-14)
        np.testing.assert_allclose(far[6 + i], 0.75 * B0[i] * dt * S, rtol=1e-13)
    assert I.shape[1] == 1
    # and the frozen-padding jump did reach the points next to the padding
    assert abs(st.reshape(24, n, n, n)[9, 3, 3, 3] - B0[0] * R) > 1e-6


@pytest.mark.parametrize("method", ["rk3", "rk3_planted"])
def test_rk3_closed_form_linear_gauge_mode(method):
    """The same linear gauge mode through one step of the SSP-RK3 of
    rkBSSN.cpp:691-762: a third-order method reproduces the cubic Taylor
    polynomial of the linear system, so B is multiplied by
    R3(z) = 1 + z + z^2/2 + z^3/6 (z = -eta dt) and beta gains
    3/4 B0 dt (1 + z/2 + z^2/6). A copy of the stage weights with the 3/4 and
    1/4 of the second stage swapped (the 'planted' case, applied to the
    oracle's own output by recomputing the step from its stages) fails."""
    n = 31
    t = W.single_block_19(0)
    t["sz"][0] = (n, n, n)
    t["pmax"][0] = t["pmin"][0] + (n - 1) * 0.1
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=0, amp=0.0)
    B0 = np.array([0.3, -0.2, 0.1])
    for i in range(3):
        st[9 + i] = B0[i]
    eta, dt = 2.0, 0.05
    p = O.default_params(ko_sigma=0.4, eta_central=eta, eta_outer=eta)
    z = -eta * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6
    S = 1 + z / 2 + z ** 2 / 6
    if method == "rk3":
        O.rk3_step(st, t, dt, p, nthreads=1)
        far = st.reshape(24, n, n, n)[:, 12:n - 12, 12:n - 12, 12:n - 12]
        for i in range(3):
            np.testing.assert_allclose(far[9 + i], B0[i] * R, rtol=1e-14)
            np.testing.assert_allclose(far[6 + i], 0.75 * B0[i] * dt * S, rtol=1e-13)
    else:
        # the stages by hand for dB/dt = -eta B with the second-stage weights swapped
        u0 = B0[0]
        u1 = u0 + dt * (-eta * u0)
        u2 = 0.25 * u0 + 0.75 * u1 + 0.25 * dt * (-eta * u1)
        u3 = u0 / 3 + 2 * u2 / 3 + 2 * dt / 3 * (-eta * u2)
        assert abs(u3 - u0 * R) > 1e-6


def _uniform_gauge_block(n=13):
    """One n^3 block holding flat data with a uniform B, and a neighbour table
    that links the block to itself in every direction (its padding refreshed
    from its own interior): the fields stay uniform, every derivative is 0, and
    the right-hand side is the ODE B' = -eta B, beta' = 3/4 B."""
    t = W.single_block_19(0)
    t["sz"][0] = (n, n, n)
    t["pmax"][0] = t["pmin"][0] + (n - 1) * 0.1
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=0, amp=0.0)
    B0 = np.array([0.3, -0.2, 0.1])
    for i in range(3):
        st[9 + i] = B0[i]
    nbr = np.zeros((1, 27), np.int32)
    return t, st, B0, nbr


# Fehlberg's embedded pair: the fifth-order solution's stability function is
# the degree-5 Taylor polynomial of e^z plus z^6/2080 (Hairer, Norsett and
# Wanner, Solving ODEs I, the RKF45 tableau)
def _r5(z):
    return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 + z ** 5 / 120 + z ** 6 / 2080


@pytest.mark.parametrize("method", ["rk45", "rk45_planted"])
def test_rk45_closed_form_linear_gauge_mode(method):
    """One accepted step of the reference's Runge-Kutta 4(5) (rkBSSN.cpp:
    861-1085) on the uniform gauge mode: B is multiplied by R5(z), z = -eta h,
    and beta gains 3/4 B0 h (R5(z) - 1) / z (the integral of the same stages).
    The 'planted' case recomputes the scalar stages by hand with one tableau
    entry changed (-845/4101 for -845/4104) and shows the check fails."""
    t, st, B0, nbr = _uniform_gauge_block()
    eta, h = 2.0, 0.05
    p = O.default_params(ko_sigma=0.4, eta_central=eta, eta_outer=eta)
    z = -eta * h
    if method == "rk45":
        r = O.rk45_step(st, t, h, 1.0, p, nbr=nbr, general=False, nthreads=1)
        assert r["accepted"] and r["attempts"] == 1 and r["h_used"] == h
        inner = st.reshape(24, 13, 13, 13)[:, 3:10, 3:10, 3:10]
        for i in range(3):
            np.testing.assert_allclose(inner[9 + i], B0[i] * _r5(z), rtol=1e-14)
            np.testing.assert_allclose(inner[6 + i], 0.75 * B0[i] * h * (_r5(z) - 1) / z, rtol=1e-13)
    else:
        from fractions import Fraction as F
        U = [[], [F(1, 4)], [F(3, 32), F(9, 32)], [F(1932, 2197), F(-7200, 2197), F(7296, 2197)],
             [F(439, 216), F(-8), F(3680, 513), F(-845, 4101)],
             [F(-8, 27), F(2), F(-3544, 2565), F(1859, 4104), F(-11, 40)]]
        C5 = [F(16, 135), 0, F(6656, 12825), F(28561, 56430), F(-9, 50), F(2, 55)]
        k = []
        for j in range(6):
            y = B0[0] + h * sum(float(U[j][m]) * k[m] for m in range(j))
            k.append(-eta * y)
        y5 = B0[0] + h * sum(float(C5[j]) * k[j] for j in range(6))
        assert abs(y5 - B0[0] * _r5(z)) > 1e-9


def test_rk45_error_estimate_and_step_control():
    """The error estimate of the reference's pair: with its fourth-order
    weight of stage 3 at 2197/4101 (rkBSSN.h:59; Fehlberg's is 2197/4104) the
    two solutions differ at order h, |B4 - B5| = |B0| |z| d (1 + c3 z + O(z^2))
    with d = 2197/4101 - 2197/4104 and c3 = 12/13 the stage's node (the
    Fehlberg difference itself is O(z^5)). Step control (rkBSSN.cpp:1062-1080):
    with that estimate proportional to h, a tolerance below it is met after
    the retries h <- 0.8 h (tol/err)^(1/4) predicts, and the next step is
    0.8 h (tol/err)^(1/5)."""
    eta, h = 2.0, 0.01
    p = O.default_params(ko_sigma=0.4, eta_central=eta, eta_outer=eta)
    t, st, B0, nbr = _uniform_gauge_block()
    r = O.rk45_step(st, t, h, 1.0, p, nbr=nbr, general=False, nthreads=1)
    z, d = -eta * h, 2197.0 / 4101.0 - 2197.0 / 4104.0
    expect = abs(B0).max() * abs(z) * d * (1 + 12.0 / 13.0 * z)
    assert r["err_norm"] == pytest.approx(expect, rel=1e-3)
    # the same model of the estimate predicts the retries and the accepted step
    def est(hh):
        return abs(B0).max() * eta * hh * d * (1 - 12.0 / 13.0 * eta * hh)
    tol = r["err_norm"] / 20
    hh, tries = h, 1
    while est(hh) > tol:
        hh = 0.8 * hh * (tol / est(hh)) ** 0.25
        tries += 1
    t, st, B0, nbr = _uniform_gauge_block()
    r2 = O.rk45_step(st, t, h, tol, p, nbr=nbr, general=False, nthreads=1)
    assert r2["accepted"] and r2["attempts"] == tries > 1
    assert r2["h_used"] == pytest.approx(hh, rel=1e-3)
    assert r2["err_norm"] <= tol
    assert r2["h_next"] == pytest.approx(0.8 * r2["h_used"] * (tol / r2["err_norm"]) ** 0.2, rel=1e-12)
    inner = st.reshape(24, 13, 13, 13)[:, 3:10, 3:10, 3:10]
    np.testing.assert_allclose(inner[9], B0[0] * _r5(-eta * r2["h_used"]), rtol=1e-14)


def test_enforce_constraints_projects_and_is_idempotent():
    rng = np.random.default_rng(3)
    st = np.zeros((24, 1))
    g = np.eye(3) + 0.1 * rng.uniform(-1, 1, (3, 3))
    g = 1.3 * (g + g.T) / 2
    A = rng.uniform(-1, 1, (3, 3))
    A = (A + A.T) / 2
    comp = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (i, j) in enumerate(comp):
        st[12 + c, 0] = g[i, j]
        st[18 + c, 0] = A[i, j]
    st[0, 0], st[1, 0] = 0.05, 0.02
    assert O.enforce_point(st, 0, 0.1) == 0
    g2 = np.array([[st[12 + comp.index((min(i, j), max(i, j))), 0] for j in range(3)] for i in range(3)])
    A2 = np.array([[st[18 + comp.index((min(i, j), max(i, j))), 0] for j in range(3)] for i in range(3)])
    assert abs(np.linalg.det(g2) - 1.0) < 1e-14
    assert abs(np.trace(np.linalg.inv(g2) @ A2)) < 1e-14
    assert st[0, 0] == 0.1 and st[1, 0] == 0.1
    before = st.copy()
    assert O.enforce_point(st, 0, 0.1) == 0
    np.testing.assert_allclose(st, before, rtol=1e-14, atol=1e-15)


def test_enforce_constraints_keeps_nan_in_alpha_and_chi():
    """The floors are the reference's comparisons (bssn_constraints.h:149-171:
    chi < floor -> floor; alpha = std::max(alpha, floor), which returns its
    first argument unless it is below the second): a NaN in alpha or chi
    stays a NaN, so the after-step NaN check still sees it; values below the
    floor are raised to it."""
    st = np.zeros((24, 2))
    for c, val in enumerate((1.0, 0.0, 0.0, 1.0, 0.0, 1.0)):
        st[12 + c, :] = val
    st[0, 0], st[1, 0] = np.nan, np.nan
    st[0, 1], st[1, 1] = 0.05, 0.02
    O.enforce_point(st, 0, 0.1)
    O.enforce_point(st, 1, 0.1)
    assert np.isnan(st[0, 0]) and np.isnan(st[1, 0])
    assert st[0, 1] == 0.1 and st[1, 1] == 0.1


def test_enforce_constraints_rejects_negative_determinant():
    st = np.zeros((24, 1))
    st[12, 0], st[15, 0], st[17, 0] = 1.0, 1.0, -1.0
    assert O.enforce_point(st, 0, 0.1) == -1


@pytest.mark.parametrize("bflag", [0, 63])
def test_harness_recipe_oracle_matches_reference_equations(bflag):
    """The reference GPU test harness's case (cuda/src/rhs_test.cu: 13^3 blocks
    of spacing 1e-2/13/(b % 3 + 1), fake_initial_data, KO sigma 0.4): the
    oracle's block right-hand side equals the one evaluated with the
    reference's own generated equations to 1e-12 normwise per variable."""
    if O.ref_lib() is None:
        pytest.skip("reference checkout not present")
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from gpu_util import fake_initial_data, harness_blocks
    t = harness_blocks(6, bflag=bflag)
    st = fake_initial_data(t)
    p = O.default_params(ko_sigma=0.4)
    a = O.bssnrhs_batch(st, t, p, nthreads=1)
    with O.reference_block_algebra():
        b = O.bssnrhs_batch(st, t, p, nthreads=1)
    m = O.interior_mask(t, st.shape[1])
    assert np.isfinite(a[:, m]).all()
    err = normwise(b[:, m], a[:, m], axis=1)
    assert err.max() < 1e-12, err


def test_reference_block_algebra_matches_oracle_block():
    """The reference's generated equations run over whole blocks (the CPU
    baseline's configuration) give the oracle's block right-hand side, also
    with faces flagged and on uneven blocks, and the phase timers add up."""
    if O.ref_lib() is None:
        pytest.skip("reference checkout not present")
    for t in (W.single_block_19(63), W.uniform_13(4)):
        st = np.zeros((24, W.total_padded(t)))
        O.init_perturbed_minkowski(st, t, seed=3, amp=1e-3)
        p = O.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=0.25, eta_width=1.0)
        a = O.bssnrhs_batch(st, t, p, nthreads=1)
        O.phase_times(reset=True)
        with O.reference_block_algebra():
            b = O.bssnrhs_batch(st, t, p, nthreads=1)
        ph = O.phase_times(reset=True)
        m = O.interior_mask(t, st.shape[1])
        err = normwise(b[:, m], a[:, m], axis=1)
        assert err.max() < 1e-12, err
        assert all(v >= 0 for v in ph.values()) and ph["deriv"] > 0 and ph["rhs"] > 0 and ph["ko"] > 0
