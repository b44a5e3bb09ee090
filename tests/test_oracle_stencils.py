"""Pins the oracle's stencils (orc_deriv644_*, orc_ko_deriv42_*) to things that
are not the oracle: finite-difference weights derived here from scratch with
Fornberg's algorithm in exact rational arithmetic, the summation-by-parts
structure of the Kreiss-Oliger operator (-H^-1 D3^T D3 / 64h), polynomial
exactness, and the exact set of points each routine writes (reference loop
bounds, BSSN_GR/src/derivs.cpp:2628-2633, 2729-2734, 2843-2848, 2957-2962,
985-990)."""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def fornberg(x0, xs, m):
    """Weights of the m-th derivative at x0 on nodes xs (Fornberg 1988), exact."""
    x0 = Fraction(x0)
    xs = [Fraction(x) for x in xs]
    n = len(xs)
    c = [[Fraction(0)] * (m + 1) for _ in range(n)]
    c[0][0] = Fraction(1)
    c1 = Fraction(1)
    c4 = xs[0] - x0
    for i in range(1, n):
        mn = min(i, m)
        c2 = Fraction(1)
        c5 = c4
        c4 = xs[i] - x0
        for j in range(i):
            c3 = xs[i] - xs[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i][k] = c1 * (k * c[i - 1][k - 1] - c5 * c[i - 1][k]) / c2
                c[i][0] = -c1 * c5 * c[i - 1][0] / c2
            for k in range(mn, 0, -1):
                c[j][k] = (c4 * c[j][k] - k * c[j][k - 1]) / c3
            c[j][0] = c4 * c[j][0] / c3
        c1 = c2
    return [c[i][m] for i in range(n)]


def d_matrix(kind, n, h, lo, hi):
    """Expected 1-D operator as an n x n matrix; rows not written are NaN."""
    M = np.full((n, n), np.nan)
    ie = n - 3
    order = 1 if kind == "d1" else 2
    rows = {}
    for a in range(3, ie):
        rows[a] = list(range(-3, 4))
    if kind in ("d1", "d2"):
        if lo:
            rows[3] = list(range(0, 5)) if kind == "d1" else list(range(0, 6))
            rows[4] = list(range(-1, 4)) if kind == "d1" else list(range(-1, 5))
            rows[5] = list(range(-2, 3))
        if hi:
            rows[ie - 1] = list(range(-4, 1)) if kind == "d1" else list(range(-5, 1))
            rows[ie - 2] = list(range(-3, 2)) if kind == "d1" else list(range(-4, 2))
            rows[ie - 3] = list(range(-2, 3))
        for a, offs in rows.items():
            w = fornberg(0, offs, order)
            M[a, :] = 0.0
            for o, wk in zip(offs, w):
                M[a, a + o] = float(wk) / h ** order
        return M
    # Kreiss-Oliger: -(D3^T D3 u)_a / (64 h w_a), D3 the undivided third difference
    # on the points a face leaves usable; w = 59/48, 43/48, 49/48 at a flagged face.
    lo_pt = 3 if lo else 0
    hi_pt = ie if hi else n
    seg = list(range(lo_pt, hi_pt))
    D3 = np.zeros((len(seg) - 3, n))
    for r in range(len(seg) - 3):
        for o, c in enumerate((-1.0, 3.0, -3.0, 1.0)):
            D3[r, seg[r] + o] = c
    G = D3.T @ D3
    wts = np.ones(n)
    if lo:
        wts[3:6] = [59 / 48, 43 / 48, 49 / 48]
    if hi:
        wts[ie - 3:ie] = [49 / 48, 43 / 48, 59 / 48]
    for a in range(3, ie):
        M[a, :] = -G[a, :] / (64.0 * h * wts[a])
    return M


# routine -> (kind, axis, lower bound of the two other indices)
ROUTINES = {
    "deriv644_x": ("d1", 0, (0, 0)),
    "deriv644_y": ("d1", 1, (3, 0)),
    "deriv644_z": ("d1", 2, (3, 3)),
    "deriv644_xx": ("d2", 0, (3, 3)),
    "deriv644_yy": ("d2", 1, (3, 3)),
    "deriv644_zz": ("d2", 2, (3, 3)),
    "ko_deriv42_x": ("ko", 0, (3, 3)),
    "ko_deriv42_y": ("ko", 1, (3, 3)),
    "ko_deriv42_z": ("ko", 2, (3, 3)),
}


def expected(name, u, h, bflag):
    kind, axis, (lo_a, lo_b) = ROUTINES[name]
    nz, ny, nx = u.shape
    n_ax = (nx, ny, nz)[axis]
    M = d_matrix(kind, n_ax, h, (bflag >> (2 * axis)) & 1, (bflag >> (2 * axis + 1)) & 1)
    out = np.full(u.shape, np.nan)
    Mz = np.nan_to_num(M)
    valid_rows = ~np.isnan(M[:, 0])
    if axis == 0:
        full = np.einsum("ab,kjb->kja", Mz, u)
        sel = (slice(lo_b, nz - lo_b), slice(lo_a, ny - lo_a), valid_rows)
    elif axis == 1:
        full = np.einsum("ab,kbi->kai", Mz, u)
        sel = (slice(lo_b, nz - lo_b), valid_rows, slice(lo_a, nx - lo_a))
    else:
        full = np.einsum("ab,bji->aji", Mz, u)
        sel = (valid_rows, slice(lo_b, ny - lo_b), slice(lo_a, nx - lo_a))
    out[sel] = full[sel]
    return out


def test_fornberg_textbook():
    assert fornberg(0, range(-1, 2), 1) == [Fraction(-1, 2), 0, Fraction(1, 2)]
    assert fornberg(0, range(-1, 2), 2) == [1, -2, 1]
    assert fornberg(0, range(-3, 4), 1) == [Fraction(k, 60) for k in (-1, 9, -45, 0, 45, -9, 1)]


@pytest.mark.parametrize("name", list(ROUTINES))
@pytest.mark.parametrize("bflag", [0, 63, 1, 2, 4, 8, 16, 32, 21, 42])
def test_stencil_matches_independent_weights(name, bflag):
    rng = np.random.default_rng(7 + bflag)
    nx, ny, nz = 13, 14, 15  # distinct edges so an axis mix-up cannot pass
    u = rng.uniform(-1.0, 1.0, (nz, ny, nx))
    h = 0.037
    got = O.stencil(name, u, h, (nx, ny, nz), bflag)
    exp = expected(name, u, h, bflag)
    assert np.array_equal(np.isnan(got), np.isnan(exp)), "written point set differs"
    m = ~np.isnan(exp)
    scale = np.abs(exp[m]).max()
    assert np.abs(got[m] - exp[m]).max() <= 1e-13 * scale


def _grid(n, h, x0=0.3):
    x = x0 + h * np.arange(n)
    return np.meshgrid(x, x, x, indexing="ij")  # Z, Y, X order below


@pytest.mark.parametrize("deg", range(0, 7))
def test_first_derivative_exact_on_polynomials(deg):
    n, h = 15, 0.05
    z, y, x = _grid(n, h)
    u = (x + 0.1) ** deg * (1.0 + 0.5 * y) + z
    exact = deg * (x + 0.1) ** max(deg - 1, 0) * (1.0 + 0.5 * y)
    for bflag in (0, 3):
        got = O.stencil("deriv644_x", u, h, (n, n, n), bflag)
        m = ~np.isnan(got)
        if bflag and deg > 4:
            m[:, :, 3:6] = False  # 4th-order closures: exact only up to degree 4
            m[:, :, n - 6:n - 3] = False
        assert np.allclose(got[m], exact[m], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("deg", range(0, 8))
def test_second_derivative_exact_on_polynomials(deg):
    n, h = 15, 0.05
    z, y, x = _grid(n, h)
    u = (y - 0.2) ** deg
    exact = deg * (deg - 1) * (y - 0.2) ** max(deg - 2, 0)
    for bflag in (0, 12):
        got = O.stencil("deriv644_yy", u, h, (n, n, n), bflag)
        m = ~np.isnan(got)
        if bflag and deg > 4:
            m[:, 3:6, :] = False
            m[:, n - 6:n - 3, :] = False
        assert np.allclose(got[m], exact[m], rtol=1e-8, atol=1e-8)


def test_mixed_derivative_composition_exact():
    n, h = 15, 0.05
    z, y, x = _grid(n, h)
    u = x ** 3 * y ** 2 * (1 + z)
    dx = O.stencil("deriv644_x", u, h, (n, n, n), 0)
    dxy = O.stencil("deriv644_y", np.nan_to_num(dx), h, (n, n, n), 0)
    m = np.zeros_like(u, dtype=bool)
    m[3:n - 3, 3:n - 3, 3:n - 3] = True
    exact = 6 * x ** 2 * y * (1 + z)
    assert np.allclose(dxy[m], exact[m], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("deg", range(0, 6))
def test_ko_annihilates_low_degree(deg):
    n, h = 15, 0.05
    z, y, x = _grid(n, h)
    u = (z + 0.4) ** deg
    for bflag in (0, 48):
        got = O.stencil("ko_deriv42_z", u, h, (n, n, n), bflag)
        m = ~np.isnan(got)
        if bflag and deg > 2:
            m[3:6] = False
            m[n - 6:n - 3] = False
        assert np.abs(got[m]).max() < 1e-9
