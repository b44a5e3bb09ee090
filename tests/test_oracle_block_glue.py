"""Pins of the oracle's block glue: the radiative boundary condition
(BSSN_GR/src/rhs.cpp:272-330, bssn_bcs :511-631), the RIT eta(r) damping
(rhs.cpp:209-215) and the curvature-adjusted KO split (CAKO, rhs.cpp:356-433).

Each check compares the oracle's block right-hand side with a closed form
that follows from the mathematics, not from the oracle's own formula:
  * f = f_inf + c r^-p is annihilated by the radiative operator
    -(1/r)(x.grad f + falloff (f - f_inf)) exactly when p = falloff, and
    otherwise gives -(falloff - p) c r^-(p+1);
  * on flat data with zero shift and a uniform B^i, B_rhs^i = -eta(r) B^i;
  * the 7-point KO stencil on a product of sines A prod_a sin(k_a x_a + phi_a)
    is -sum_a sin^6(k_a h_a / 2) / h_a times the product.
Then every check is run again on copies of the oracle with one plausible
mistake planted (a wrong falloff, a flipped sign, a wrong asymptotic value, a
wrong eta exponent, a swapped gauge set) to show that it catches it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

# rhs.cpp:275-327: (falloff, asymptotic value) of each variable's bssn_bcs call
BC_TABLE = {
    "alpha": (1, 1), "chi": (1, 1), "K": (1, 0),
    "Gt0": (2, 0), "Gt1": (2, 0), "Gt2": (2, 0),
    "beta0": (1, 0), "beta1": (1, 0), "beta2": (1, 0),
    "B0": (1, 0), "B1": (1, 0), "B2": (1, 0),
    "gt0": (1, 1), "gt1": (1, 0), "gt2": (1, 0), "gt3": (1, 1), "gt4": (1, 0), "gt5": (1, 1),
    "At0": (2, 0), "At1": (2, 0), "At2": (2, 0), "At3": (2, 0), "At4": (2, 0), "At5": (2, 0),
}
# rhs.cpp:361-432: variables whose KO term takes BSSN_EPSILON_CAKO_GAUGE
GAUGE = {"alpha", "beta0", "beta1", "beta2", "B0", "B1", "B2"}


def coords(n, pmin, h):
    i = np.arange(n)
    z, y, x = np.meshgrid(pmin[2] + i * h, pmin[1] + i * h, pmin[0] + i * h, indexing="ij")
    return x, y, z


def block_rhs(fields, pmin, h, bflag, **params):
    """Oracle right-hand side of one cubic block; fields (24, n, n, n)."""
    n = fields.shape[1]
    blk = O.make_blocks([n], pmin=[pmin], h=[h], bflag=[bflag])
    st = np.ascontiguousarray(fields.reshape(24, -1))
    rhs = O.bssnrhs_batch(st, blk, O.default_params(**params), nthreads=1)
    return rhs.reshape(24, n, n, n)


# ---------------------------------------------------------------------------
# radiative boundary condition
# ---------------------------------------------------------------------------
BC_N, BC_H, BC_PMIN, BC_C = 19, 0.02, (2.1, 1.7, 1.3), 0.1


def falloff_fields(shift):
    """f_v = f_inf + c r^-(falloff + shift) for every variable."""
    x, y, z = coords(BC_N, BC_PMIN, BC_H)
    r = np.sqrt(x * x + y * y + z * z)
    f = np.empty((24, BC_N, BC_N, BC_N))
    p = np.empty(24)
    for v, name in enumerate(O.VAR_NAMES):
        fo, fa = BC_TABLE[name]
        p[v] = fo + shift
        f[v] = fa + BC_C * r ** -p[v]
    return f, p, r


def face_mask(n, bflag):
    m = np.zeros((n, n, n), dtype=bool)
    lo, hi = 3, n - 4
    sl = slice(3, n - 3)
    if bflag & 1: m[sl, sl, lo] = True
    if bflag & 2: m[sl, sl, hi] = True
    if bflag & 4: m[sl, lo, sl] = True
    if bflag & 8: m[sl, hi, sl] = True
    if bflag & 16: m[lo, sl, sl] = True
    if bflag & 32: m[hi, sl, sl] = True
    return m


def check_bc():
    """The face values of every variable against the closed form, for the
    field's own falloff (zero) and for falloffs one above and one below."""
    face = face_mask(BC_N, 63)
    for shift in (0, 1, -1):
        f, p, r = falloff_fields(shift)
        rhs = block_rhs(f, BC_PMIN, BC_H, 63, ko_sigma=0.0)
        for v, name in enumerate(O.VAR_NAMES):
            fo, _ = BC_TABLE[name]
            want = -(fo - p[v]) * BC_C * r ** -(p[v] + 1)
            scale = BC_C * r[face].min() ** -(fo + 1)   # size of a falloff-one-off term
            err = np.abs(rhs[v][face] - want[face]).max() / scale
            # one-sided 4th-order closures at h = 0.02, r >= 2.9: truncation
            # measured 5.3e-8 of a falloff-one-off term
            assert err < 5e-7, (name, shift, err)


def check_bc_point_set():
    """Only the flagged face planes take the boundary value: with each field at
    its own falloff the face values vanish to truncation, while the equations
    one plane further in do not (they see a non-trivial metric)."""
    f, _, _ = falloff_fields(0)
    for bflag in (1, 2, 4, 8, 16, 32, 21, 42):
        rhs = block_rhs(f, BC_PMIN, BC_H, bflag, ko_sigma=0.0)
        face = face_mask(BC_N, bflag)
        gt0 = O.VAR_NAMES.index("gt0")
        assert np.abs(rhs[gt0][face]).max() < 1e-7, bflag
        interior = np.zeros_like(face)
        interior[3:-3, 3:-3, 3:-3] = True
        others = interior & ~face
        # -2 alpha At_00 + ... is O(c^2) = 1e-3 away from the faces
        assert np.abs(rhs[gt0][others]).min() > 1e-5, bflag


def test_bc_closed_form():
    check_bc()


def test_bc_point_set():
    check_bc_point_set()


def test_bc_adds_ko_after_boundary_value():
    """KO on the face is added to the boundary value (rhs.cpp:333-334 run after
    the boundary pass): the difference between sigma and sigma = 0 is sigma
    times the KO sum, evaluated with the closure stencils the face uses."""
    f, _, _ = falloff_fields(1)
    face = face_mask(BC_N, 63)
    r0 = block_rhs(f, BC_PMIN, BC_H, 63, ko_sigma=0.0)
    r1 = block_rhs(f, BC_PMIN, BC_H, 63, ko_sigma=0.4)
    sz = (BC_N,) * 3
    for v in (0, 3, 18):
        ko = sum(O.stencil(nm, f[v], BC_H, sz, 63) for nm in ("ko_deriv42_x", "ko_deriv42_y", "ko_deriv42_z"))
        d = (r1[v] - r0[v])[face]
        np.testing.assert_allclose(d, 0.4 * ko[face], rtol=1e-9, atol=1e-15)


# ---------------------------------------------------------------------------
# RIT eta(r)
# ---------------------------------------------------------------------------
ETA_N, ETA_H, ETA_PMIN = 19, 0.1, (-0.3, -0.3, -0.3)
ETA_C, ETA_O, ETA_W = 2.0, 0.25, 1.0
B_UNIFORM = (0.5, -0.25, 0.75)


def check_eta():
    """Flat data, beta = 0, uniform B: B_rhs = -eta(r) B with
    eta = (c - o) exp(-(r/w)^4) + o; the block spans r in [0, 2.08 w]."""
    f = np.zeros((24, ETA_N, ETA_N, ETA_N))
    for name in ("alpha", "chi", "gt0", "gt3", "gt5"):
        f[O.VAR_NAMES.index(name)] = 1.0
    for i, b in enumerate(B_UNIFORM):
        f[O.VAR_NAMES.index(f"B{i}")] = b
    rhs = block_rhs(f, ETA_PMIN, ETA_H, 0, ko_sigma=0.4, eta_central=ETA_C, eta_outer=ETA_O,
                    eta_width=ETA_W)
    x, y, z = coords(ETA_N, ETA_PMIN, ETA_H)
    r = np.sqrt(x * x + y * y + z * z)
    eta = (ETA_C - ETA_O) * np.exp(-(r / ETA_W) ** 4) + ETA_O
    s = slice(3, ETA_N - 3)
    assert r[s, s, s].min() < 1e-12 and r[s, s, s].max() > 2 * ETA_W
    for i, b in enumerate(B_UNIFORM):
        got = rhs[O.VAR_NAMES.index(f"B{i}")][s, s, s]
        np.testing.assert_allclose(got, -eta[s, s, s] * b, rtol=1e-13, atol=0)
        # beta_rhs = 3/4 (lambda_f0 + lambda_f1 alpha) B (bssn.py b_rhs)
        np.testing.assert_allclose(rhs[O.VAR_NAMES.index(f"beta{i}")][s, s, s], 0.75 * b, rtol=1e-15)


def test_eta_closed_form():
    check_eta()


# ---------------------------------------------------------------------------
# CAKO: sigma = sqrt(chi) eps_gauge / eps_other
# ---------------------------------------------------------------------------
KO_N, KO_H, KO_PMIN = 17, 0.1, (0.4, -0.2, 0.1)
K_WAVE = np.array([9.1, 12.3, 15.7])  # 1/length: k h = 0.91 .. 1.57 (little cancellation)
CHI0, EPS_G, EPS_O = 0.64, 0.9, 0.3


def plane_wave():
    """A product of one sine per axis: the KO stencil along x sees only the
    x factor, so ko_x = -sin^6(k_x h/2)/h times the product, as for a plane
    wave, while every factor's argument stays small (no rounding noise that
    the 6th difference would amplify)."""
    x, y, z = coords(KO_N, KO_PMIN, KO_H)
    return 1e-3 * np.sin(K_WAVE[0] * x + 0.7) * np.sin(K_WAVE[1] * y - 0.3) * np.sin(K_WAVE[2] * z + 1.1)


def check_cako():
    """The same plane wave on all 24 variables (over a flat background with
    chi = 0.64): the KO part of every right-hand side (CAKO on minus
    dissipation off) is sqrt(chi) eps_v times the closed-form KO sum of the
    wave, with eps_v the gauge value for alpha, beta, B and the other value
    for the rest."""
    P = plane_wave()
    f = np.tile(P, (24, 1, 1, 1))
    f[O.VAR_NAMES.index("chi")] += CHI0
    for name in ("alpha", "gt0", "gt3", "gt5"):
        f[O.VAR_NAMES.index(name)] += 1.0
    on = block_rhs(f, KO_PMIN, KO_H, 0, cako_enabled=1, eps_cako_gauge=EPS_G, eps_cako_other=EPS_O)
    off = block_rhs(f, KO_PMIN, KO_H, 0, ko_sigma=0.0)
    ko = -(np.sin(K_WAVE * KO_H / 2) ** 6).sum() / KO_H * P
    s = slice(3, KO_N - 3)
    sq = np.sqrt(f[O.VAR_NAMES.index("chi")])
    for v, name in enumerate(O.VAR_NAMES):
        eps = EPS_G if name in GAUGE else EPS_O
        want = (sq * eps * ko)[s, s, s]
        got = (on[v] - off[v])[s, s, s]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-10, (name, err)  # measured 1.9e-12 (rounding of the 7-term sums)


def test_cako_closed_form():
    check_cako()


def test_ko_plane_wave_closed_form_plain_sigma():
    """Without CAKO every variable takes the same sigma (rhs.cpp:434-489)."""
    P = plane_wave()
    f = np.tile(P, (24, 1, 1, 1))
    for name in ("alpha", "chi", "gt0", "gt3", "gt5"):
        f[O.VAR_NAMES.index(name)] += 1.0
    on = block_rhs(f, KO_PMIN, KO_H, 0, ko_sigma=0.4)
    off = block_rhs(f, KO_PMIN, KO_H, 0, ko_sigma=0.0)
    ko = -(np.sin(K_WAVE * KO_H / 2) ** 6).sum() / KO_H * P
    s = slice(3, KO_N - 3)
    for v in range(24):
        np.testing.assert_allclose((on[v] - off[v])[s, s, s], 0.4 * ko[s, s, s], rtol=1e-9,
                                   atol=1e-12 * np.abs(ko).max())


# ---------------------------------------------------------------------------
# the checks catch planted mistakes (copies of the oracle, built here)
# ---------------------------------------------------------------------------
HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "bssn_oracle.c")

# (name, text in bssn_oracle.c, replacement, the checks that must fail)
MUTANTS = [
    ("Gt falloff 1", "case ORC_GT0: case ORC_GT0 + 1: case ORC_GT0 + 2: *falloff = 2.0;",
     "case ORC_GT0: case ORC_GT0 + 1: case ORC_GT0 + 2: *falloff = 1.0;", ("bc",)),
    ("BC sign", "f_rhs[pp] = -inv_r * (", "f_rhs[pp] = inv_r * (", ("bc",)),
    ("gt asymptote", "*asym = (c == 0 || c == 3 || c == 5) ? 1.0 : 0.0;",
     "*asym = (c == 0 || c == 2 || c == 5) ? 1.0 : 0.0;", ("bc",)),
    ("BC plane", "const int fixed = (face & 1) ? ((axis == 0 ? ie : axis == 1 ? je : ke) - 1) : PW;",
     "const int fixed = (face & 1) ? ((axis == 0 ? ie : axis == 1 ? je : ke) - 2) : PW;", ("bc", "bc_set")),
    ("eta exponent 2", "const double arg = -wr * wr * wr * wr;", "const double arg = -wr * wr;", ("eta",)),
    ("eta centre/outer", "const double eta = (p->eta_central - p->eta_outer) * exp(arg) + p->eta_outer;",
     "const double eta = (p->eta_central - p->eta_outer) * exp(arg) + p->eta_central;", ("eta",)),
    ("B not gauge", "return v == ORC_ALPHA || (v >= ORC_BETA0 && v <= ORC_B0 + 2);",
     "return v == ORC_ALPHA || (v >= ORC_BETA0 && v <= ORC_BETA0 + 2);", ("cako",)),
    ("CAKO chi power", "sigma = sqrt(u[ORC_CHI][pp]);", "sigma = u[ORC_CHI][pp];", ("cako",)),
]
CHECKS = {"bc": check_bc, "bc_set": check_bc_point_set, "eta": check_eta, "cako": check_cako}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not found")
    src = open(SRC).read()
    d = tmp_path_factory.mktemp("oracle_mutants")
    libs = {}
    for i, (name, old, new, _) in enumerate(MUTANTS):
        assert src.count(old) == 1, f"mutation site of '{name}' not found exactly once"
        c = d / f"m{i}.c"
        c.write_text(src.replace(old, new))
        so = d / f"libm{i}.so"
        subprocess.run([cc, "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                        "-I", os.path.dirname(SRC), "-o", str(so), str(c), "-lm"], check=True)
        libs[name] = str(so)
    return libs


@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_checks_catch_planted_mistake(mutant, mutant_libs):
    name, _, _, must_fail = mutant
    with O.using_library(mutant_libs[name]):
        for key in must_fail:
            with pytest.raises(AssertionError):
                CHECKS[key]()
