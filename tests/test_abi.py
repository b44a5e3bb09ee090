"""CPU checks of the product library: it loads, exports every symbol the
public headers declare, rejects invalid block tables before touching a GPU,
and its generated equations (compiled for the host from the same header the
HIP kernels include) match the reference's golden vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dgr
from dgr import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    for h in ("dgr_bssn.h", "dgr_bssn_shim.hpp"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        names |= set(re.findall(r"\b(dgr_[a-z_0-9]+)\s*\(", txt))
    return names


def test_library_exports_declared_symbols():
    L = dgr.load()
    missing = [s for s in sorted(declared_symbols()) if not hasattr(L, s)]
    assert not missing, missing
    assert set(dgr.api.SYMBOLS) <= declared_symbols()


def test_reference_signature_shim_exported():
    out = os.popen(f"nm -D --defined-only {dgr.lib_path()}").read()
    # void bssnrhs(double**, const double**, const unsigned&, const double*, const double*,
    #              const unsigned*, const unsigned&, double, const double**)  (rhs.h:32-36)
    assert "_Z7bssnrhsPPdPPKdRKjS2_S2_PS4_S5_dS3_" in out


def test_default_params_match_reference_defaults():
    p = dgr.default_params()
    assert list(p.lambda_) == [1, 1, 1, 1] and list(p.lambda_f) == [1.0, 0.0]
    assert (p.eta_central, p.eta_outer, p.eta_width) == (2.0, 0.25, 40.0)
    assert p.ko_sigma == 0.01 and p.chi_floor == 0.1 and p.cako_enabled == 0


@pytest.mark.parametrize("mutate,msg", [
    (lambda t: t["sz"].__setitem__(0, (12, 13, 13)), "size"),
    (lambda t: t["offset"].__setitem__(0, 10 ** 12), "szpdof"),
    (lambda t: t["bflag"].__setitem__(0, 64), "face bits"),
    (lambda t: t["pmax"].__setitem__(0, t["pmin"][0]), "pmax"),
])
def test_plan_rejects_invalid_blocks_without_gpu(mutate, msg):
    t = W.single_block_19()
    mutate(t)
    with pytest.raises(RuntimeError, match=msg):
        dgr.Plan(t, W.total_padded(W.single_block_19()))


def test_entry_points_reject_null_plans_without_gpu():
    """Entry points return DGR_E_INVALID (-1) with a message for a null plan or
    missing arguments, before any HIP call (so this runs without a GPU)."""
    L = dgr.load()
    p = dgr.default_params()
    dt = C.c_double(0.01)
    calls = {
        "dgr_rk45_step": lambda: L.dgr_rk45_step(None, None, C.byref(dt), 1e-3, C.byref(p), 5, None, None, None, None),
        "dgr_rk3_step": lambda: L.dgr_rk3_step(None, None, 0.01, C.byref(p), None),
        "dgr_rk4_step": lambda: L.dgr_rk4_step(None, None, 0.01, C.byref(p), 1, None),
        "dgr_bssn_rhs": lambda: L.dgr_bssn_rhs(None, None, None, 0.0, C.byref(p), None),
        "dgr_padding_fill": lambda: L.dgr_padding_fill(None, None, None),
        "dgr_halo_fill": lambda: L.dgr_halo_fill(None, None, None),
        "dgr_variable_range": lambda: L.dgr_variable_range(None, None, None, None, None, None),
    }
    for name, call in calls.items():
        assert call() == -1, name
        assert L.dgr_last_error().decode(), name


def _host_lib():
    return C.CDLL(os.path.join(ROOT, "dendro-gr_amd", "lib", "libdgr_algebra_host.so"))


@pytest.mark.parametrize("pset", ["A", "B"])
def test_generated_equations_match_reference_vectors(pset):
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    lam, lf = {"A": ((1, 1, 1, 1), (1.0, 0.0)), "B": ((2, 3, 5, 7), (0.7, 0.3))}[pset]
    inp = np.ascontiguousarray(g["inputs"])
    out = np.zeros((len(inp), 24))
    la, lff = np.array(lam, float), np.array(lf, float)
    dp = C.POINTER(C.c_double)
    _host_lib().dgr_host_bssn_points(inp.ctypes.data_as(dp), len(inp), la.ctypes.data_as(dp),
                                     lff.ctypes.data_as(dp), out.ctypes.data_as(dp))
    ref = g[f"out_{pset}"]
    err = np.abs(out - ref).max(0) / np.abs(ref).max(0)
    assert err.max() < 1e-13, err


@pytest.mark.parametrize("name", ["rochester", "rochester_etafunc", "ssl", "ssl_etafunc", "rochester_ssl",
                                  "rochester_ssl_etafunc"])
def test_generated_gauge_equations_match_reference_vectors(name):
    """The generated equations with the gauge variants (bssn_rhs_gauge_generated.h,
    the program the split path runs and, in phase order, the fused kernel's
    gauge instantiations) built on the host against the reference's own
    Rochester-gauge and slow-start-lapse builds (tests/golden/
    algebra_points_gauge.npz); with the eta function, in[162] holds the
    oracle's eta of chi's gradient (pinned to the reference's eta-function build)."""
    import oracle as O
    from test_oracle_golden import gauge_params
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    ref = np.load(os.path.join(GOLD, "algebra_points_gauge.npz"))[f"out_{name}"]
    p = gauge_params(name)
    inp = np.ascontiguousarray(g["inputs"].copy())
    if p.eta_func:
        inp[:, 162] = O.eta_func_points(inp, p)
    out = np.zeros((len(inp), 24))
    la, lff, xi = (np.array(list(v), float) for v in (p.lambda_, p.lambda_f, p.xi))
    dp = C.POINTER(C.c_double)
    L = _host_lib()
    L.dgr_host_bssn_points_gauge.argtypes = [dp, C.c_int64, dp, dp, dp, C.c_int, C.c_int, C.c_double, dp]
    L.dgr_host_bssn_points_gauge(inp.ctypes.data_as(dp), len(inp), la.ctypes.data_as(dp), lff.ctypes.data_as(dp),
                                 xi.ctypes.data_as(dp), p.rochester, p.ssl, p.time, out.ctypes.data_as(dp))
    live = np.abs(ref).max(0) > 0
    err = np.abs(out - ref).max(0)[live] / np.abs(ref).max(0)[live]
    assert err.max() < 1e-13, err
    np.testing.assert_array_equal(out[:, ~live], 0.0)


def test_generated_ham_matches_reference_constraint_vectors():
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    ref = np.load(os.path.join(GOLD, "ham_points.npz"))["ham"]
    inp = np.ascontiguousarray(g["inputs"])
    out = np.zeros(len(inp))
    dp = C.POINTER(C.c_double)
    _host_lib().dgr_host_ham_points(inp.ctypes.data_as(dp), len(inp), out.ctypes.data_as(dp))
    assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-13


def test_generated_constraints_match_reference_constraint_vectors():
    """The generated constraint header's six outputs (ham, mom0..2, psi4 real
    and imaginary) against the reference's physconeqs.cpp vectors
    (CodeGen/constraints.py:42-282) at the fixture's coordinates."""
    g = np.load(os.path.join(GOLD, "algebra_points.npz"))
    ref = np.load(os.path.join(GOLD, "ham_points.npz"))
    inp = np.ascontiguousarray(g["inputs"])
    xyz = np.ascontiguousarray(ref["xyz"])
    out = np.zeros((len(inp), 6))
    dp = C.POINTER(C.c_double)
    _host_lib().dgr_host_constraint_points(inp.ctypes.data_as(dp), xyz.ctypes.data_as(dp), len(inp),
                                           out.ctypes.data_as(dp))
    assert np.abs(out[:, 0] - ref["ham"]).max() / np.abs(ref["ham"]).max() < 1e-13
    for name, cols in (("mom", slice(1, 4)), ("psi4", slice(4, 6))):
        err = np.abs(out[:, cols] - ref[name]).max(0) / np.abs(ref[name]).max(0)
        assert err.max() < 1e-13, (name, err)
