"""ORACLE — test infrastructure only.

ctypes loader for the CPU restatement (liboracle.so) and, where it was built,
the reference's generated equations compiled in place (_ref/libdgr_ref.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NIN = 163  # per-point algebra input width (fields, grad, grad2, eta)
NVARS = 24

VAR_NAMES = [
    "alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2", "B0", "B1", "B2",
    "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "At0", "At1", "At2", "At3", "At4", "At5",
]


class Params(C.Structure):
    _fields_ = [
        ("lambda_", C.c_double * 4),
        ("lambda_f", C.c_double * 2),
        ("eta_central", C.c_double),
        ("eta_outer", C.c_double),
        ("eta_width", C.c_double),
        ("ko_sigma", C.c_double),
        ("chi_floor", C.c_double),
        ("cako_enabled", C.c_int32),
        ("eta_func", C.c_int32),
        ("eps_cako_gauge", C.c_double),
        ("eps_cako_other", C.c_double),
        ("eta_r0", C.c_double),
        ("eta_power", C.c_double * 2),
        ("rochester", C.c_int32),
        ("ssl", C.c_int32),
        ("xi", C.c_double * 3),
        ("time", C.c_double),
    ]


class Block(C.Structure):
    _fields_ = [
        ("offset", C.c_uint64),
        ("sz", C.c_uint32 * 3),
        ("bflag", C.c_uint32),
        ("pmin", C.c_double * 3),
        ("pmax", C.c_double * 3),
    ]


BLOCK_DTYPE = np.dtype([
    ("offset", "<u8"), ("sz", "<u4", (3,)), ("bflag", "<u4"),
    ("pmin", "<f8", (3,)), ("pmax", "<f8", (3,)),
])
assert BLOCK_DTYPE.itemsize == C.sizeof(Block) == 72

_lib = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _configure(L):
    """Argument and return types of the oracle's C entry points on a loaded library."""
    dp = C.POINTER(C.c_double)
    L.orc_default_params.argtypes = [C.POINTER(Params)]
    L.orc_bssn_points.argtypes = [dp, C.c_int64, C.POINTER(Params), dp]
    L.orc_eta_func.argtypes = [dp, C.POINTER(Params)]
    L.orc_eta_func.restype = C.c_double
    L.orc_bssnrhs_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                    C.POINTER(Params), C.c_int]
    L.orc_bssnrhs_batch.restype = C.c_int
    L.orc_rk4_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double,
                               C.POINTER(Params), C.c_int, C.c_int]
    L.orc_rk4_step.restype = C.c_int
    L.orc_halo_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.orc_rk4_step_halo.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_double,
                                    C.POINTER(Params), C.c_int, C.c_int]
    L.orc_rk4_step_halo.restype = C.c_int
    L.orc_padding_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int]
    L.orc_rk4_step_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_double,
                                    C.POINTER(Params), C.c_int, C.c_int]
    L.orc_rk4_step_fill.restype = C.c_int
    L.orc_lagrange6_half.argtypes = [C.c_int, dp]
    L.orc_rk3_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_double,
                               C.POINTER(Params), C.c_int]
    L.orc_rk3_step.restype = C.c_int
    L.orc_rk45_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                C.POINTER(Params), C.c_int, C.c_int, dp, dp, dp, C.POINTER(C.c_int)]
    L.orc_rk45_step.restype = C.c_int
    L.orc_ham_points.argtypes = [dp, C.c_int64, dp]
    L.orc_ham_batch.argtypes = [dp, C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
    L.orc_ham_batch.restype = C.c_int
    L.orc_constraint_points.argtypes = [dp, C.c_int64, dp, dp]
    L.orc_constraint_points_xyz.argtypes = [dp, dp, C.c_int64, dp, dp, dp]
    L.orc_constraints_batch.argtypes = [dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.c_int]
    L.orc_constraints_batch.restype = C.c_int
    L.orc_constraint_l2_norm.argtypes = [dp, C.c_void_p, C.c_int64, dp, C.c_double, C.c_double]
    L.orc_constraint_l2_norm.restype = C.c_double
    L.orc_constraint_l2_norm_nbr.argtypes = [dp, C.c_void_p, C.c_int64, C.c_void_p, dp, C.c_double, C.c_double]
    L.orc_constraint_l2_norm_nbr.restype = C.c_double
    L.orc_enforce_constraints_point.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_enforce_constraints_point.restype = C.c_int
    for nm in ("orc_deriv644_x", "orc_deriv644_y", "orc_deriv644_z", "orc_deriv644_xx",
               "orc_deriv644_yy", "orc_deriv644_zz", "orc_ko_deriv42_x", "orc_ko_deriv42_y",
               "orc_ko_deriv42_z"):
        f = getattr(L, nm)
        f.argtypes = [dp, dp, C.c_double, C.POINTER(C.c_uint32), C.c_uint32]
    L.orc_init_perturbed_minkowski.argtypes = [C.c_void_p, C.c_void_p, C.c_int64,
                                               C.c_uint64, C.c_double]
    L.orc_init_brill_lindquist.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64,
                                           C.c_double, C.c_double]
    L.orc_set_block_algebra.argtypes = [C.c_void_p]
    L.orc_phase_times.argtypes = [dp, C.c_int]
    return L


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = _configure(C.CDLL(path))
    return _lib


@contextlib.contextmanager
def using_library(path: str):
    """Runs the block in this module against another build of the oracle
    (tests that plant a mistake in a copy of the source use it to show that
    the pins catch it); the default build is restored afterwards."""
    global _lib
    saved = _lib
    _lib = _configure(C.CDLL(os.path.abspath(path)))
    try:
        yield _lib
    finally:
        _lib = saved


def ref_lib():
    """The reference's generated equations compiled in place, or None where absent."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libdgr_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        R.ref_bssn_points.argtypes = [dp, C.c_int64, dp, dp, dp]
        R.ref_bssn_points_etafunc.argtypes = [dp, C.c_int64, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        R.ref_bssn_points_variant.argtypes = [dp, C.c_int64, dp, dp, C.c_double, C.c_double, C.c_double, dp,
                                              C.c_double, C.c_int, dp]
        R.ref_ham_points.argtypes = [dp, C.c_int64, dp]
        R.ref_constraint_points.argtypes = [dp, C.c_int64, dp, dp]
        R.ref_constraints_xyz.argtypes = [dp, dp, C.c_int64, dp, dp, dp]
        _ref = R
    return _ref


PHASES = ("deriv", "rhs", "bc", "ko")


def phase_times(reset: bool = True) -> dict:
    """Seconds per phase of the block right-hand side (summed over blocks and
    threads) since the last reset, at the reference's timer boundaries."""
    out = (C.c_double * 4)()
    lib().orc_phase_times(out, int(reset))
    return {k: out[i] for i, k in enumerate(PHASES)}


@contextlib.contextmanager
def reference_block_algebra():
    """Within the block, the block right-hand side evaluates the reference's own
    generated equations (bssneqs_eta_const_standard_gauge.cpp compiled in
    place, oracle/_ref) over each block instead of the oracle's tensor loops."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref not built (needs the reference checkout)")
    lib().orc_set_block_algebra(C.cast(R.ref_bssn_block_algebra, C.c_void_p))
    try:
        yield
    finally:
        lib().orc_set_block_algebra(None)


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**kw) -> Params:
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "lambda_":
            for i in range(4):
                p.lambda_[i] = v[i]
        elif k == "lambda_f":
            p.lambda_f[0], p.lambda_f[1] = v
        elif k == "eta_power":
            p.eta_power[0], p.eta_power[1] = v
        elif k == "xi":
            p.xi[0], p.xi[1], p.xi[2] = v
        else:
            setattr(p, k, v)
    return p


def bssn_points(inp: np.ndarray, p: Params) -> np.ndarray:
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    out = np.zeros((n, 24))
    lib().orc_bssn_points(_dp(inp), n, C.byref(p), _dp(out))
    return out


def ref_bssn_points(inp: np.ndarray, lam, lamf) -> np.ndarray:
    R = ref_lib()
    assert R is not None, "oracle/_ref not built (needs the reference checkout)"
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    out = np.zeros((n, 24))
    la = np.asarray(lam, dtype=np.float64)
    lf = np.asarray(lamf, dtype=np.float64)
    R.ref_bssn_points(_dp(inp), n, _dp(la), _dp(lf), _dp(out))
    return out


def ref_bssn_points_etafunc(inp: np.ndarray, lam, lamf, r0: float, p0: float, p1: float) -> np.ndarray:
    """The reference's equations of its BSSN_ETA_FUNCTION build
    (bssneqs_eta_func_standard_gauge.cpp) per point; in[162] is not read."""
    R = ref_lib()
    assert R is not None, "oracle/_ref not built (needs the reference checkout)"
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    out = np.zeros((n, 24))
    la = np.asarray(lam, dtype=np.float64)
    lf = np.asarray(lamf, dtype=np.float64)
    R.ref_bssn_points_etafunc(_dp(inp), n, _dp(la), _dp(lf), r0, p0, p1, _dp(out))
    return out


def ref_bssn_points_variant(inp: np.ndarray, p: Params) -> np.ndarray:
    """The reference's equations of the build p selects (rhs.cpp:219-265:
    eta function, Rochester gauge, slow-start lapse at p.time) per point; with
    the eta function in[162] is not read. The Rochester builds write no B_rhs
    (0 here)."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("reference library oracle/_ref/libdgr_ref.so not built")
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    out = np.zeros((n, 24))
    la = np.array(list(p.lambda_), dtype=np.float64)
    lf = np.array(list(p.lambda_f), dtype=np.float64)
    xi = np.array(list(p.xi), dtype=np.float64)
    variant = (1 if p.eta_func else 0) | (2 if p.rochester else 0) | (4 if p.ssl else 0)
    rc = R.ref_bssn_points_variant(_dp(inp), n, _dp(la), _dp(lf), p.eta_r0, p.eta_power[0], p.eta_power[1],
                                   _dp(xi), p.time, variant, _dp(out))
    assert rc == 0
    return out


def eta_func_points(inp: np.ndarray, p: Params) -> np.ndarray:
    """orc_eta_func at each point of an algebra-input array."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    return np.array([lib().orc_eta_func(_dp(inp[q]), C.byref(p)) for q in range(inp.shape[0])])


def ham_points(inp: np.ndarray) -> np.ndarray:
    """Hamiltonian constraint per point (input layout of bssn_points)."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros(inp.shape[0])
    lib().orc_ham_points(_dp(inp), inp.shape[0], _dp(out))
    return out


def ref_ham_points(inp: np.ndarray) -> np.ndarray:
    R = ref_lib()
    assert R is not None, "oracle/_ref not built (needs the reference checkout)"
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros(inp.shape[0])
    R.ref_ham_points(_dp(inp), inp.shape[0], _dp(out))
    return out


def constraint_points(inp: np.ndarray):
    """(ham, mom) per point: C_HAM and the momentum constraints, mom shaped (n, 3)."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    ham = np.zeros(inp.shape[0])
    mom = np.zeros((inp.shape[0], 3))
    lib().orc_constraint_points(_dp(inp), inp.shape[0], _dp(ham), _dp(mom))
    return ham, mom


def ref_constraint_points(inp: np.ndarray):
    """The reference's physconeqs.cpp: (ham, mom) with mom shaped (n, 3)."""
    R = ref_lib()
    assert R is not None, "oracle/_ref not built (needs the reference checkout)"
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    ham = np.zeros(inp.shape[0])
    mom = np.zeros((inp.shape[0], 3))
    R.ref_constraint_points(_dp(inp), inp.shape[0], _dp(ham), _dp(mom))
    return ham, mom


def constraint_points_xyz(inp: np.ndarray, xyz: np.ndarray):
    """(ham, mom, psi4) per point at coordinates xyz (n, 3): mom (n, 3),
    psi4 (n, 2) real and imaginary part."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    assert xyz.shape == (inp.shape[0], 3)
    ham, mom, psi4 = np.zeros(inp.shape[0]), np.zeros((inp.shape[0], 3)), np.zeros((inp.shape[0], 2))
    lib().orc_constraint_points_xyz(_dp(inp), _dp(xyz), inp.shape[0], _dp(ham), _dp(mom), _dp(psi4))
    return ham, mom, psi4


def ref_constraints_xyz(inp: np.ndarray, xyz: np.ndarray):
    """The reference's physconeqs.cpp at coordinates xyz: (ham, mom, psi4)."""
    R = ref_lib()
    assert R is not None, "oracle/_ref not built (needs the reference checkout)"
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    assert xyz.shape == (inp.shape[0], 3)
    ham, mom, psi4 = np.zeros(inp.shape[0]), np.zeros((inp.shape[0], 3)), np.zeros((inp.shape[0], 2))
    R.ref_constraints_xyz(_dp(inp), _dp(xyz), inp.shape[0], _dp(ham), _dp(mom), _dp(psi4))
    return ham, mom, psi4


def stencil(name: str, u: np.ndarray, h: float, sz, bflag: int) -> np.ndarray:
    """Apply one derivative routine to a (nz,ny,nx) array; untouched points stay NaN."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.full(u.shape, np.nan)
    szc = (C.c_uint32 * 3)(*sz)
    getattr(lib(), "orc_" + name)(_dp(out), _dp(u), h, szc, bflag)
    return out


def make_blocks(n_or_sizes, offsets=None, pmin=None, h=None, bflag=None) -> np.ndarray:
    """Block table (BLOCK_DTYPE). sizes are padded cube edges."""
    sizes = list(n_or_sizes)
    nb = len(sizes)
    blk = np.zeros(nb, dtype=BLOCK_DTYPE)
    off = 0
    for b, n in enumerate(sizes):
        blk["offset"][b] = off if offsets is None else offsets[b]
        off += n ** 3
        blk["sz"][b] = (n, n, n)
        blk["bflag"][b] = 0 if bflag is None else bflag[b]
        pm = np.zeros(3) if pmin is None else np.asarray(pmin[b], dtype=np.float64)
        hb = 0.1 if h is None else h[b]
        blk["pmin"][b] = pm
        blk["pmax"][b] = pm + (n - 1) * hb
    return blk


def total_points(blocks: np.ndarray) -> int:
    return int(max(int(b["offset"]) + int(np.prod(b["sz"])) for b in blocks))


def _ptrs(vars_: list[np.ndarray]):
    arr = (C.c_void_p * 24)(*[v.ctypes.data for v in vars_])
    return arr


def state_views(state: np.ndarray) -> list[np.ndarray]:
    """state: (24, N) contiguous; returns the 24 row views."""
    assert state.shape[0] == 24 and state.flags.c_contiguous
    return [state[v] for v in range(24)]


def init_perturbed_minkowski(state: np.ndarray, blocks: np.ndarray, seed: int, amp: float = 1e-3):
    blocks = np.ascontiguousarray(blocks)
    lib().orc_init_perturbed_minkowski(_ptrs(state_views(state)), blocks.ctypes.data,
                                       len(blocks), seed, amp)


def init_brill_lindquist(state: np.ndarray, blocks: np.ndarray, seed: int, noise: float = 1e-6,
                         chi_floor: float = 0.1):
    blocks = np.ascontiguousarray(blocks)
    lib().orc_init_brill_lindquist(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks),
                                   seed, noise, chi_floor)


def bssnrhs_batch(state: np.ndarray, blocks: np.ndarray, p: Params, nthreads: int = 0,
                  rhs: np.ndarray | None = None) -> np.ndarray:
    if rhs is None:
        rhs = np.zeros_like(state)
    blocks = np.ascontiguousarray(blocks)
    err = lib().orc_bssnrhs_batch(_ptrs(state_views(rhs)), _ptrs(state_views(state)),
                                  blocks.ctypes.data, len(blocks), C.byref(p), nthreads)
    if err != 0:
        raise RuntimeError("orc_bssnrhs_batch failed")
    return rhs


def rk4_step(state: np.ndarray, blocks: np.ndarray, dt: float, p: Params,
             enforce_per_stage: bool = True, nthreads: int = 0) -> None:
    """In place: interior points advance one RK4 step, padding stays."""
    blocks = np.ascontiguousarray(blocks)
    err = lib().orc_rk4_step(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks), dt,
                             C.byref(p), int(enforce_per_stage), nthreads)
    if err != 0:
        raise RuntimeError("orc_rk4_step failed (constraint violation)")


def rk3_step(state: np.ndarray, blocks: np.ndarray, dt: float, p: Params, nbr: np.ndarray | None = None,
             general: bool = True, nthreads: int = 0) -> None:
    """In place: one step of the reference's SSP-RK3 (rkBSSN.cpp:691-762),
    constraints enforced after every stage; with a neighbour table the padding
    is refreshed before every stage and after the step, else it stays."""
    blocks = np.ascontiguousarray(blocks)
    nb = None
    if nbr is not None:
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        assert nb.shape == (len(blocks), 27)
    err = lib().orc_rk3_step(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks),
                             None if nb is None else nb.ctypes.data, int(general), dt, C.byref(p), nthreads)
    if err != 0:
        raise RuntimeError("orc_rk3_step failed (constraint violation)")


def rk45_step(state: np.ndarray, blocks: np.ndarray, h: float, tol: float, p: Params,
              nbr: np.ndarray | None = None, general: bool = True, max_attempts: int = 50,
              nthreads: int = 0) -> dict:
    """In place: one step of the reference's embedded Runge-Kutta 4(5) with
    step-size control (rkBSSN.cpp:861-1085). Returns the step's h_used,
    h_next, err_norm (largest |y4 - y5|), attempts and accepted."""
    blocks = np.ascontiguousarray(blocks)
    nb = None
    if nbr is not None:
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        assert nb.shape == (len(blocks), 27)
    hu, hn, en = np.zeros(1), np.zeros(1), np.zeros(1)
    att = C.c_int(0)
    rc = lib().orc_rk45_step(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks),
                             None if nb is None else nb.ctypes.data, int(general), h, tol, C.byref(p),
                             max_attempts, nthreads, hu.ctypes.data_as(C.POINTER(C.c_double)), hn.ctypes.data_as(C.POINTER(C.c_double)),
                             en.ctypes.data_as(C.POINTER(C.c_double)), C.byref(att))
    if rc < 0:
        raise RuntimeError("orc_rk45_step failed (constraint violation)")
    return {"h_used": float(hu[0]), "h_next": float(hn[0]), "err_norm": float(en[0]), "attempts": att.value,
            "accepted": rc == 0}


def halo_fill(state: np.ndarray, blocks: np.ndarray, nbr: np.ndarray) -> None:
    """In place: padding points refreshed from the same-level neighbours in
    nbr (nblocks x 27 int32, -1 for none)."""
    blocks = np.ascontiguousarray(blocks)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    assert nbr.shape == (len(blocks), 27)
    lib().orc_halo_fill(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks), nbr.ctypes.data)


def padding_fill(state: np.ndarray, blocks: np.ndarray, nbr: np.ndarray, general: bool = True) -> None:
    """In place: orc_halo_fill, then (general) every other padding point from
    the block whose interior holds it: same spacing or half of it (copy) or
    twice it (order-6 interpolation in that block's element)."""
    blocks = np.ascontiguousarray(blocks)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    assert nbr.shape == (len(blocks), 27)
    lib().orc_padding_fill(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks), nbr.ctypes.data,
                           int(general))


def lagrange6_half(code: int) -> np.ndarray:
    w = np.zeros(7)
    lib().orc_lagrange6_half(int(code), _dp(w))
    return w


def rk4_step_fill(state: np.ndarray, blocks: np.ndarray, nbr: np.ndarray, dt: float, p: Params,
                  general: bool = True, enforce_per_stage: bool = True, nthreads: int = 0) -> None:
    """rk4_step_halo with padding_fill(general) as the refresh."""
    blocks = np.ascontiguousarray(blocks)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    assert nbr.shape == (len(blocks), 27)
    err = lib().orc_rk4_step_fill(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks), nbr.ctypes.data,
                                  int(general), dt, C.byref(p), int(enforce_per_stage), nthreads)
    if err != 0:
        raise RuntimeError("orc_rk4_step_fill failed (constraint violation)")


def rk4_step_halo(state: np.ndarray, blocks: np.ndarray, nbr: np.ndarray, dt: float, p: Params,
                  enforce_per_stage: bool = True, nthreads: int = 0) -> None:
    """In place: one RK4 step with the padding refreshed from the neighbours
    before every stage and after the step."""
    blocks = np.ascontiguousarray(blocks)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    assert nbr.shape == (len(blocks), 27)
    err = lib().orc_rk4_step_halo(_ptrs(state_views(state)), blocks.ctypes.data, len(blocks), nbr.ctypes.data,
                                  dt, C.byref(p), int(enforce_per_stage), nthreads)
    if err != 0:
        raise RuntimeError("orc_rk4_step_halo failed (constraint violation)")


def ham_batch(state: np.ndarray, blocks: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """C_HAM at the interior points of every block (padding left 0)."""
    ham = np.zeros(state.shape[1])
    blocks = np.ascontiguousarray(blocks)
    err = lib().orc_ham_batch(_dp(ham), _ptrs(state_views(state)), blocks.ctypes.data, len(blocks),
                              nthreads)
    if err != 0:
        raise RuntimeError("orc_ham_batch failed")
    return ham


def constraints_batch(state: np.ndarray, blocks: np.ndarray, nthreads: int = 0, psi4: bool = False):
    """(ham, mom) at the interior points of every block: C_HAM (szpdof) and the
    momentum constraints (3, szpdof); with psi4=True also psi4 (2, szpdof:
    real, imaginary) as a third element; padding left 0."""
    ham = np.zeros(state.shape[1])
    mom = np.zeros((3, state.shape[1]))
    ps = np.zeros((2, state.shape[1])) if psi4 else None
    blocks = np.ascontiguousarray(blocks)
    mp = (C.c_void_p * 3)(*[mom[c].ctypes.data for c in range(3)])
    pp = (C.c_void_p * 2)(*[ps[c].ctypes.data for c in range(2)]) if psi4 else None
    err = lib().orc_constraints_batch(_dp(ham), mp, pp, _ptrs(state_views(state)), blocks.ctypes.data,
                                      len(blocks), nthreads)
    if err != 0:
        raise RuntimeError("orc_constraints_batch failed")
    return (ham, mom, ps) if psi4 else (ham, mom)


def constraint_l2_norm(c: np.ndarray, blocks: np.ndarray, bh=((3.0, 0, 0), (-3.0, 0, 0)),
                       r1: float = 5.0, r2: float = 5.0, nbr: np.ndarray | None = None) -> float:
    """computeConstraintL2Norm over interior points outside r1/r2 of the two BHs;
    with a neighbour table (nblocks x 27 int32) a node shared by neighbouring
    blocks counts once."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    blocks = np.ascontiguousarray(blocks)
    b = np.ascontiguousarray(np.asarray(bh, dtype=np.float64).ravel())
    if nbr is None:
        return float(lib().orc_constraint_l2_norm(_dp(c), blocks.ctypes.data, len(blocks), _dp(b), r1, r2))
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    assert nbr.shape == (len(blocks), 27)
    return float(lib().orc_constraint_l2_norm_nbr(_dp(c), blocks.ctypes.data, len(blocks), nbr.ctypes.data,
                                                  _dp(b), r1, r2))


def enforce_point(state: np.ndarray, node: int, chi_floor: float) -> int:
    return lib().orc_enforce_constraints_point(_ptrs(state_views(state)), node, chi_floor)


def interior_mask(blocks: np.ndarray, npts: int) -> np.ndarray:
    m = np.zeros(npts, dtype=bool)
    for b in blocks:
        nx, ny, nz = (int(s) for s in b["sz"])
        idx = np.zeros((nz, ny, nx), dtype=bool)
        idx[3:nz - 3, 3:ny - 3, 3:nx - 3] = True
        off = int(b["offset"])
        m[off:off + nx * ny * nz] = idx.ravel()
    return m
