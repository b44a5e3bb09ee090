"""ctypes binding of include/dgr_bssn.h."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .workloads import BLOCK_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
NVARS = 24
VAR_NAMES = [
    "alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2", "B0", "B1", "B2",
    "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "At0", "At1", "At2", "At3", "At4", "At5",
]

# exported C symbols (include/dgr_bssn.h)
SYMBOLS = [
    "dgr_default_params", "dgr_last_error", "dgr_plan_create", "dgr_plan_destroy",
    "dgr_plan_interior_points", "dgr_bssn_rhs", "dgr_rk4_step", "dgr_enforce_constraints",
    "dgr_plan_status", "dgr_plan_set_timing", "dgr_plan_kernel_times",
    "dgr_ham_constraint", "dgr_constraint_l2_norm", "dgr_rk4_steps", "dgr_plan_set_precision",
    "dgr_plan_autotune", "dgr_plan_tiling", "dgr_plan_set_tiling", "dgr_constraints",
    "dgr_plan_link_neighbors", "dgr_halo_fill", "dgr_plan_set_halo_refresh",
    "dgr_constraint_l2_norm_noalloc", "dgr_padding_fill", "dgr_plan_fill_sources", "dgr_rk3_step",
    "dgr_rk45_step", "dgr_variable_range",
]
PRECISION = {"fp64": 0, "mixed": 1}
TILINGS = ("wt", "w", "full")  # DGR_TILING_WT, _W, _FULL


class BssnParams(C.Structure):
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


def lib_path() -> str:
    """lib/libdgr_bssn.so of this package, or the build named by DGR_LIBRARY
    (for comparing builds made with different compiler options)."""
    return os.environ.get("DGR_LIBRARY") or os.path.join(PKG, "lib", "libdgr_bssn.so")


_lib = None


def load():
    """Loads lib/libdgr_bssn.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with dendro-gr_amd/build.py")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.dgr_default_params.argtypes = [C.POINTER(BssnParams)]
        L.dgr_last_error.restype = C.c_char_p
        L.dgr_plan_create.argtypes = [C.c_int, vp, C.c_uint32, C.c_uint64, C.POINTER(vp)]
        L.dgr_plan_destroy.argtypes = [vp]
        L.dgr_plan_interior_points.argtypes = [vp]
        L.dgr_plan_interior_points.restype = C.c_int64
        L.dgr_bssn_rhs.argtypes = [vp, vp, vp, C.c_double, C.POINTER(BssnParams), vp]
        L.dgr_rk4_step.argtypes = [vp, vp, C.c_double, C.POINTER(BssnParams), C.c_int, vp]
        L.dgr_enforce_constraints.argtypes = [vp, vp, C.c_double, vp]
        L.dgr_plan_status.argtypes = [vp, vp]
        L.dgr_plan_set_timing.argtypes = [vp, C.c_int]
        L.dgr_plan_kernel_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.dgr_plan_set_precision.argtypes = [vp, C.c_int]
        L.dgr_rk4_steps.argtypes = [vp, vp, C.c_double, C.POINTER(BssnParams), C.c_int, C.c_int, vp]
        L.dgr_rk3_step.argtypes = [vp, vp, C.c_double, C.POINTER(BssnParams), vp]
        L.dgr_variable_range.argtypes = [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int64), vp]
        L.dgr_rk45_step.argtypes = [vp, vp, C.POINTER(C.c_double), C.c_double, C.POINTER(BssnParams), C.c_int,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), vp]
        L.dgr_ham_constraint.argtypes = [vp, vp, vp, vp]
        L.dgr_constraints.argtypes = [vp, vp, vp, vp, vp, vp]
        L.dgr_plan_link_neighbors.argtypes = [vp, vp, C.POINTER(C.c_int64)]
        L.dgr_halo_fill.argtypes = [vp, vp, vp]
        L.dgr_plan_set_halo_refresh.argtypes = [vp, C.c_int]
        L.dgr_padding_fill.argtypes = [vp, vp, vp]
        L.dgr_plan_fill_sources.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.dgr_plan_autotune.argtypes = [vp, vp, vp, C.POINTER(BssnParams), vp]
        L.dgr_plan_tiling.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        L.dgr_plan_set_tiling.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        L.dgr_constraint_l2_norm.argtypes = [vp, vp, C.POINTER(C.c_double), C.c_double, C.c_double,
                                             C.POINTER(C.c_double), vp]
        L.dgr_constraint_l2_norm_noalloc.argtypes = [vp, vp, C.POINTER(C.c_double), C.c_double, C.c_double,
                                                     vp, vp]
        _lib = L
    return _lib


def default_params(**kw) -> BssnParams:
    p = BssnParams()
    load().dgr_default_params(C.byref(p))
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


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {load().dgr_last_error().decode()}")


def var_pointers(state) -> C.Array:
    """24 device pointers into a (24, szpdof) float64 CUDA tensor."""
    assert state.dim() == 2 and state.shape[0] == NVARS and state.is_contiguous()
    assert str(state.dtype) == "torch.float64" and state.is_cuda
    base = state.data_ptr()
    stride = state.stride(0) * 8
    return (C.c_void_p * NVARS)(*[base + v * stride for v in range(NVARS)])


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Plan:
    """A block batch prepared on the device (dgr_plan)."""

    def __init__(self, blocks: np.ndarray, szpdof: int, device: int = 0):
        blocks = np.ascontiguousarray(blocks, dtype=BLOCK_DTYPE)
        self._blocks = blocks
        self.szpdof = int(szpdof)
        h = C.c_void_p()
        L = load()
        _check(L.dgr_plan_create(device, blocks.ctypes.data, len(blocks), self.szpdof, C.byref(h)),
               "dgr_plan_create")
        self._h = h

    @property
    def interior_points(self) -> int:
        return int(load().dgr_plan_interior_points(self._h))

    def rhs(self, out, state, params: BssnParams, t: float = 0.0, stream=None):
        _check(load().dgr_bssn_rhs(self._h, var_pointers(out), var_pointers(state), t,
                                   C.byref(params), _stream_ptr(stream)), "dgr_bssn_rhs")

    def rk4_step(self, state, dt: float, params: BssnParams, enforce_per_stage: bool = True,
                 stream=None):
        _check(load().dgr_rk4_step(self._h, var_pointers(state), dt, C.byref(params),
                                   int(enforce_per_stage), _stream_ptr(stream)), "dgr_rk4_step")

    def rk4_steps(self, state, dt: float, params: BssnParams, nsteps: int, enforce_per_stage: bool = True,
                  stream=None):
        """nsteps RK4 steps replayed from a captured HIP graph."""
        _check(load().dgr_rk4_steps(self._h, var_pointers(state), dt, C.byref(params), int(enforce_per_stage),
                                    int(nsteps), _stream_ptr(stream)), "dgr_rk4_steps")

    def rk3_step(self, state, dt: float, params: BssnParams, stream=None):
        """One step of the reference's SSP-RK3 (BSSN_RK_TYPE = 0), in place."""
        _check(load().dgr_rk3_step(self._h, var_pointers(state), dt, C.byref(params), _stream_ptr(stream)),
               "dgr_rk3_step")

    def rk45_step(self, state, dt: float, tol: float, params: BssnParams, max_attempts: int = 50, stream=None):
        """One step of the reference's embedded Runge-Kutta 4(5) with step-size
        control (BSSN_RK_TYPE = 2), in place. Returns the step's h_used,
        h_next, err_norm (max |y4 - y5|), attempts and accepted (False when
        max_attempts tries all exceeded tol; the state then holds the last)."""
        h = C.c_double(dt)
        hu, en, att = C.c_double(0.0), C.c_double(0.0), C.c_int(0)
        rc = load().dgr_rk45_step(self._h, var_pointers(state), C.byref(h), tol, C.byref(params), max_attempts,
                                  C.byref(hu), C.byref(en), C.byref(att), _stream_ptr(stream))
        if rc not in (0, -5):  # DGR_E_STEP: not accepted, reported below
            _check(rc, "dgr_rk45_step")
        return {"h_used": hu.value, "h_next": h.value, "err_norm": en.value, "attempts": att.value,
                "accepted": rc == 0}

    def variable_range(self, var, stream=None):
        """(min, max, non-finite count) of one variable (a szpdof float64 CUDA
        tensor, e.g. state[0] for alpha) over the interior points: the
        reference's after-step check (bssnCtx.cpp:1240-1262)."""
        assert var.dim() == 1 and var.is_cuda and str(var.dtype) == "torch.float64" and var.is_contiguous()
        lo, hi, bad = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
        _check(load().dgr_variable_range(self._h, C.c_void_p(var.data_ptr()), C.byref(lo), C.byref(hi),
                                         C.byref(bad), _stream_ptr(stream)), "dgr_variable_range")
        return lo.value, hi.value, bad.value

    def enforce(self, state, chi_floor: float, stream=None):
        _check(load().dgr_enforce_constraints(self._h, var_pointers(state), chi_floor,
                                              _stream_ptr(stream)), "dgr_enforce_constraints")

    def ham(self, out, state, stream=None):
        """Hamiltonian constraint at the interior points into `out` (a szpdof
        float64 CUDA tensor)."""
        assert out.dim() == 1 and out.is_cuda and str(out.dtype) == "torch.float64"
        _check(load().dgr_ham_constraint(self._h, C.c_void_p(out.data_ptr()), var_pointers(state),
                                         _stream_ptr(stream)), "dgr_ham_constraint")

    def constraints(self, ham, mom, state, psi4=None, stream=None):
        """Hamiltonian constraint into `ham`, the three momentum constraints
        into `mom` (a (3, szpdof) tensor or a sequence of three 1-D tensors,
        entries may be None) and psi4 (real, imaginary) into `psi4` (likewise,
        two entries); each argument may be None."""
        def ptr(t):
            if t is None:
                return None
            assert t.dim() == 1 and t.is_cuda and str(t.dtype) == "torch.float64"
            return t.data_ptr()
        mp = pp = None
        if mom is not None:
            mp = (C.c_void_p * 3)(*[ptr(mom[i]) for i in range(3)])
        if psi4 is not None:
            pp = (C.c_void_p * 2)(*[ptr(psi4[i]) for i in range(2)])
        _check(load().dgr_constraints(self._h, C.c_void_p(ptr(ham)), mp, pp, var_pointers(state),
                                      _stream_ptr(stream)), "dgr_constraints")

    def link_neighbors(self):
        """Same-level neighbour table (nblocks x 27 int32, -1 for none) as the
        library found it from the block geometry."""
        import numpy as np
        n = len(self._blocks)
        out = np.full((n, 27), -1, dtype=np.int32)
        links = C.c_int64()
        _check(load().dgr_plan_link_neighbors(self._h, C.c_void_p(out.ctypes.data), C.byref(links)),
               "dgr_plan_link_neighbors")
        assert int(links.value) == int((out >= 0).sum())
        return out

    def halo_fill(self, state, stream=None):
        """Padding of `state` refreshed in place from same-level neighbours."""
        _check(load().dgr_halo_fill(self._h, var_pointers(state), _stream_ptr(stream)), "dgr_halo_fill")

    def padding_fill(self, state, stream=None):
        """All of `state`'s padding refreshed in place: from same-level neighbours,
        and elsewhere from the block of the same, half or twice the spacing
        whose interior holds the point (copy, injection, order-6 interpolation)."""
        _check(load().dgr_padding_fill(self._h, var_pointers(state), _stream_ptr(stream)), "dgr_padding_fill")

    def fill_sources(self):
        """(copied, interpolated) padding points of the refresh beyond the
        same-level neighbours."""
        a, b = C.c_int64(), C.c_int64()
        _check(load().dgr_plan_fill_sources(self._h, C.byref(a), C.byref(b)), "dgr_plan_fill_sources")
        return int(a.value), int(b.value)

    def set_halo_refresh(self, on):
        """RK4 entry points refresh the padding before every stage: False/0 off,
        True/1 from same-level neighbours, 2 or "full" also across level jumps
        and from blocks of other sizes."""
        mode = 2 if on == "full" else int(on)
        _check(load().dgr_plan_set_halo_refresh(self._h, mode), "dgr_plan_set_halo_refresh")

    def constraint_norm(self, c, bh=((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0)), r1: float = 5.0,
                        r2: float = 5.0, stream=None) -> float:
        b = (C.c_double * 6)(*[float(v) for p in bh for v in p])
        res = C.c_double()
        _check(load().dgr_constraint_l2_norm(self._h, C.c_void_p(c.data_ptr()), b, r1, r2, C.byref(res),
                                             _stream_ptr(stream)), "dgr_constraint_l2_norm")
        return res.value

    def constraint_norm_noalloc(self, out, c, bh=((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0)), r1: float = 5.0,
                                r2: float = 5.0, stream=None):
        """dgr_constraint_l2_norm_noalloc: the norm into the device double out[0],
        asynchronously, with no allocation (needs link_neighbors() first)."""
        b = (C.c_double * 6)(*[float(v) for p in bh for v in p])
        _check(load().dgr_constraint_l2_norm_noalloc(self._h, C.c_void_p(c.data_ptr()), b, r1, r2,
                                                     C.c_void_p(out.data_ptr()), _stream_ptr(stream)),
               "dgr_constraint_l2_norm_noalloc")

    def status(self, stream=None) -> int:
        return int(load().dgr_plan_status(self._h, _stream_ptr(stream)))

    def set_precision(self, precision: str):
        """'fp64' (default) or 'mixed' (FP32 derivatives/equations, FP64 state)."""
        _check(load().dgr_plan_set_precision(self._h, PRECISION[precision]), "dgr_plan_set_precision")

    def autotune(self, out, state, params: BssnParams, stream=None):
        """Times the tilings per block size for the current precision and keeps
        the fastest (dgr_plan_autotune); `out` is overwritten."""
        _check(load().dgr_plan_autotune(self._h, var_pointers(out), var_pointers(state), C.byref(params),
                                        _stream_ptr(stream)), "dgr_plan_autotune")

    def set_tiling(self, precision: str, size, tiling: str):
        """Pins the tiling ('wt', 'w' or 'full') of blocks of `size` (nx, ny, nz)."""
        _check(load().dgr_plan_set_tiling(self._h, PRECISION[precision], *[int(n) for n in size],
                                          TILINGS.index(tiling)), "dgr_plan_set_tiling")

    def tiling(self, precision: str = "fp64") -> dict:
        """{(nx, ny, nz): 'wt' | 'w' | 'full'} for the blocks without flagged faces."""
        L = load()
        n = L.dgr_plan_tiling(self._h, PRECISION[precision], 0, None, None)
        _check(min(n, 0), "dgr_plan_tiling")
        sizes = (C.c_uint32 * (3 * max(n, 1)))()
        sch = (C.c_int * max(n, 1))()
        L.dgr_plan_tiling(self._h, PRECISION[precision], n, sizes, sch)
        return {(sizes[3 * i], sizes[3 * i + 1], sizes[3 * i + 2]): TILINGS[sch[i]] for i in range(n)}

    def set_timing(self, enable: bool = True):
        _check(load().dgr_plan_set_timing(self._h, int(enable)), "dgr_plan_set_timing")

    TIMING_KINDS = ("derivs", "rhs", "face", "fused_wide", "fused_thin", "fused_face", "fused_full")

    def kernel_times(self):
        """{kind: (total ms, launches)} since the previous call."""
        nk = len(self.TIMING_KINDS)
        ms = (C.c_double * nk)()
        n = (C.c_int64 * nk)()
        _check(load().dgr_plan_kernel_times(self._h, ms, n), "dgr_plan_kernel_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(self.TIMING_KINDS)}

    def close(self):
        if getattr(self, "_h", None):
            load().dgr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
