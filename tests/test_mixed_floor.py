"""What FP32 arithmetic can reach on the perturbed-Minkowski block, measured
on the CPU with the oracle's stencils and the host build of the generated
equations: the floor under any mixed-precision design whose derivatives or
linear terms pass through FP32.

On this data every non-background variable changes by up to ~9x its own size
in one RK4 step (dt = 0.025 with modes of k h up to ~2), so the state error
per variable is the right-hand side's relative error times ~9. Rounding only
the 138 derivative values to FP32 (everything else FP64) already moves the
per-variable state estimate dt |dRHS| / |u| above 1e-7; rounding only the
field values moves it above 1e-7 too. A per-variable bound of 1e-8 against
FP64 is therefore out of reach for FP32 on these inputs, while the error
against the state's scale (the background fields are O(1)) stays ~1e-8."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from dgr import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "dendro-gr_amd", "lib", "libdgr_algebra_host.so")
D2V = [0, 1, 6, 7, 8, 12, 13, 14, 15, 16, 17]  # alpha, chi, beta0..2, gt0..5


def point_inputs(st, blk, eta=2.0):
    """(points, 163) inputs of the pointwise equations at the block's interior."""
    n = int(blk["sz"][0])
    h = float((blk["pmax"][0] - blk["pmin"][0]) / (n - 1))
    sz = (n, n, n)
    f = st.reshape(24, n, n, n)
    d1 = np.zeros((24, 3, n, n, n))
    d2 = np.zeros((11, 6, n, n, n))
    for v in range(24):
        for a, nm in enumerate(("deriv644_x", "deriv644_y", "deriv644_z")):
            d1[v, a] = O.stencil(nm, f[v], h, sz, 0)
    for m, v in enumerate(D2V):
        d2[m, 0] = O.stencil("deriv644_xx", f[v], h, sz, 0)
        d2[m, 1] = O.stencil("deriv644_y", np.nan_to_num(d1[v, 0]), h, sz, 0)
        d2[m, 2] = O.stencil("deriv644_z", np.nan_to_num(d1[v, 0]), h, sz, 0)
        d2[m, 3] = O.stencil("deriv644_yy", f[v], h, sz, 0)
        d2[m, 4] = O.stencil("deriv644_z", np.nan_to_num(d1[v, 1]), h, sz, 0)
        d2[m, 5] = O.stencil("deriv644_zz", f[v], h, sz, 0)
    s = slice(3, n - 3)
    cols = [f[:, s, s, s].reshape(24, -1), d1[:, :, s, s, s].reshape(72, -1),
            d2[:, :, s, s, s].reshape(66, -1), np.full((1, (n - 6) ** 3), eta)]
    return np.ascontiguousarray(np.concatenate(cols).T)


def algebra(inp):
    L = C.CDLL(HOST)
    dp = C.POINTER(C.c_double)
    out = np.zeros((inp.shape[0], 24))
    lam, lamf = np.ones(4), np.array([1.0, 0.0])
    L.dgr_host_bssn_points(inp.ctypes.data_as(dp), C.c_int64(inp.shape[0]), lam.ctypes.data_as(dp),
                           lamf.ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


@pytest.fixture(scope="module")
def block():
    if not os.path.exists(HOST):
        pytest.skip("host build of the equations missing (dendro-gr_amd/build.py)")
    t = W.single_block_19()
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    inp = point_inputs(st, t[0])
    return inp, algebra(inp)


def state_error_estimate(r, r64, umax, dt=0.025):
    """Per variable, dt max|RHS - RHS_64| / max|u|: the state error one step adds."""
    return (np.abs(r - r64).max(axis=0) * dt / umax).max()


def test_derivatives_rounded_to_fp32_exceed_1e8_per_variable(block):
    inp, r64 = block
    umax = np.abs(inp[:, :24]).max(axis=0)
    x = inp.copy()
    x[:, 24:162] = x[:, 24:162].astype(np.float32).astype(np.float64)
    est = state_error_estimate(algebra(x), r64, umax)
    assert est > 1e-7, est
    # the same error against the state's scale is three orders smaller
    assert (np.abs(algebra(x) - r64).max() * 0.025 / np.abs(inp[:, :24]).max()) < 3e-9


def test_fields_rounded_to_fp32_exceed_1e8_per_variable(block):
    inp, r64 = block
    umax = np.abs(inp[:, :24]).max(axis=0)
    x = inp.copy()
    x[:, :24] = x[:, :24].astype(np.float32).astype(np.float64)
    assert state_error_estimate(algebra(x), r64, umax) > 1e-7


def test_variables_change_by_several_times_their_size_per_step(block):
    inp, r64 = block
    umax = np.abs(inp[:, :24]).max(axis=0)
    growth = 0.025 * np.abs(r64).max(axis=0) / umax
    assert growth.max() > 5.0  # K, Gt, B, At: dt |RHS| / |u| up to ~9


def test_puncture_noise_floor_exceeds_1e8_per_variable():
    """The octree data (Brill-Lindquist background, 1e-6 white noise on K,
    Gt, beta, B, At): the noise is not smooth, so no offset or low-order fit
    subtracted per tile shrinks it. Rounding to FP32 only the derivative values
    of the variables other than alpha and chi (which the mixed kernels stage
    and differentiate in FP64) already moves the state one step later by more
    than 1e-8 of its size per variable (Gt, B and the off-diagonal gt, ~4e-8:
    their right-hand sides take derivatives of the noise)."""
    t = W.puncture_octree(0.42)
    b = int(np.nonzero(t["sz"][:, 0] == 19)[0][0])
    one = t[b:b + 1].copy()
    one["offset"][0] = 0
    st = np.zeros((24, W.total_padded(one)))
    O.init_brill_lindquist(st, one, seed=5, noise=1e-6)
    h = float((one["pmax"][0, 0] - one["pmin"][0, 0]) / 18)
    inp = point_inputs(st, one[0], eta=2.0)
    r64 = algebra(inp)
    dt = 0.25 * h
    x = inp.copy()
    keep = set(range(24, 30)) | set(range(96, 108))  # alpha's and chi's first and second derivatives
    cols = [c for c in range(24, 162) if c not in keep]
    x[:, cols] = x[:, cols].astype(np.float32).astype(np.float64)
    # per variable against the size of the state after one (Euler) step, for
    # the variables that are not identically zero
    after = np.abs(inp[:, :24] + dt * r64).max(axis=0)
    live = after > 0
    est = (np.abs(algebra(x) - r64).max(axis=0)[live] * dt / after[live]).max()
    assert est > 1e-8, est
