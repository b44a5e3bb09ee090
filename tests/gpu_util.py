"""Shared helpers of the GPU tests (not collected by pytest)."""
import sys

import numpy as np

import oracle as O
import dgr
from dgr import workloads as W


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def host_state(blocks, init="minkowski", seed=1, amp=1e-3, noise=1e-6):
    st = np.zeros((24, W.total_padded(blocks)))
    if init == "minkowski":
        O.init_perturbed_minkowski(st, blocks, seed=seed, amp=amp)
    elif init == "puncture":
        O.init_brill_lindquist(st, blocks, seed=seed, noise=noise)
    else:
        raise ValueError(init)
    return st


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def gpu_rhs(blocks, st, params, fill=np.nan):
    import torch
    plan = dgr.Plan(blocks, st.shape[1])
    u = to_dev(st)
    out = torch.full_like(u, fill)
    plan.rhs(out, u, params)
    torch.cuda.synchronize()
    assert plan.status() == 0
    res = out.cpu().numpy()
    plan.close()
    return res


def oracle_params(p):
    """Oracle parameter struct with the same values as a dgr.BssnParams."""
    return O.default_params(lambda_=list(p.lambda_), lambda_f=list(p.lambda_f),
                            eta_central=p.eta_central, eta_outer=p.eta_outer,
                            eta_width=p.eta_width, ko_sigma=p.ko_sigma, chi_floor=p.chi_floor,
                            cako_enabled=p.cako_enabled, eps_cako_gauge=p.eps_cako_gauge,
                            eps_cako_other=p.eps_cako_other, eta_func=p.eta_func, eta_r0=p.eta_r0,
                            eta_power=list(p.eta_power), rochester=p.rochester, ssl=p.ssl,
                            xi=list(p.xi), time=p.time)


def interior_values(arr, blocks):
    """(24, total interior points) from a (24, szpdof) array."""
    parts = []
    for b in blocks:
        nx, ny, nz = (int(s) for s in b["sz"])
        off = int(b["offset"])
        blk = arr[:, off:off + nx * ny * nz].reshape(24, nz, ny, nx)
        parts.append(blk[:, 3:nz - 3, 3:ny - 3, 3:nx - 3].reshape(24, -1))
    return np.concatenate(parts, axis=1)


def padding_mask(blocks, szpdof):
    m = np.zeros(szpdof, dtype=bool)
    for b in blocks:
        nx, ny, nz = (int(s) for s in b["sz"])
        off = int(b["offset"])
        blk = np.ones((nz, ny, nx), dtype=bool)
        blk[3:nz - 3, 3:ny - 3, 3:nx - 3] = False
        m[off:off + nx * ny * nz] = blk.ravel()
    return m


def normwise(a, b):
    """Per-variable max|a-b| / max|b| over axis 1."""
    den = np.abs(b).max(axis=1)
    den = np.where(den == 0, 1.0, den)
    return np.abs(a - b).max(axis=1) / den


def harness_blocks(nblocks, bflag=0, n=13):
    """The block set of the reference's GPU test harness (cuda/src/rhs_test.cu:
    28-66): nblocks blocks of n^3 at (-0.5, -0.5, -0.5), spacing
    1e-2 / n / (b % 3 + 1), one face flag word for all."""
    t = np.zeros(nblocks, dtype=W.BLOCK_DTYPE)
    for b in range(nblocks):
        h = 1.0e-2 / n / (b % 3 + 1)
        t["offset"][b] = b * n ** 3
        t["sz"][b] = (n, n, n)
        t["bflag"][b] = bflag
        t["pmin"][b] = (-0.5, -0.5, -0.5)
        t["pmax"][b] = (-0.5 + (n - 1) * h,) * 3
    return t


def fake_initial_data(blocks):
    """The reference test harness's input (grUtils.cpp:1368-1620,
    bssn::fake_initial_data, evaluated at each padded point's coordinates as
    test_utils.h:28-41 does), restated in numpy: closed-form fields of every
    variable, then gt scaled to unit determinant (flat where it is negative)
    and At made trace-free, as that function ends."""
    st = np.zeros((24, W.total_padded(blocks)))
    for b in range(len(blocks)):
        n = [int(s) for s in blocks["sz"][b]]
        h = (blocks["pmax"][b] - blocks["pmin"][b]) / (np.array(n) - 1)
        k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
        x = (blocks["pmin"][b][0] + i * h[0]).ravel()
        y = (blocks["pmin"][b][1] + j * h[1]).ravel()
        z = (blocks["pmin"][b][2] + k * h[2]).ravel()
        f1, f2 = 31.0 / 17.0, 37.0 / 11.0
        sin, cos, exp = np.sin, np.cos, np.exp
        u = np.zeros((24, x.size))
        u[0] = 1.0 - 0.25 * sin(f1 * x)
        u[6] = 4.0 / 17.0 * sin(x) * cos(z)
        u[7] = np.pi / 5.0 * cos(y) * sin(z + x)
        u[8] = 4.0 / 17.0 * sin(f2 * x) * sin(z)
        u[9] = 31.0 * x * cos(f1 * z + y)
        u[10] = 7.0 * y * sin(f1 * x + y) + 3.0 * cos(z)
        u[11] = 5.0 * z * cos(f1 * x + y) + 7.0 * sin(z + y + x) + 1.0
        u[3] = 5.0 * cos(x) / (10.0 * sin(x + z) + 26.0 - 1.0 * cos(x * z) * cos(x))
        u[4] = -5.0 * sin(y) / (25.0 + 10.0 * cos(y + z) + cos(y) * cos(y * z))
        u[5] = -5.0 * sin(z) / (25.0 + 10.0 * cos(y + x) + cos(y * x) * cos(z))
        e = exp(-4.0 * cos(x) * sin(y))
        u[1] = 1.0 + e
        u[12] = 1.00 + 0.2 * sin(x + z) * cos(y)
        u[15] = 1.00 + 0.2 * cos(y) * cos(z + x)
        u[17] = 1.00 / (u[12] + u[15])
        u[13] = 0.07 * (2.0 + cos(x * x + y * y))
        u[14] = 0.1 * (3.0 + sin(z) * cos(x))
        u[16] = 0.15 * (1.751 - sin(x * x) * cos(y) * cos(z))
        br = (5.0 * e / (5.0 + sin(x)) * cos(x) + 5.0 * e / (5.0 + cos(y)) * cos(y)
              + 0.4 * (25.0 + 5.0 * cos(y) + 5.0 * sin(x) + sin(x) * cos(y)) * e * cos(z))
        u[2] = br * 0.01234
        br4 = (5.0 * e / (5.0 + sin(x)) * cos(x) + 5.0 * e / (5.0 + cos(y)) * cos(y)
               + 0.04 * (25.0 + 5.0 * cos(y) + 5.0 * sin(x) + sin(x) * cos(y)) * e * cos(z))
        ie = exp(4.0 * cos(x) * sin(y))
        u[18] = e * (cos(x) - 0.3333333333 * ie * (1.0 + 0.2 * sin(x)) * br4)
        u[19] = 1.0 + x * z / (0.1 + x * x + y * y + z * z)
        u[20] = 1.3 - x * y / (3.0 + x * x + 2.0 * y * y + z * z) * (x * x + z * z)
        u[21] = e * (cos(y) - 0.33333333330 * ie * (1 + 0.2 * cos(y)) * br4)
        u[22] = -1.0 + y * z / (1.0 + 3.0 * x * x + y * y + z * z)
        u[23] = e * (cos(z) - 0.3333333333 * ie / (1 + 0.2 * sin(x)) / (1 + 0.2 * cos(y)) * br4)
        # unit determinant and trace-free At (the end of fake_initial_data)
        sym = ((0, 1, 2), (1, 3, 4), (2, 4, 5))
        g = np.stack([np.stack([u[12 + sym[a][c]] for c in range(3)]) for a in range(3)])
        A = np.stack([np.stack([u[18 + sym[a][c]] for c in range(3)]) for a in range(3)])
        det = np.linalg.det(np.moveaxis(g, -1, 0))
        flat = det < 0.0
        g[:, :, flat] = np.eye(3)[:, :, None]
        det[flat] = 1.0
        g = g / det ** (1.0 / 3.0)
        gi = np.moveaxis(np.linalg.inv(np.moveaxis(g, -1, 0)), 0, -1)
        tr = np.einsum("abp,abp->p", A, gi) / 3.0
        A = A - tr * g
        for a in range(3):
            for c in range(a, 3):
                u[12 + sym[a][c]] = g[a, c]
                u[18 + sym[a][c]] = A[a, c]
        o = int(blocks["offset"][b])
        st[:, o:o + x.size] = u
    return st
