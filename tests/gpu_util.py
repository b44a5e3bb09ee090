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
                            eps_cako_other=p.eps_cako_other)


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
