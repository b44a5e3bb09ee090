"""Synthetic block data generated on the device (for benchmarks at full size).

Same recipe as the CPU generator the tests use: Minkowski plus amp * sum of 4
sine modes per (block, variable), mode parameters drawn from a splitmix64
stream keyed by (seed, block, variable); Brill-Lindquist puncture data for the
uneven octree. Values agree with the CPU generator up to the last bits of
sin/pow.
"""
from __future__ import annotations

import numpy as np

from .workloads import total_padded

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_K1 = np.uint64(0x632BE59BD9B4E019)
_K2 = np.uint64(0x8CB92BA72F3D8DD7)
BACKGROUND = np.array([1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0],
                      dtype=np.float64)


def _splitmix(state: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    with np.errstate(over="ignore"):
        state = state + _GOLD
        z = state.copy()
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return state, z


def sine_modes(nblocks: int, seed: int, lengths: np.ndarray):
    """Per (block, var, mode): amplitude, wave vector, phase. lengths: (nb, 3)."""
    b = np.arange(nblocks, dtype=np.uint64)[:, None]
    v = np.arange(24, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        st = np.uint64(seed) + _K1 * (b + np.uint64(1)) + _K2 * (v + np.uint64(1))
    twopi = 6.283185307179586476925286766559
    A = np.zeros((nblocks, 24, 4))
    kv = np.zeros((nblocks, 24, 4, 3))
    ph = np.zeros((nblocks, 24, 4))
    for m in range(4):
        st, z = _splitmix(st)
        A[:, :, m] = 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0
        for a in range(3):
            st, z = _splitmix(st)
            kv[:, :, m, a] = (1 + (z & np.uint64(3)).astype(np.float64)) * twopi / lengths[:, None, a]
        st, z = _splitmix(st)
        ph[:, :, m] = twopi * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)
    return A, kv, ph


def _coords(blocks, idx, n, torch, dev):
    pmin = torch.from_numpy(blocks["pmin"][idx]).to(dev)
    h = torch.from_numpy((blocks["pmax"][idx] - blocks["pmin"][idx]) / (n - 1)).to(dev)
    i = torch.arange(n, device=dev, dtype=torch.float64)
    xs = [pmin[:, a, None] + i[None, :] * h[:, a, None] for a in range(3)]
    return xs


def _scatter(state, blocks, idx, n, vals, torch, dev):
    # vals: (nbg, 24, n^3) -> state[:, offset_b : offset_b + n^3]
    offs = torch.from_numpy(blocks["offset"][idx].astype(np.int64)).to(dev)
    cols = (offs[:, None] + torch.arange(n ** 3, device=dev)[None, :]).reshape(-1)
    state[:, cols] = vals.permute(1, 0, 2).reshape(24, -1)


def perturbed_minkowski(blocks: np.ndarray, seed: int = 1, amp: float = 1e-3, device="cuda",
                        group: int = 512):
    import torch
    dev = torch.device(device)
    state = torch.zeros((24, total_padded(blocks)), dtype=torch.float64, device=dev)
    L = blocks["pmax"] - blocks["pmin"]
    A, kv, ph = sine_modes(len(blocks), seed, L)
    bg = torch.from_numpy(BACKGROUND).to(dev)
    sizes = blocks["sz"][:, 0].astype(np.int64)
    for n in np.unique(sizes):
        all_idx = np.nonzero(sizes == n)[0]
        for g0 in range(0, len(all_idx), group):
            idx = all_idx[g0:g0 + group]
            x, y, z = _coords(blocks, idx, int(n), torch, dev)
            s = torch.zeros((len(idx), 24, n, n, n), dtype=torch.float64, device=dev)
            for m in range(4):
                k = torch.from_numpy(kv[idx, :, m, :]).to(dev)      # (nbg, 24, 3)
                a = torch.from_numpy(A[idx, :, m]).to(dev)          # (nbg, 24)
                p = torch.from_numpy(ph[idx, :, m]).to(dev)
                arg = (k[:, :, 0, None, None, None] * x[:, None, None, None, :]
                       + k[:, :, 1, None, None, None] * y[:, None, None, :, None]
                       + k[:, :, 2, None, None, None] * z[:, None, :, None, None]
                       + p[:, :, None, None, None])
                s += a[:, :, None, None, None] * torch.sin(arg)
            vals = bg[None, :, None] + amp * s.reshape(len(idx), 24, -1)
            _scatter(state, blocks, idx, int(n), vals, torch, dev)
    return state


def brill_lindquist(blocks: np.ndarray, seed: int = 1, noise: float = 1e-6, chi_floor: float = 0.1,
                    device="cuda", group: int = 256):
    """Two punctures m1 = m2 = 0.5 at (+-3, 0, 0): alpha = psi^-2, chi = psi^-4 (each
    floored), gt = delta; K, Gt, beta, B, At get noise * U[-1,1] (torch's
    generator seeded with `seed`)."""
    import torch
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    state = torch.zeros((24, total_padded(blocks)), dtype=torch.float64, device=dev)
    sizes = blocks["sz"][:, 0].astype(np.int64)
    noisy = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 18, 19, 20, 21, 22, 23]
    for n in np.unique(sizes):
        all_idx = np.nonzero(sizes == n)[0]
        for g0 in range(0, len(all_idx), group):
            idx = all_idx[g0:g0 + group]
            x, y, z = _coords(blocks, idx, int(n), torch, dev)
            X = x[:, None, None, :]
            Y = y[:, None, :, None]
            Z = z[:, :, None, None]
            r1 = torch.sqrt((X - 3.0) ** 2 + Y ** 2 + Z ** 2)
            r2 = torch.sqrt((X + 3.0) ** 2 + Y ** 2 + Z ** 2)
            psi = 1.0 + 0.5 / (2.0 * r1) + 0.5 / (2.0 * r2)
            vals = torch.zeros((len(idx), 24, n ** 3), dtype=torch.float64, device=dev)
            vals[:, 0] = torch.clamp(1.0 / (psi * psi), min=chi_floor).reshape(len(idx), -1)
            vals[:, 1] = torch.clamp(1.0 / psi ** 4, min=chi_floor).reshape(len(idx), -1)
            for c in (12, 15, 17):
                vals[:, c] = 1.0
            r = torch.rand((len(idx), len(noisy), n ** 3), generator=gen, dtype=torch.float64, device=dev)
            vals[:, noisy] = noise * (2.0 * r - 1.0)
            _scatter(state, blocks, idx, int(n), vals, torch, dev)
    return state
