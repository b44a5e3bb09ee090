"""Block tables for the synthetic workloads the path is measured and tested on.

A block table is a numpy structured array with the layout of the C-ABI's
``dgr_block`` (offset, padded size, face flags, coordinates of the padded
corner points), i.e. the arguments BSSN_GR/src/rhs.cpp:49-70 derives per
ot::Block before calling bssnrhs().

Workloads (points are interior points, summed over blocks):
  single_block_19   one 19^3 block, dx = 0.1, corner (0.25, 0.05, -0.15)
  uniform_13        N blocks of 13^3 (7^3 interior), dx = 0.1*(1 + b mod 3)
  puncture_octree   uneven 55^3/31^3/19^3/13^3 blocks placed by distance to two
                    punctures at (+-3,0,0), counts 150/1000/8000/60000 x scale
"""
from __future__ import annotations

import numpy as np

BLOCK_DTYPE = np.dtype([
    ("offset", "<u8"), ("sz", "<u4", (3,)), ("bflag", "<u4"),
    ("pmin", "<f8", (3,)), ("pmax", "<f8", (3,)),
])

PW = 3  # padding width (BSSN_GR/src/parameters.cpp:15-16)


def _table(sizes, pmins, hs, bflags=None, align=1):
    nb = len(sizes)
    t = np.zeros(nb, dtype=BLOCK_DTYPE)
    off = 0
    for b in range(nb):
        n = int(sizes[b])
        t["offset"][b] = off
        off += ((n ** 3 + align - 1) // align) * align
        t["sz"][b] = (n, n, n)
        t["bflag"][b] = 0 if bflags is None else int(bflags[b])
        pm = np.asarray(pmins[b], dtype=np.float64)
        t["pmin"][b] = pm
        t["pmax"][b] = pm + (n - 1) * float(hs[b])
    return t


def _morton(a: int, b: int, c: int) -> int:
    key = 0
    for bit in range(21):
        key |= ((a >> bit) & 1) << (3 * bit) | ((b >> bit) & 1) << (3 * bit + 1) | ((c >> bit) & 1) << (3 * bit + 2)
    return key


def lattice(shape=(3, 3, 2), n: int = 13, h: float = 0.05, origin=(-0.47, -0.31, -0.22), holes=(),
            extra_level: bool = False, order: str = "shuffled") -> np.ndarray:
    """Blocks of n^3 points on a regular lattice of shape (nbx, nby, nbz):
    neighbours share their boundary nodes, so block (a,b,c) starts at
    origin + (a,b,c) (n-7) h. `holes` lists lattice cells left out; with
    extra_level, one block of spacing h/2 is appended (it links to nothing).
    order: "shuffled" (the table order is not the lattice order) or "morton"
    (Z-order, as Dendro stores an octree's blocks along its space-filling
    curve)."""
    hole = set(holes)
    cells = [(a, b, c) for c in range(shape[2]) for b in range(shape[1]) for a in range(shape[0])
             if (a, b, c) not in hole]
    if order == "morton":
        cells.sort(key=lambda x: _morton(*x))
    elif order == "shuffled":
        perm = np.random.default_rng(1).permutation(len(cells))
        cells = [cells[i] for i in perm]
    else:
        raise ValueError(order)
    step = (n - 7) * h
    pmins = [tuple(origin[d] + cells[q][d] * step for d in range(3)) for q in range(len(cells))]
    hs = [h] * len(cells)
    if extra_level:
        pmins.append(tuple(origin[d] + shape[d] * step + 1.0 for d in range(3)))
        hs.append(h / 2)
    return _table([n] * len(pmins), pmins, hs)


def neighbors(t: np.ndarray, tol: float = 1e-6) -> np.ndarray:
    """Same-level neighbour table (nblocks x 27 int32, index
    (di+1) + 3(dj+1) + 9(dk+1), -1 for none, the centre entry -1): block nb
    is b's neighbour in direction d when both have the same sizes and spacing
    and pmin_nb = pmin_b + d (n-7) h within tol h. Brute force over pairs."""
    nb = len(t)
    out = np.full((nb, 27), -1, dtype=np.int32)
    sz = t["sz"].astype(np.int64)
    h = (t["pmax"] - t["pmin"]) / (sz - 1)
    for b in range(nb):
        for c in range(nb):
            if c == b or not np.array_equal(sz[b], sz[c]) or np.abs(h[b] - h[c]).max() > tol * h[b].min():
                continue
            d = (t["pmin"][c] - t["pmin"][b]) / ((sz[b] - 7) * h[b])
            r = np.rint(d)
            if np.abs(d - r).max() * (sz[b] - 7).max() > tol or np.abs(r).max() > 1 or not r.any():
                continue
            di, dj, dk = (int(x) for x in r)
            out[b, (di + 1) + 3 * (dj + 1) + 9 * (dk + 1)] = c
    return out


def two_level_lattice(h: float = 0.05, origin=(-0.41, -0.29, -0.17), fine_cell=(1, 1, 1)) -> np.ndarray:
    """A 2x2x2 lattice of 13^3 blocks of spacing 2h (interiors 12h wide,
    neighbours sharing boundary nodes) with one cell refined: in place of the
    coarse block at `fine_cell`, 2x2x2 blocks of 13^3 of spacing h (interiors
    6h wide). Faces between the refined cell and its coarse neighbours are
    level jumps; blocks sharing a level and a size link as neighbours."""
    o = np.asarray(origin, dtype=np.float64)
    sizes, pmins, hs = [], [], []
    for c in range(2):
        for b in range(2):
            for a in range(2):
                cell = np.array([a, b, c], dtype=np.float64)
                if (a, b, c) == tuple(fine_cell):
                    for r in range(2):
                        for q in range(2):
                            for p in range(2):
                                sizes.append(13)
                                hs.append(h)
                                pmins.append(o + 12 * h * cell + 6 * h * np.array([p, q, r]) - PW * h)
                else:
                    sizes.append(13)
                    hs.append(2 * h)
                    pmins.append(o + 12 * h * cell - PW * 2 * h)
    return _table(sizes, pmins, hs)


def mixed_size_lattice(h: float = 0.05, origin=(-0.43, -0.31, -0.19)) -> np.ndarray:
    """Blocks of one spacing h but two sizes: a 19^3 block (interior 12h wide)
    at the origin, eight 13^3 blocks (interiors 6h wide) filling the 12h cube
    next to it along +x, and one more 19^3 block beyond them; every 19^3 face
    towards the 13^3 blocks meets four of them."""
    o = np.asarray(origin, dtype=np.float64)
    sizes, pmins, hs = [19], [o - PW * h], [h]
    for r in range(2):
        for q in range(2):
            for p in range(2):
                sizes.append(13)
                hs.append(h)
                pmins.append(o + np.array([12 * h, 0.0, 0.0]) + 6 * h * np.array([p, q, r]) - PW * h)
    sizes.append(19)
    hs.append(h)
    pmins.append(o + np.array([24 * h, 0.0, 0.0]) - PW * h)
    return _table(sizes, pmins, hs)


def refined_octree(base: int = 4, max_level: int = 9, shell: float = 4.0, half_width: float = 16.0,
                   punctures=((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0)), boundary_faces: bool = True) -> np.ndarray:
    """A 2:1-balanced octree over [-half_width, half_width]^3 refined toward
    the punctures, one 13^3 block per leaf octant (its 7^3 interior nodes are
    the octant's order-6 element; neighbours share their boundary nodes).
    Every octant of level `base` is a leaf to start with; an octant of level l
    < max_level is split into 8 when its distance to the nearest puncture is
    below shell x its edge; then octants are split until every leaf's face,
    edge and corner neighbours are at most one level coarser (2:1 balance).
    Padding across a level jump lies in a coarser or finer neighbour: the
    workload of the full padding refresh. Blocks are in Morton order of their
    octants (Dendro's block order along its space-filling curve). With
    boundary_faces, a block's faces on the domain boundary are flagged
    (OCT_DIR_* bits), as Dendro flags them, so the radiative boundary
    condition applies there."""
    import itertools
    P = np.asarray(punctures, dtype=np.float64)
    size = lambda lev: 2.0 * half_width / 2 ** lev
    leaves = {(base, i, j, k) for i in range(2 ** base) for j in range(2 ** base) for k in range(2 ** base)}

    def dist(key):
        lev, i, j, k = key
        lo = np.array([i, j, k], dtype=np.float64) * size(lev) - half_width
        d = np.maximum(0.0, np.maximum(lo - P, P - (lo + size(lev))))
        return float(np.sqrt((d * d).sum(axis=1)).min())

    def split(key):
        leaves.remove(key)
        lev, i, j, k = key
        for a, b, c in itertools.product((0, 1), repeat=3):
            leaves.add((lev + 1, 2 * i + a, 2 * j + b, 2 * k + c))

    for lev in range(base, max_level):
        for key in [q for q in leaves if q[0] == lev and dist(q) < shell * size(lev)]:
            split(key)
    dirs = [d for d in itertools.product((-1, 0, 1), repeat=3) if d != (0, 0, 0)]
    while True:
        mark = set()
        for lev, i, j, k in leaves:
            n = 2 ** lev
            for dx, dy, dz in dirs:
                x, y, z = i + dx, j + dy, k + dz
                if not (0 <= x < n and 0 <= y < n and 0 <= z < n):
                    continue
                for up in range(lev + 1):
                    key = (lev - up, x >> up, y >> up, z >> up)
                    if key in leaves:
                        if up >= 2:
                            mark.add(key)
                        break
        if not mark:
            break
        for key in mark:
            split(key)
    top = max(q[0] for q in leaves)
    order = sorted(leaves, key=lambda q: _morton(q[1] << (top - q[0]), q[2] << (top - q[0]), q[3] << (top - q[0])))
    sizes, pmins, hs, flags = [], [], [], []
    for lev, i, j, k in order:
        h = size(lev) / 6.0
        sizes.append(13)
        hs.append(h)
        pmins.append(np.array([i, j, k], dtype=np.float64) * size(lev) - half_width - PW * h)
        f = 0
        if boundary_faces:
            last = 2 ** lev - 1
            for a, c in enumerate((i, j, k)):
                f |= (1 << (2 * a)) if c == 0 else 0
                f |= (1 << (2 * a + 1)) if c == last else 0
        flags.append(f)
    return _table(sizes, pmins, hs, flags)


def total_padded(t: np.ndarray) -> int:
    """Length each variable array needs (max offset + block size)."""
    if len(t) == 0:
        return 0
    ends = t["offset"].astype(np.int64) + np.prod(t["sz"].astype(np.int64), axis=1)
    return int(ends.max())


def interior_points(t: np.ndarray) -> int:
    return int(np.prod(t["sz"].astype(np.int64) - 2 * PW, axis=1).sum())


def single_block_19(bflag: int = 0) -> np.ndarray:
    return _table([19], [(0.25, 0.05, -0.15)], [0.1], [bflag])


def uniform_13(nblocks: int = 4096, align: int = 1) -> np.ndarray:
    b = np.arange(nblocks)
    pmins = np.stack([0.25 + 4.0 * (b % 16), 0.05 + 4.0 * ((b // 16) % 16),
                      -0.15 + 4.0 * (b // 256)], axis=1)
    hs = 0.1 * (1 + b % 3)
    return _table([13] * nblocks, pmins, hs, align=align)


PUNCTURE_CLASSES = ((55, 150, 0.02), (31, 1000, 0.04), (19, 8000, 0.08), (13, 60000, 0.16))


def puncture_octree(scale: float = 0.42, align: int = 1) -> np.ndarray:
    """Uneven blocks: unit cells ordered by distance to the nearest puncture get
    55^3 (nearest), then 31^3, 19^3, 13^3 blocks; dx halves toward the punctures so
    every block spans ~0.96 in each direction. Blocks are emitted in cell order
    (sizes interleaved, as an SFC-ordered octree would hand them over)."""
    counts = [max(1, int(round(c * scale))) for _, c, _ in PUNCTURE_CLASSES]
    total = sum(counts)
    R = int(np.ceil((total * 3.0) ** (1.0 / 3.0) / 2.0)) + 4
    ax = np.arange(-R, R)
    cz, cy, cx = np.meshgrid(ax, ax, ax, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    centers = cells + 0.5
    d1 = np.linalg.norm(centers - np.array([3.0, 0.0, 0.0]), axis=1)
    d2 = np.linalg.norm(centers - np.array([-3.0, 0.0, 0.0]), axis=1)
    order = np.argsort(np.minimum(d1, d2), kind="stable")[:total]
    cls = np.repeat(np.arange(4), counts)
    chosen = cells[order]
    # emit in cell (z, y, x) order so sizes interleave
    emit = np.lexsort((chosen[:, 0], chosen[:, 1], chosen[:, 2]))
    shift = np.array([0.0123, 0.0234, 0.0345])
    sizes, pmins, hs = [], [], []
    for e in emit:
        n, _, h = PUNCTURE_CLASSES[cls[e]]
        sizes.append(n)
        hs.append(h)
        pmins.append(chosen[e] + shift + 0.02 - PW * h)
    return _table(sizes, pmins, hs, align=align)


def stage_bytes(t: np.ndarray) -> float:
    """Algorithmic HBM bytes of one fused RK stage over all blocks: each padded
    FP64 state read once (192 n^3 per block), per interior point the RHS
    produced once (192) and the stage update (read u0, read and write the
    accumulator, write the next stage: 768)."""
    n3 = np.prod(t["sz"].astype(np.float64), axis=1)
    m3 = np.prod(t["sz"].astype(np.float64) - 2 * PW, axis=1)
    return float((192.0 * n3 + 960.0 * m3).sum())


def fused_kind_bytes(t: np.ndarray, precision: str = "fp64") -> dict:
    """stage_bytes split over the fused launches in proportion to the interior
    points each launch covers within every block."""
    out = {k: 0.0 for k in fused_kind_points(t[:0], precision)}
    for b in range(len(t)):
        share = fused_kind_points(t[b:b + 1], precision)
        m = sum(share.values())
        bb = stage_bytes(t[b:b + 1])
        for k, v in share.items():
            out[k] += bb * v / m
    return out


def fused_kind_points(t: np.ndarray, precision: str = "fp64") -> dict:
    """Interior points each launch of the fused path covers, by tile kind, with
    the plan's tiling rule: blocks with flagged faces -> 'fused_face'; in FP64
    other blocks are cut along z into parts of 5 points, a remainder of 1-2
    points forming 'fused_thin' tiles and everything else 'fused_wide' tiles;
    in mixed precision other blocks go to 'fused_full' (7 x 7 x 7 tiles)."""
    if precision == "fp64":
        pts = {"fused_wide": 0, "fused_thin": 0, "fused_face": 0}
    else:
        pts = {"fused_full": 0, "fused_face": 0}
    for b in range(len(t)):
        m = (t["sz"][b].astype(np.int64) - 2 * PW)
        xy = int(m[0] * m[1])
        if t["bflag"][b]:
            pts["fused_face"] += xy * int(m[2])
            continue
        if precision != "fp64":
            pts["fused_full"] += xy * int(m[2])
            continue
        r = int(m[2]) % 5
        thin = r if r in (1, 2) else 0
        pts["fused_thin"] += xy * thin
        pts["fused_wide"] += xy * (int(m[2]) - thin)
    return pts
