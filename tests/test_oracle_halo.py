"""The oracle's padding refresh from same-level neighbours (the copy part of
the reference's unzip, bssnCtx.cpp:75) and RK4 with that refresh between
stages (rkBSSN.cpp:764-860). dendrolib, which holds the reference's unzip, is
absent, so these are pinned by what the geometry fixes: a smooth global
field is reproduced at every linked padding point, unlinked padding is left
alone, an empty table gives the frozen-padding RK4 bit for bit, and with the
refresh neighbouring blocks agree on their shared boundary nodes after a
step."""
import numpy as np
import pytest

import oracle as O
from dgr import workloads as W

BG = np.array([1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0], float)


def coords(t, b):
    n = int(t["sz"][b][0])
    h = (t["pmax"][b] - t["pmin"][b]) / (n - 1)
    ax = [t["pmin"][b][a] + h[a] * np.arange(n) for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def smooth_state(t, amp=1e-3, interior_only=False):
    """Flat space plus one smooth global mode per variable, sampled at every
    node (or only at interior nodes, NaN elsewhere)."""
    st = np.zeros((24, W.total_padded(t)))
    for b in range(len(t)):
        n = int(t["sz"][b][0])
        x, y, z = coords(t, b)
        inner = np.zeros((n, n, n), bool)
        inner[3:n - 3, 3:n - 3, 3:n - 3] = True
        off = int(t["offset"][b])
        for v in range(24):
            f = BG[v] + amp * np.sin(1.3 * x + 0.7 * (v % 5) * y - 0.9 * z + 0.3 * v)
            if interior_only:
                f = np.where(inner, f, np.nan)
            st[v, off:off + n ** 3] = f.ravel()
    return st


def test_lattice_links():
    t = W.lattice(shape=(3, 3, 2))
    nbr = W.neighbors(t)
    # a 3x3x2 lattice: sum over blocks of the neighbours inside the lattice
    per_axis = {3: [2, 3, 2], 2: [2, 2]}
    want = sum(a * b * c for a in per_axis[3] for b in per_axis[3] for c in per_axis[2]) - 18
    assert (nbr >= 0).sum() == want
    assert (nbr[:, 13] == -1).all()
    t2 = W.lattice(shape=(3, 3, 2), holes=[(1, 1, 0)], extra_level=True)
    nbr2 = W.neighbors(t2)
    assert (nbr2 >= 0).sum() == want - 2 * 17  # the hole had 17 neighbours
    assert (nbr2[-1] == -1).all() and not (nbr2 == len(t2) - 1).any()  # other level: no links
    # symmetry: nb is b's neighbour in d iff b is nb's in -d
    for b in range(len(t2)):
        for d in range(27):
            if nbr2[b, d] >= 0:
                assert nbr2[nbr2[b, d], 26 - d] == b


def test_fill_reproduces_a_smooth_field_and_leaves_unlinked_padding():
    t = W.lattice(shape=(3, 3, 2), holes=[(1, 1, 0)], extra_level=True)
    nbr = W.neighbors(t)
    st = smooth_state(t, interior_only=True)
    full = smooth_state(t)
    O.halo_fill(st, t, nbr)
    filled = 0
    for b in range(len(t)):
        n = int(t["sz"][b][0])
        off = int(t["offset"][b])
        got = st[:, off:off + n ** 3]
        want = full[:, off:off + n ** 3]
        m = ~np.isnan(got)
        assert np.abs(got[m] - want[m]).max() < 1e-14
        filled += int(m[0].sum()) - (n - 6) ** 3
        # which padding points were filled: exactly those whose direction has a neighbour
        g = m[0].reshape(n, n, n)
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    d = [(-1 if c < 3 else (1 if c >= n - 3 else 0)) for c in (i, j, k)]
                    if d == [0, 0, 0]:
                        continue
                    linked = nbr[b, (d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1)] >= 0
                    assert g[k, j, i] == linked
    assert filled > 0


def test_empty_table_gives_the_frozen_padding_step():
    t = W.uniform_13(3)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=2, amp=1e-3)
    p = O.default_params(ko_sigma=0.4)
    a, b = st.copy(), st.copy()
    O.rk4_step(a, t, 0.02, p, enforce_per_stage=True, nthreads=1)
    O.rk4_step_halo(b, t, np.full((len(t), 27), -1, np.int32), 0.02, p, enforce_per_stage=True, nthreads=1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("refresh", [True, False])
def test_refresh_keeps_shared_boundary_nodes_consistent(refresh):
    """Neighbours share their boundary nodes (b's i = n-4 is nb's i = 3). With
    the refresh both copies see the same stencil data at every stage and
    agree after a step to rounding; with the padding frozen they part by
    O(dt * change of the neighbour) - the check fails without the refresh."""
    t = W.lattice(shape=(2, 2, 2))
    nbr = W.neighbors(t)
    st = smooth_state(t, amp=1e-2)
    p = O.default_params(ko_sigma=0.4)
    if refresh:
        O.rk4_step_halo(st, t, nbr, 0.02, p, enforce_per_stage=True, nthreads=4)
    else:
        O.rk4_step(st, t, 0.02, p, enforce_per_stage=True, nthreads=4)
    n = 13
    worst = 0.0
    for b in range(len(t)):
        nb = nbr[b, 2 + 3 + 9]  # +x neighbour
        if nb < 0:
            continue
        A = st[:, t["offset"][b]:t["offset"][b] + n ** 3].reshape(24, n, n, n)
        B = st[:, t["offset"][nb]:t["offset"][nb] + n ** 3].reshape(24, n, n, n)
        worst = max(worst, np.abs(A[:, 3:10, 3:10, n - 4] - B[:, 3:10, 3:10, 3]).max())
    if refresh:
        assert worst < 1e-13
    else:
        assert worst > 1e-9


def test_morton_order_links_the_same_lattice():
    """The bench's Morton-ordered lattice is the same set of blocks: the same
    number of links, each block keeping its neighbours by position."""
    a = W.lattice(shape=(4, 4, 2), order="morton")
    b = W.lattice(shape=(4, 4, 2), order="shuffled")
    na, nb = W.neighbors(a), W.neighbors(b)
    assert (na >= 0).sum() == (nb >= 0).sum()
    pos_a = {tuple(np.round(p, 9)): i for i, p in enumerate(a["pmin"])}
    pos_b = {tuple(np.round(p, 9)): i for i, p in enumerate(b["pmin"])}
    assert pos_a.keys() == pos_b.keys()
    for key, i in pos_a.items():
        j = pos_b[key]
        for d in range(27):
            ka = tuple(np.round(a["pmin"][na[i, d]], 9)) if na[i, d] >= 0 else None
            kb = tuple(np.round(b["pmin"][nb[j, d]], 9)) if nb[j, d] >= 0 else None
            assert ka == kb
    # Z-order: the first 8 blocks form the first 2x2x2 cube
    h = (a["pmax"][0] - a["pmin"][0]) / 12
    cells = np.rint((a["pmin"][:8] - a["pmin"][0]) / (6 * h)).astype(int)
    assert {tuple(c) for c in cells} == {(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)}


def _distinct_nodes(t, n, h, origin, f):
    """Values of f at the distinct interior nodes of lattice t (keyed by their
    integer position), and the padded array of f over every block."""
    c = np.zeros(W.total_padded(t))
    nodes = {}
    for b in t:
        off = int(b["offset"])
        i = np.arange(n)
        z, y, x = np.meshgrid(*(b["pmin"][a] + i * h for a in (2, 1, 0)), indexing="ij")
        fv = f(x, y, z)
        c[off:off + n ** 3] = fv.ravel()
        for k in range(3, n - 3):
            for j in range(3, n - 3):
                for ii in range(3, n - 3):
                    key = tuple(np.rint((np.array([x[k, j, ii], y[k, j, ii], z[k, j, ii]]) - origin) / h).astype(int))
                    nodes.setdefault(key, (fv[k, j, ii], (x[k, j, ii], y[k, j, ii], z[k, j, ii])))
    return c, nodes


def test_norm_counts_shared_nodes_once():
    """computeConstraintL2Norm counts each zipped node once (grUtils.tcc:60-110):
    on lattices whose neighbours share their boundary nodes (a full 2x2x2 one,
    and ones with holes, where a node on an edge is shared by blocks that are
    diagonal neighbours only), the norm with the neighbour table equals a
    brute-force norm over the distinct nodes (by their integer lattice
    position), with and without the mask; the norm without the table counts
    a face node twice and differs."""
    h, n = 0.05, 13
    origin = np.array([-0.37, -0.21, -0.13])
    f = lambda x, y, z: np.cos(3.1 * x + 0.4) * np.sin(2.3 * y - 0.2) + 0.7 * z * z + 0.1
    for shape, holes in (((2, 2, 2), ()), ((2, 2, 2), ((0, 0, 0),)), ((3, 2, 2), ((1, 0, 0), (0, 1, 1)))):
        t = W.lattice(shape=shape, n=n, h=h, origin=tuple(origin), holes=holes)
        nbr = W.neighbors(t)
        c, nodes = _distinct_nodes(t, n, h, origin, f)
        if not holes:
            assert len(nodes) == (2 * (n - 7) + 1) ** 3  # distinct nodes of the lattice
        for bh, r in ((((30.0, 0, 0), (-30.0, 0, 0)), 5.0), (((3.0, 0, 0), (-30.0, 0, 0)), 3.0),
                      (((0.2, 0.1, 0.0), (0.0, 0.0, 0.3)), 0.12)):
            vals = [v for v, p in nodes.values()
                    if np.linalg.norm(np.subtract(p, bh[0])) >= r and np.linalg.norm(np.subtract(p, bh[1])) >= r]
            want = float(np.sqrt(np.mean(np.square(vals))))
            got = O.constraint_l2_norm(c, t, bh=bh, r1=r, r2=r, nbr=nbr)
            assert abs(got - want) <= 1e-14 * want, (shape, holes, got, want)
            loose = O.constraint_l2_norm(c, t, bh=bh, r1=r, r2=r)
            assert abs(loose - want) > 1e-6 * want
