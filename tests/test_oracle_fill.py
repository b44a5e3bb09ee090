"""The oracle's general padding fill (the copy and interpolation parts of the
reference's unzip, BSSN_GR/src/bssnCtx.cpp:75; the octree code is in
dendrolib, absent): pinned by what the geometry and the mathematics fix.

  * the order-6 Lagrange weights at the half-node positions reproduce
    polynomials of degree <= 6 and sum to one; at the nodes they select;
  * on a two-level lattice (one coarse cell refined), a field that is a
    polynomial of degree <= 6 along each axis is reproduced at every padding
    point that lies inside another block: same-level copies and injections
    from the finer blocks exactly, interpolations from the coarser blocks to
    rounding; a degree-7 term is not reproduced there (the interpolation is
    order 6 and is used);
  * the filled set is exactly the padding inside other blocks' interiors
    (domain-boundary padding keeps its value);
  * blocks of one spacing and different sizes copy from each other.
"""
import numpy as np
import pytest

import oracle as O
from dgr import workloads as W


def coords(b):
    n = int(b["sz"][0])
    h = (b["pmax"] - b["pmin"]) / (n - 1)
    i = np.arange(n)
    z, y, x = np.meshgrid(*(b["pmin"][a] + i * h[a] for a in (2, 1, 0)), indexing="ij")
    return x, y, z


def poly6(x, y, z):
    px = 1.0 + 0.3 * x - 0.2 * x ** 2 + 0.05 * x ** 3 - 0.04 * x ** 5 + 0.01 * x ** 6
    py = 0.7 - 0.1 * y + 0.25 * y ** 2 + 0.02 * y ** 4 - 0.03 * y ** 6
    pz = 1.1 + 0.4 * z - 0.06 * z ** 3 + 0.02 * z ** 6
    return px * py * pz + x ** 4 * y ** 2 * z - 0.5 * x * y ** 6


def state_from(blocks, f, pad=np.nan):
    """Every variable: f at the interior points of each block, `pad` at its
    padding points."""
    st = np.full((24, W.total_padded(blocks)), pad)
    for b in blocks:
        n = int(b["sz"][0])
        x, y, z = coords(b)
        v = f(x, y, z)
        v[:3], v[-3:], v[:, :3], v[:, -3:], v[:, :, :3], v[:, :, -3:] = (pad,) * 6
        off = int(b["offset"])
        st[:, off:off + n ** 3] = v.ravel()[None, :] * (1.0 + 0.01 * np.arange(24))[:, None]
    return st


def inside_other(blocks, b):
    """Padding points of block b that lie (on its spacing's lattice) inside
    another block's interior: the expected filled set."""
    n = int(blocks["sz"][b][0])
    x, y, z = coords(blocks[b])
    pad = np.ones((n, n, n), dtype=bool)
    pad[3:-3, 3:-3, 3:-3] = False
    got = np.zeros_like(pad)
    for c, nb in enumerate(blocks):
        if c == b:
            continue
        m = int(nb["sz"][0])
        h = (nb["pmax"][0] - nb["pmin"][0]) / (m - 1)
        lo = nb["pmin"] + 3 * h - 1e-9
        hi = nb["pmin"] + (m - 4) * h + 1e-9
        got |= (x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1]) & (z >= lo[2]) & (z <= hi[2])
    return pad & got


@pytest.mark.parametrize("code", range(13))
def test_lagrange6_weights(code):
    w = O.lagrange6_half(code)
    t = code / 2
    assert abs(w.sum() - 1.0) < 1e-14
    for d in range(7):
        assert abs((w * np.arange(7) ** d).sum() - t ** d) <= 1e-12 * max(1.0, t ** d)
    if code % 2 == 0:
        assert np.array_equal(w, np.eye(7)[code // 2])


def test_two_level_fill_reproduces_degree_6_polynomials():
    t = W.two_level_lattice()
    assert len(t) == 15
    nbr = W.neighbors(t)
    st = state_from(t, poly6)
    O.padding_fill(st, t, nbr, general=True)
    nfilled = {1: 0, 2: 0}
    for b, blk in enumerate(t):
        n = int(blk["sz"][0])
        off = int(blk["offset"])
        got = st[0, off:off + n ** 3].reshape(n, n, n)
        x, y, z = coords(blk)
        want = poly6(x, y, z)
        exp = inside_other(t, b)
        filled = ~np.isnan(got)
        filled[3:-3, 3:-3, 3:-3] = False
        assert np.array_equal(filled, exp), b
        err = np.abs(got[exp] - want[exp]).max() / np.abs(want[exp]).max()
        assert err < 1e-13, (b, err)
        # every variable (scaled copies of the field) filled alike
        assert np.allclose(st[5, off:off + n ** 3].reshape(n, n, n)[exp], 1.05 * want[exp], rtol=1e-13)
        h = (blk["pmax"][0] - blk["pmin"][0]) / (n - 1)
        nfilled[1 if h < 0.075 else 2] += int(exp.sum())
    assert nfilled[1] > 0 and nfilled[2] > 0


def test_two_level_fill_interpolates_at_order_6():
    """A degree-7 term along x is reproduced by the copies and injections but
    not by the interpolation from the coarse blocks (and the fine blocks'
    padding facing the coarse ones is interpolated, not copied)."""
    t = W.two_level_lattice()
    nbr = W.neighbors(t)
    f7 = lambda x, y, z: (x + 0.3) ** 7 + y  # noqa: E731
    st = state_from(t, f7)
    O.padding_fill(st, t, nbr, general=True)
    worst_fine = worst_coarse = 0.0
    for b, blk in enumerate(t):
        n = int(blk["sz"][0])
        off = int(blk["offset"])
        got = st[0, off:off + n ** 3].reshape(n, n, n)
        x, y, z = coords(blk)
        exp = inside_other(t, b)
        e = np.abs(got[exp] - f7(x, y, z)[exp]).max()
        h = (blk["pmax"][0] - blk["pmin"][0]) / (n - 1)
        if h < 0.075:
            worst_fine = max(worst_fine, e)
        else:
            worst_coarse = max(worst_coarse, e)
    assert worst_coarse < 1e-13     # coarse padding: copies and injections only
    assert worst_fine > 1e-9        # fine padding facing coarse blocks: interpolated


def test_mixed_size_blocks_copy_from_each_other():
    t = W.mixed_size_lattice()
    nbr = W.neighbors(t)
    st = state_from(t, poly6)
    O.padding_fill(st, t, nbr, general=True)
    for b, blk in enumerate(t):
        n = int(blk["sz"][0])
        off = int(blk["offset"])
        got = st[0, off:off + n ** 3].reshape(n, n, n)
        x, y, z = coords(blk)
        exp = inside_other(t, b)
        assert exp.sum() > 0
        assert np.abs(got[exp] - poly6(x, y, z)[exp]).max() < 1e-13 * np.abs(poly6(x, y, z)).max()


def test_same_size_fill_unchanged_without_general_sources():
    """On a same-size lattice the general fill adds nothing to orc_halo_fill."""
    t = W.lattice(shape=(2, 2, 2), n=13, h=0.05, origin=(-0.37, -0.21, -0.13))
    nbr = W.neighbors(t)
    rnd = np.random.default_rng(2).standard_normal((24, W.total_padded(t)))
    a, b = rnd.copy(), rnd.copy()
    O.halo_fill(a, t, nbr)
    O.padding_fill(b, t, nbr, general=True)
    assert np.array_equal(a, b)


def test_refined_octree_is_balanced_and_shares_nodes():
    """The refined-octree workload: every pair of blocks whose interiors touch
    (face, edge or corner) differs by at most one level, the leaves tile the
    domain, and same-level neighbours share their boundary nodes (the link
    table finds them)."""
    t = W.refined_octree(2, 4, 1.0)
    h = (t["pmax"][:, 0] - t["pmin"][:, 0]) / 12
    lo = t["pmin"] + 3 * h[:, None]
    hi = lo + 6 * h[:, None]
    assert abs(float((6 * h) ** 3 @ np.ones(len(t))) - 32.0 ** 3) < 1e-9  # leaves tile [-16, 16]^3
    touch = np.all((lo[:, None, :] <= hi[None, :, :] + 1e-9) & (lo[None, :, :] <= hi[:, None, :] + 1e-9), axis=2)
    np.fill_diagonal(touch, False)
    ratio = h[:, None] / h[None, :]
    assert np.all(np.isclose(ratio[touch], 0.5) | np.isclose(ratio[touch], 1.0) | np.isclose(ratio[touch], 2.0))
    assert (np.isclose(ratio[touch], 2.0)).any()  # level jumps exist
    nbr = W.neighbors(t)
    assert (nbr >= 0).sum() > 0


def test_refined_octree_fill_reproduces_degree_6_polynomials():
    """On the refined octree (three levels) the general fill reproduces a
    polynomial of degree <= 6 at every padding point inside another block,
    and leaves the padding outside the domain untouched."""
    t = W.refined_octree(2, 4, 1.0)
    nbr = W.neighbors(t)
    f = lambda x, y, z: poly6(x / 16.0, y / 16.0, z / 16.0)  # noqa: E731
    st = state_from(t, f)
    O.padding_fill(st, t, nbr, general=True)
    worst, nfill = 0.0, 0
    for b, blk in enumerate(t):
        n = int(blk["sz"][0])
        off = int(blk["offset"])
        got = st[0, off:off + n ** 3].reshape(n, n, n)
        x, y, z = coords(blk)
        filled = ~np.isnan(got)
        filled[3:-3, 3:-3, 3:-3] = False
        inside = (np.abs(x) <= 16 + 1e-9) & (np.abs(y) <= 16 + 1e-9) & (np.abs(z) <= 16 + 1e-9)
        pad = np.ones((n, n, n), bool)
        pad[3:-3, 3:-3, 3:-3] = False
        assert np.array_equal(filled, pad & inside), b  # all in-domain padding has a source
        if filled.any():
            worst = max(worst, np.abs(got[filled] - f(x, y, z)[filled]).max())
            nfill += int(filled.sum())
    assert nfill > 0 and worst < 1e-12, worst
