"""Pins the oracle's Z-order, tree and neighbour stages by closed forms and
brute force over small grids.

- Morton codes (zorder/generate_morton_code.cuh:7-16, compact.cuh:9-21): bit
  interleaving written here directly, over every (x, y) of a 64 x 64 grid and
  random 15-bit pairs; level starts (utilities/get_lvl_idx.cuh:11-24).
- Sorted orders (zorder/get_sorting_indices.cu:3-62): rev_z_order[m] is the
  row-major index of Morton code m; rev_row_major is its inverse.
- Tree (mra/regularisation.cu, extra_significance.cu,
  traverse_tree_of_sig_details.cu): every ancestor of a significant node is
  significant; a leaf is a node whose ancestors are all significant and which
  is not (or is at the finest level); the leaves tile the grid once, in
  increasing Morton order.
- Neighbours (neighbours/find_neighbours.cu:87-108, compaction.cu:109-151):
  brute force over finest cells: N/E from the leaf's last (north-east) finest
  cell, S/W from its first (south-west) cell; the neighbour is the node at the
  coarser of the two leaf levels; grid and DEM edges give -1.
- Encode: the oracle's hierarchy at every level equals the closed-form
  analysis operator applied to its children (to round-off).
"""
import numpy as np
import pytest

import cases
import oracle


def interleave(x, y, bits=16):
    c = 0
    for b in range(bits):
        c |= ((x >> b) & 1) << (2 * b)
        c |= ((y >> b) & 1) << (2 * b + 1)
    return c


def test_morton_closed_form():
    for x in range(64):
        for y in range(64):
            c = oracle.morton(x, y)
            assert c == interleave(x, y)
            assert oracle.compact(c) == x and oracle.compact(c >> 1) == y
    rng = np.random.default_rng(1)
    for x, y in rng.integers(0, 1 << 15, size=(500, 2)):
        assert oracle.morton(int(x), int(y)) == interleave(int(x), int(y))


def test_level_starts():
    for l in range(15):
        assert oracle.lvl_idx(l) == sum(4 ** k for k in range(l)) + 3
        if l >= 1:
            assert oracle.lvl_idx(l) % 4 == 0


def lvl(l):
    return ((1 << (2 * l)) - 1) // 3 + 3


@pytest.fixture(scope="module")
def stepped(tmp_path_factory):
    """A few oracle steps of a seeded wet/dry case with a ragged DEM."""
    d = str(tmp_path_factory.mktemp("tree"))
    out = {}
    for name, kw in (("square", dict(L=5, seed=21)), ("ragged", dict(L=5, seed=8, xsz=27, ysz=19))):
        par = cases.random_case(f"{d}/{name}", t_end=1.0, manning=0.02, **kw)
        o = oracle.Oracle(par, "f64")
        o.step(); o.step(); o.step()
        out[name] = o
    yield out
    for o in out.values():
        o.close()


@pytest.mark.parametrize("name", ["square", "ragged"])
def test_sorted_orders_closed_form(stepped, name):
    o = stepped[name]
    L = int(o.scalar("L"))
    n = 1 << L
    m = np.arange(n * n)
    x = np.array([oracle.compact(int(c)) for c in m])
    y = np.array([oracle.compact(int(c) >> 1) for c in m])
    rz = o.get("rev_z_order")
    assert np.array_equal(rz, x + n * y)
    rr = o.get("rev_row_major")
    assert np.array_equal(rr[rz], m)


def leaf_arrays(o):
    n = int(o.scalar("length"))
    act = o.get("assem.act")[:n].astype(np.int64)
    lev = o.get("assem.lev")[:n].astype(np.int64)
    code = act - np.array([lvl(int(l)) for l in lev])
    return act, lev, code


@pytest.mark.parametrize("name", ["square", "ragged"])
def test_leaves_tile_grid_in_morton_order(stepped, name):
    o = stepped[name]
    L = int(o.scalar("L"))
    act, lev, code = leaf_arrays(o)
    size = 4 ** (L - lev)
    start = code * size
    assert size.sum() == 4 ** L
    assert start[0] == 0
    assert np.array_equal(start[1:], start[:-1] + size[:-1])


@pytest.mark.parametrize("name", ["square", "ragged"])
def test_tree_is_closed_and_leaves_are_its_frontier(stepped, name):
    o = stepped[name]
    L = int(o.scalar("L"))
    sig = o.get("sig")
    for l in range(1, L):
        for c in range(4 ** l):
            if sig[lvl(l) + c]:
                assert sig[lvl(l - 1) + (c >> 2)], (l, c)
    act, lev, code = leaf_arrays(o)
    for a, l, c in zip(act, lev, code):
        for k in range(l):
            assert sig[lvl(k) + (c >> (2 * (l - k)))], (l, c, k)
        if l < L:
            assert not sig[a]


@pytest.mark.parametrize("name", ["square", "ragged"])
def test_neighbours_brute_force(stepped, name):
    o = stepped[name]
    L = int(o.scalar("L"))
    n = 1 << L
    xsz = int(o.scalar("xsz"))
    ysz = int(o.scalar("ysz"))
    act, lev, code = leaf_arrays(o)
    size = 4 ** (L - lev)
    start = code * size
    # finest Morton code -> leaf
    owner = np.repeat(np.arange(act.size), size)
    got = o.neighbours()
    for i in range(act.size):
        l = lev[i]
        first, last = int(start[i]), int(start[i] + size[i] - 1)
        xs, ys = oracle.compact(first), oracle.compact(first >> 1)
        xe, ye = oracle.compact(last), oracle.compact(last >> 1)
        cand = {0: (xe, ye + 1, ye == n - 1 or ye == ysz - 1),     # N
                1: (xe + 1, ye, xe == n - 1 or xe == xsz - 1),     # E
                2: (xs, ys - 1, ys == 0),                          # S
                3: (xs - 1, ys, xs == 0)}                          # W
        for k, (xn, yn, border) in cand.items():
            if border:
                exp = -1
            else:
                j = owner[interleave(xn, yn)]
                mlev = min(l, lev[j])
                exp = lvl(mlev) + (interleave(xn, yn) >> (2 * (L - mlev)))
            assert got[k][i] == exp, (name, i, k, got[k][i], exp)


@pytest.mark.parametrize("name", ["square"])
def test_encoded_hierarchy_matches_analysis_operator(stepped, name):
    o = stepped[name]
    L = int(o.scalar("L"))
    perm = [0, 2, 1, 3]
    F = oracle.filters()
    M = np.zeros((3, 12))
    for i in range(3):
        for k in range(4):
            for j in range(3):
                M[i, 3 * perm[k] + j] = F[0, k, i, j]
    z = np.stack([o.get("z0"), o.get("z1x"), o.get("z1y")])
    # the bed hierarchy is encoded for every parent at set-up
    for l in range(L):
        for c in range(4 ** l):
            kids = np.concatenate([z[:, lvl(l + 1) + 4 * c + q] for q in range(4)])
            par = M @ kids / 2.0
            assert np.allclose(par, z[:, lvl(l) + c], atol=1e-12, rtol=0), (l, c)
