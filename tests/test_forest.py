"""Host forest + stencil flattener against the brute-force restatement of
refine_levelset_cf / my_p4est_nodes_new / construct_neighbors (oracle/pyoracle.py)."""
import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O


def build(dim, shape, lmin, lmax, lip=2.4):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=lip).build_nodes()
    return f, C.flat_stencil_from_forest(f)


CASES = [
    (2, ls.Star(), 2, 5),
    (2, ls.Star(center=(0.45, 0.52), r0=0.2, amplitude=0.07, lobes=3), 1, 6),
    (2, ls.Sphere((0.31, 0.4), 0.22), 3, 6),
    (3, ls.Sphere((0.4, 0.55, 0.5), 0.27), 1, 4),
    (3, ls.random_spheres(4, seed=7, rmin=0.1, rmax=0.2), 2, 4),
]


@pytest.mark.parametrize("dim,shape,lmin,lmax", CASES)
def test_stencils_match_brute_force(dim, shape, lmin, lmax):
    f, tab = build(dim, shape, lmin, lmax)
    xyz, lvl, local = f.leaves()
    assert tab.n_nodes < 6000
    nbr, d, dtr, wall, nodes = O.brute_stencil(xyz, lvl, dim)
    ixyz, _ = f.nodes()
    # node numbering: same points in the same (p4est Morton) order
    assert ixyz.shape[0] == len(nodes)
    assert np.array_equal(ixyz, np.array(nodes, dtype=np.int32))
    # stencil indices bit-exact, distances bit-exact
    assert np.array_equal(tab.nbr, nbr)
    assert np.array_equal(tab.d, d)
    assert np.array_equal(tab.d_tr, dtr)
    assert np.array_equal(tab.wall, wall)


@pytest.mark.parametrize("dim,shape,lmin,lmax", CASES[:1] + CASES[3:4])
def test_refinement_matches_refine_levelset_cf(dim, shape, lmin, lmax):
    f, _ = build(dim, shape, lmin, lmax)
    xyz, lvl, _ = f.leaves()
    ML = O.MAXLEVEL[dim]
    for q, lv in zip(xyz, lvl):
        # a leaf is not refined ...
        assert not O.refine_decision(q, int(lv), dim, lmin, lmax, 2.4, 0.0, shape)
        # ... and its parent was
        if lv > 0:
            h = 1 << (ML - int(lv) + 1)
            parent = [int(c) // h * h for c in q]
            assert O.refine_decision(parent, int(lv) - 1, dim, lmin, lmax, 2.4, 0.0, shape)


def test_leaves_tile_domain_in_morton_order():
    for dim in (2, 3):
        f, _ = build(dim, ls.Sphere([0.5] * dim, 0.3), 2, 5 if dim == 2 else 4)
        xyz, lvl, _ = f.leaves()
        ML = O.MAXLEVEL[dim]
        vol = np.sum((1.0 / 2.0 ** lvl.astype(float)) ** dim)
        assert abs(vol - 1.0) < 1e-12
        for i in range(len(xyz) - 1):
            assert O.quadrant_compare(tuple(int(v) for v in xyz[i]), tuple(int(v) for v in xyz[i + 1]), dim) < 0


def test_uniform_grid_stencil_closed_form():
    # lmin == lmax: every interior arm has length h and the node sits on a
    # corner of the arm's cell: one transverse offset is h, the other 0 (the
    # lower cell wins the first-wins tie, so m0 = h, p0 = 0)
    f, tab = build(2, ls.Sphere((0.5, 0.5), 0.3), 4, 4)
    h = 1.0 / 16
    assert tab.n_nodes == 17 * 17
    interior = tab.wall == 0
    assert np.all(tab.d[:, interior] == h)
    assert np.all(tab.d_tr[:, 0, interior] * tab.d_tr[:, 1, interior] == 0.0)
    assert np.all(tab.d_tr[:, 0, interior] + tab.d_tr[:, 1, interior] == h)
    # wall arms: second-degree neighbor at -2h
    _, xyz = f.nodes()
    xm = (tab.wall & 1) != 0
    assert np.all(xyz[xm, 0] == 0.0)
    assert np.all(tab.d[0, xm] == -2 * h)
    # the "m00" arm of an x-min wall node points at the node two cells to the right
    for n in np.nonzero(xm)[0]:
        m = tab.nbr[0, 0, n] if tab.d_tr[0, 0, n] == 0 else tab.nbr[0, 1, n]
        assert xyz[m, 0] == 2 * h and xyz[m, 1] == xyz[n, 1]


def test_sample_matches_levelset_eval():
    f, _ = build(2, ls.Star(), 2, 6)
    _, xyz = f.nodes()
    star = ls.Star()
    assert np.array_equal(f.sample(star), star(xyz))
    # python restatement of the star formula (geom::Star-like): r = 0.3 + 0.05 cos 5t, (1 + 0.8 |x - c|^2)
    dx, dy = xyz[:, 0] - 0.5, xyz[:, 1] - 0.5
    r = 0.3 + 0.05 * np.cos(5 * np.arctan2(dy, dx))
    ref = (np.sqrt(dx * dx + dy * dy) - r) * (1 + 0.8 * (dx * dx + dy * dy))
    assert np.max(np.abs(f.sample(star) - ref)) < 1e-15


def test_invalid_arguments_raise():
    with pytest.raises(ValueError):
        C.Forest(4)
    f = C.Forest(2)
    with pytest.raises(C.CaslError):
        f.refine_levelset(ls.Spheres([[0.5, 0.5, 0.5]], [0.1]), 1, 3)  # 3D shape on a 2D forest
    with pytest.raises(C.CaslError):
        C.flat_stencil_from_forest(f)  # nodes not built


BRICKS = [
    # dim, box, trees per axis, periodic, shape, lmin, lmax
    (2, ([0.0, 0.0], [2.0, 1.0]), [2, 1], [0, 0], ls.Sphere((0.9, 0.45), 0.3), 1, 5),
    (2, ([0.0, 0.0], [1.0, 1.0]), [1, 1], [1, 1], ls.Sphere((0.1, 0.5), 0.25), 2, 5),
    (2, ([-1.0, 0.0], [2.0, 2.0]), [3, 2], [1, 0], ls.Star(center=(0.0, 0.9), r0=0.5, amplitude=0.1, lobes=5), 1, 4),
    (3, ([0.0, 0.0, 0.0], [2.0, 1.0, 1.0]), [2, 1, 1], [0, 1, 0], ls.Sphere((0.95, 0.1, 0.5), 0.3), 1, 3),
    (3, ([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]), [1, 2, 2], [1, 1, 1], ls.Sphere((0.05, 0.5, 0.95), 0.3), 1, 3),
]


def build_brick(dim, box, n_trees, periodic, shape, lmin, lmax, lip=2.4):
    f = C.Forest(dim, box[0], box[1], n_trees=n_trees, periodic=periodic)
    f.refine_levelset(shape, lmin, lmax, lip=lip).build_nodes()
    return f, C.flat_stencil_from_forest(f)


@pytest.mark.parametrize("dim,box,n_trees,periodic,shape,lmin,lmax", BRICKS,
                         ids=["2x1", "periodic_xy", "3x2_periodic_x", "3d_2x1x1_periodic_y", "3d_1x2x2_periodic"])
def test_brick_stencils_match_brute_force(dim, box, n_trees, periodic, shape, lmin, lmax):
    """Bricks of trees, periodic or not: nodes (owner tree and coordinates,
    in order), neighbour ids, face and transverse distances and wall bits
    bit for bit against the brute-force brick restatement."""
    f, tab = build_brick(dim, box, n_trees, periodic, shape, lmin, lmax)
    assert tab.n_nodes < 6000
    xyz, lvl, _ = f.leaves()
    pos, leaf_tree, node_tree = f.trees()
    nbr, d, dtr, wall, nodes = O.brute_stencil_brick(xyz, lvl, leaf_tree, pos, dim, box[0], box[1], n_trees,
                                                      periodic)
    ixyz, _ = f.nodes()
    assert len(nodes) == tab.n_nodes
    assert np.array_equal(node_tree, np.array([t for t, _ in nodes], dtype=np.int32))
    assert np.array_equal(ixyz, np.array([c for _, c in nodes], dtype=np.int32))
    assert np.array_equal(tab.nbr, nbr)
    assert np.array_equal(tab.d, d)
    assert np.array_equal(tab.d_tr, dtr)
    assert np.array_equal(tab.wall, wall)
    # periodic axes have no wall; every face distance is positive off the walls
    for a in range(dim):
        if periodic[a]:
            assert not np.any(tab.wall & (3 << (2 * a)))
    assert np.all(tab.d[:, tab.wall == 0] > 0)


@pytest.mark.parametrize("dim,box,n_trees,periodic,shape,lmin,lmax", BRICKS[:1] + BRICKS[3:4], ids=["2d", "3d"])
def test_brick_refinement_matches_refine_levelset_cf(dim, box, n_trees, periodic, shape, lmin, lmax):
    f, _ = build_brick(dim, box, n_trees, periodic, shape, lmin, lmax)
    xyz, lvl, _ = f.leaves()
    pos, leaf_tree, _ = f.trees()
    vmin, ext, _ = O.brick_geometry(dim, box[0], box[1], n_trees)
    ML = O.MAXLEVEL[dim]
    for q, lv, t in zip(xyz, lvl, leaf_tree):
        tmin = [vmin[a][pos[t][a]] for a in range(dim)]
        text = [ext[a][pos[t][a]] for a in range(dim)]
        assert not O.refine_decision(q, int(lv), dim, lmin, lmax, 2.4, 0.0, shape, tree_min=tmin, tree_ext=text)
        if lv > 0:
            h = 1 << (ML - int(lv) + 1)
            parent = [int(c) // h * h for c in q]
            assert O.refine_decision(parent, int(lv) - 1, dim, lmin, lmax, 2.4, 0.0, shape, tree_min=tmin,
                                     tree_ext=text)


def test_brick_tree_order_and_leaf_order():
    """Trees follow a Morton curve of their brick positions; leaves run tree
    by tree, each tree's leaves in Morton order, and tile the domain."""
    f, _ = build_brick(2, ([0.0, 0.0], [3.0, 2.0]), [3, 2], [0, 0], ls.Sphere((1.5, 1.0), 0.6), 1, 4)
    pos, leaf_tree, _ = f.trees()
    assert sorted(map(tuple, pos)) == [(i, j) for i in range(3) for j in range(2)]
    for t in range(len(pos) - 1):
        a, b = tuple(int(v) for v in pos[t]), tuple(int(v) for v in pos[t + 1])
        assert O.quadrant_compare(a, b, 2) < 0
    xyz, lvl, _ = f.leaves()
    assert np.all(np.diff(leaf_tree) >= 0)
    for i in range(len(xyz) - 1):
        if leaf_tree[i] == leaf_tree[i + 1]:
            assert O.quadrant_compare(tuple(int(v) for v in xyz[i]), tuple(int(v) for v in xyz[i + 1]), 2) < 0
    vol = np.sum((1.0 / 2.0 ** lvl.astype(float)) ** 2)
    assert abs(vol - 6.0) < 1e-12


def test_brick_rejects_bad_arguments():
    with pytest.raises(C.CaslError):
        C.Forest(2, n_trees=[0, 1])
    with pytest.raises(C.CaslError):
        C.Forest(3, n_trees=[2000, 1, 1])
    with pytest.raises(ValueError):
        C.Forest(2, n_trees=[1, 1, 1])
