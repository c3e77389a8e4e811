"""The tiled encoding, built and checked on the host (casl_tile_plan): every
record decodes to its flattened stencil (the builder's own check), indices
and extension ranges pass the upload checks, and the counts agree with what
the flattened tables say independently: the nodes with a hanging face (a face
whose transverse products d_m0 d_p0 or d_0m d_0p are nonzero), at most D - 1
hanging faces per node, none on a uniform grid, and no encoding for cells
that are not cubes."""
import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls


def hanging_faces(tab):
    dim = tab.dim
    t = np.asarray(tab.d_tr).reshape(2 * dim, 2 * (dim - 1), -1)
    h = t[:, 0] * t[:, 1] != 0
    if dim == 3:
        h |= t[:, 2] * t[:, 3] != 0
    return h.sum(0)


@pytest.mark.parametrize("dim,shape,lmin,lmax", [
    (2, ls.Star(), 3, 9),
    (2, ls.Sphere([0.5, 0.5], 0.3), 1, 10),
    (3, ls.random_spheres(12, seed=4, rmin=0.06, rmax=0.12), 2, 7),
    (3, ls.Sphere([0.45, 0.5, 0.55], 0.3), 1, 7),
    (3, ls.Gyroid(), 4, 6),  # hanging faces at wall nodes too
], ids=["star2d", "circle2d_jump", "spheres3d", "sphere3d_jump", "gyroid3d_walls"])
def test_plan_counts_match_tables(dim, shape, lmin, lmax):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=2.4).build_nodes()
    tab = C.flat_stencil_from_forest(f)
    plan = C.tile_plan(tab)
    nh = hanging_faces(tab)
    assert nh.max() <= dim - 1
    assert plan["n_irregular_nodes"] == int((nh > 0).sum()) > 0
    assert plan["n_records"] >= tab.n_nodes and plan["n_irregular"] >= plan["n_irregular_nodes"]
    if dim == 2:
        assert plan["n_two_hanging"] == 0
    else:
        assert 0 < plan["n_two_hanging"] < plan["n_irregular"]


def test_uniform_grid_has_no_hanging_faces():
    f = C.Forest(3).refine_levelset(ls.Sphere([0.5, 0.5, 0.5], 0.3), 4, 4).build_nodes()
    plan = C.tile_plan(C.flat_stencil_from_forest(f))
    assert plan["n_irregular"] == 0 and plan["n_tiles"] >= 1


@pytest.mark.parametrize("dim,box_max", [(2, [2.0, 1.0]), (3, [2.0, 1.0, 0.5]), (2, [3.0, 3.0])],
                         ids=["rect2d", "brick3d", "square_3"])
def test_noncubic_or_nondyadic_cells_are_tiled(dim, box_max):
    """Boxes whose edges differ by powers of two: cells are not cubes, so a
    hanging arm's transverse extents differ from its length.  The encoding
    keeps one scale per axis; the builder's record check (every decoded
    distance, transverse offset and corner equal to the tables') passes,
    hanging faces included.  A cubic box of any edge (3 x 3) is tiled too."""
    shape = ls.Sphere([0.45 * e for e in box_max], 0.3 * min(box_max))
    f = C.Forest(dim, [0.0] * dim, box_max).refine_levelset(shape, 2, 6, lip=2.4).build_nodes()
    plan = C.tile_plan(C.flat_stencil_from_forest(f))
    assert plan["n_tiles"] >= 1 and plan["n_irregular"] > 0, plan


@pytest.mark.parametrize("dim,box_max", [(2, [3.0, 1.0]), (3, [1.5, 1.0, 0.75])], ids=["rect_3to1", "brick_nonpow2"])
def test_noncubic_nondyadic_cells_are_not_tiled(dim, box_max):
    """Cells that are not cubes on edges that are not powers of two: no
    tiled encoding (such grids run the grid-wide kernels)."""
    shape = ls.Sphere([0.45 * e for e in box_max], 0.3 * min(box_max))
    f = C.Forest(dim, [0.0] * dim, box_max).refine_levelset(shape, 2, 6, lip=2.4).build_nodes()
    with pytest.raises(C.CaslError, match="not cubes"):
        C.tile_plan(C.flat_stencil_from_forest(f))
