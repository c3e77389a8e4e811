"""Pins of the oracle's first derivatives, normals and mean curvature
(oracle/oracle.c: d_central, oracle_normals, oracle_mean_curvature)."""
import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O


def grid(dim, shape, lmin, lmax):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=2.4).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    _, xyz = f.nodes()
    return f, ngbd.tables, xyz


def hanging(tab):
    """nodes with at least one arm on a coarser cell (nonzero transverse product)"""
    F, T = 2 * tab.dim, 2 * (tab.dim - 1)
    prod = np.zeros(tab.n_nodes, dtype=bool)
    for f in range(F):
        for k in range(0, T, 2):
            prod |= (tab.d_tr[f, k] * tab.d_tr[f, k + 1]) != 0
    return prod


def test_first_derivatives_exact_for_quadratics_2d():
    f, tab, xy = grid(2, ls.Star(), 3, 7)
    x, y = xy[:, 0], xy[:, 1]
    g = 0.3 * x * x - 0.7 * y * y + 1.1 * x * y + 0.4 * x - 0.2 * y + 0.5
    fd = O.first_derivatives(tab, g)
    ex = np.stack([0.6 * x + 1.1 * y + 0.4, -1.4 * y + 1.1 * x - 0.2])
    inner = tab.wall == 0
    hang = hanging(tab) & inner
    assert hang.sum() > 50  # the correction terms are exercised
    assert np.abs(fd[:, inner] - ex[:, inner]).max() < 1e-12


def test_first_derivatives_exact_for_quadratics_3d():
    f, tab, xyz = grid(3, ls.random_spheres(4, seed=2, rmin=0.1, rmax=0.2), 2, 5)
    x, y, z = xyz.T
    g = 0.5 * x * x + 0.2 * y * y - 0.9 * z * z + 0.3 * x * y - 0.4 * y * z + 0.6 * x * z + x - z
    fd = O.first_derivatives(tab, g)
    ex = np.stack([x + 0.3 * y + 0.6 * z + 1, 0.4 * y + 0.3 * x - 0.4 * z, -1.8 * z - 0.4 * y + 0.6 * x - 1])
    inner = tab.wall == 0
    assert (hanging(tab) & inner).sum() > 50
    assert np.abs(fd[:, inner] - ex[:, inner]).max() < 1e-12


def test_first_derivatives_uniform_grid_closed_form():
    f, tab, xy = grid(2, ls.Star(), 5, 5)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(tab.n_nodes)
    fd = O.first_derivatives(tab, g)
    h = 1.0 / 32
    inner = np.where(tab.wall == 0)[0]
    face = O.face_linear(tab, g)  # the face neighbours' values (one corner of weight 1 here)
    for a in range(2):
        gp, gm = face[2 * a + 1, inner], face[2 * a, inner]
        np.testing.assert_allclose(fd[a, inner], (gp - gm) / (2 * h), rtol=1e-12, atol=1e-12 / h)


def test_normals_unit_length_and_direction():
    f, tab, xy = grid(2, ls.Star(), 3, 7)
    c = np.array([0.5, 0.5])
    phi = np.hypot(xy[:, 0] - c[0], xy[:, 1] - c[1]) - 0.3
    nrm = O.normals(tab, phi)
    far = np.hypot(xy[:, 0] - c[0], xy[:, 1] - c[1]) > 0.05
    np.testing.assert_allclose(np.hypot(nrm[0, far], nrm[1, far]), 1.0, rtol=1e-12)
    radial = (xy[far] - c) / np.hypot(*(xy[far] - c).T)[:, None]
    assert np.abs(nrm[:, far].T - radial).max() < 0.05
    # zero gradient -> zero normal
    assert np.all(O.normals(tab, np.ones(tab.n_nodes)) == 0)


@pytest.mark.parametrize("dim,R", [(2, 0.25), (3, 0.3)])
def test_mean_curvature_converges(dim, R):
    errs = []
    for lmax in ((6, 7) if dim == 2 else (5, 6)):
        shape = ls.Sphere([0.5] * dim, R)
        f, tab, xyz = grid(dim, shape, 3, lmax)
        phi = np.linalg.norm(xyz - 0.5, axis=1) - R
        kap = O.mean_curvature(tab, O.normals(tab, phi))
        band = np.abs(phi) < 1.0 / 2 ** lmax
        # the curvature field of the distance function: (dim - 1) / (R + phi)
        errs.append(np.abs(kap[band] - (dim - 1) / (R + phi[band])).max() * R)
    assert errs[1] < 0.3 * errs[0] and errs[1] < 0.01  # second order
