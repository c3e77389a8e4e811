"""Pins of the oracle's advect_in_normal_direction and its CFL time step
(oracle/oracle.c: advect_one, oracle_advect, oracle_advect_cfl_dt)."""
import numpy as np

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O


def grid(dim, shape, lmin, lmax):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=2.4).build_nodes()
    _, xyz = f.nodes()
    return C.NodeNeighbors(f).init_neighbors().tables, xyz


def test_zero_velocity_is_identity():
    tab, xy = grid(2, ls.Star(), 3, 7)
    phi = np.sin(3 * xy[:, 0]) * np.cos(2 * xy[:, 1])
    out = O.advect(tab, phi, np.zeros(tab.n_nodes), 1e-3)
    assert np.array_equal(out, phi)


def test_linear_field_moves_exactly():
    # phi = a.x + b with vn: phi_t = -vn |a| exactly (every difference and
    # second derivative of a linear field is exact, the ENO terms vanish)
    tab, xy = grid(2, ls.Star(), 3, 7)
    a = np.array([0.6, -0.8])
    phi = xy @ a + 0.1
    for vn in (2.0, -1.5):
        dt = 1e-3
        out = O.advect(tab, phi, np.full(tab.n_nodes, vn), dt)
        # away from the walls, whose Neumann mirror stage 2 reads (coarse cells are 1/8 there)
        inner = np.all((xy > 0.25) & (xy < 0.75), axis=1)
        np.testing.assert_allclose(out[inner], phi[inner] - dt * vn * 1.0, rtol=0, atol=1e-14)


def test_expanding_circle_and_sphere():
    # phi = |x - c| - R, vn = 1: the front moves outward, phi -> phi - dt
    for dim, lmax in ((2, 8), (3, 6)):
        shape = ls.Sphere([0.5] * dim, 0.25)
        tab, xyz = grid(dim, shape, 3, lmax)
        r = np.linalg.norm(xyz - 0.5, axis=1)
        phi = r - 0.25
        h = 1.0 / 2 ** lmax
        dt = 0.4 * h
        out = O.advect(tab, phi, np.ones(tab.n_nodes), dt)
        band = (np.abs(phi) < 3 * h) & (tab.wall == 0)
        err = np.abs(out[band] - (phi[band] - dt))
        assert err.max() < 0.05 * dt  # O(dt h) near a smooth front


def test_second_stage_velocity():
    tab, xy = grid(2, ls.Star(), 3, 6)
    phi = xy[:, 0] - 0.4
    inner = np.all((xy > 0.25) & (xy < 0.75), axis=1)
    out = O.advect(tab, phi, np.full(tab.n_nodes, 1.0), 1e-3, vn_np1=np.full(tab.n_nodes, 3.0))
    # first stage moves by 1 dt, the second by 3 dt from there: average 2 dt
    np.testing.assert_allclose(out[inner], phi[inner] - 2e-3, atol=1e-14)


def test_cfl_time_step_closed_form():
    tab, xy = grid(2, ls.Star(), 5, 5)
    vn = np.linspace(-2.0, 1.0, tab.n_nodes)
    assert O.advect_cfl_dt(tab, vn) == 0.8 * (1.0 / 32) / 2.0
