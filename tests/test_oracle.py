"""Pins the CPU oracle (oracle/oracle.c) with checks the mathematics fixes,
chosen so that a dropped term, a wrong sign or index, or a swapped operand in
the restatement fails one of them:

* MINMOD and interface_Location* known answers, both root branches, the EPS
  early outs and the c2 ~ 0 branch (closed-form roots);
* uniform grids: pass A is the textbook (f+ - 2 f0 + f-)/h^2;
* adaptive grids with hanging nodes: corrected second derivatives and the
  quadratic neighbour interpolation are exact for every quadratic field
  (correct_naive_second_derivatives removes exactly the linear-interpolation
  bias 0.5 d_m0 d_p0 f_yy), including the negative wall arms;
* reinitialization invariants: a linear signed distance is a fixed point, a
  steeper linear field converges to the signed distance, sign preservation,
  anchoring, band limits, masks, exact mirror symmetry, and second-order
  convergence of the band error (examples/reinitialization/main_2d.cpp:250-274).
"""
import math

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O

EPS = 1e-13


def grid(dim, shape, lmin, lmax, lip=2.4):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=lip).build_nodes()
    tab = C.flat_stencil_from_forest(f)
    _, xyz = f.nodes()
    return f, tab, xyz


# ---------------------------------------------------------------- scalars
def test_minmod_known_answers():
    assert O.minmod(1.0, 2.0) == 1.0
    assert O.minmod(-3.0, -2.0) == -2.0
    assert O.minmod(1.0, -2.0) == 0.0
    assert O.minmod(0.0, 5.0) == 0.0
    assert O.minmod(2.0, 2.0) == 2.0  # tie returns b
    assert math.copysign(1, O.minmod(-1.0, -1.0)) == -1


def test_interface_location_linear_root():
    # f(x) = 3 (x - 0.3) on [0.1, 0.5]: root 0.3
    a, b = 0.1, 0.5
    fa, fb = 3 * (a - 0.3), 3 * (b - 0.3)
    assert abs(O.interface_location(a, b, fa, fb) - 0.3) < 1e-15
    # orientation: root near a when |fa| small
    assert O.interface_location(0.0, 1.0, -0.1, 0.9) == pytest.approx(0.1, abs=1e-15)


@pytest.mark.parametrize("c", [2.5, -1.7])
@pytest.mark.parametrize("root", [0.013, 0.041])
def test_interface_location_2nd_quadratic_root(c, root):
    # f(x) = c (x - root)(x - other): exact root when f'' = 2c is supplied
    other = 0.9
    f = lambda x: c * (x - root) * (x - other)
    a, b = 0.0, 0.05
    x = O.interface_location_2nd(a, b, f(a), f(b), 2 * c, 2 * c)
    assert abs(x - root) < 1e-14
    # the minmod picks the smaller-magnitude second derivative
    x2 = O.interface_location_2nd(a, b, f(a), f(b), 2 * c, 4 * c)
    assert abs(x2 - root) < 1e-14
    # opposite second-derivative signs -> minmod 0 -> linear root
    x3 = O.interface_location_2nd(a, b, f(a), f(b), 2 * c, -2 * c)
    assert abs(x3 - O.interface_location(a, b, f(a), f(b))) < 1e-15


def test_interface_location_2nd_branches():
    # fb < 0 and fb > 0 branches give the root inside (a, b) for both orientations
    for sgn in (1.0, -1.0):
        g = lambda x: sgn * ((x - 0.02) + 5.0 * (x - 0.02) ** 2)
        a, b = 0.0, 0.04
        x = O.interface_location_2nd(a, b, g(a), g(b), sgn * 10.0, sgn * 10.0)
        assert abs(x - 0.02) < 1e-14
    # EPS early outs
    assert O.interface_location_2nd(0.0, 1.0, 1e-14, -1.0, 0.0, 0.0) == 0.0 + EPS
    assert O.interface_location_2nd(0.0, 1.0, -1.0, 1e-14, 0.0, 0.0) == 1.0 - EPS
    # |c2| < EPS -> linear: x = -c0/c1 about the midpoint
    x = O.interface_location_2nd(0.0, 0.1, -0.2, 0.2, 1e-14, 1e-14)
    assert abs(x - 0.05) < 1e-15


# ---------------------------------------------------------------- pass A
def test_uniform_grid_second_differences():
    f, tab, xyz = grid(2, ls.Sphere((0.5, 0.5), 0.3), 5, 5)
    h = 1.0 / 32
    rng = np.random.default_rng(3)
    g = rng.standard_normal(tab.n_nodes)
    dd = O.second_derivatives(tab, g)
    # textbook second difference on the interior, from a coordinate lookup
    key = {(round(x * 32), round(y * 32)): i for i, (x, y) in enumerate(xyz)}
    for (i, j), n in key.items():
        if 0 < i < 32 and 0 < j < 32:
            fxx = (g[key[(i + 1, j)]] - 2 * g[n] + g[key[(i - 1, j)]]) / h ** 2
            fyy = (g[key[(i, j + 1)]] - 2 * g[n] + g[key[(i, j - 1)]]) / h ** 2
            assert abs(dd[0, n] - fxx) <= 1e-9 * max(1.0, abs(fxx))
            assert abs(dd[1, n] - fyy) <= 1e-9 * max(1.0, abs(fyy))


def quadratic(xyz, dim, seed):
    rng = np.random.default_rng(seed)
    if dim == 2:
        a, b, c, d, e, g0 = rng.uniform(-2, 2, 6)
        x, y = xyz[:, 0], xyz[:, 1]
        return a * x * x + b * x * y + c * y * y + d * x + e * y + g0, np.array([2 * a, 2 * c])
    a, b, c, p, q, r, d, e, g, k = rng.uniform(-2, 2, 10)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    f = a * x * x + b * y * y + c * z * z + p * x * y + q * y * z + r * x * z + d * x + e * y + g * z + k
    return f, np.array([2 * a, 2 * b, 2 * c])


GRADED = [
    (2, ls.Star(), 2, 7),
    (2, ls.Sphere((0.3, 0.61), 0.17), 1, 8),
    (3, ls.Sphere((0.45, 0.5, 0.55), 0.25), 2, 5),
    (3, ls.random_spheres(3, seed=11, rmin=0.08, rmax=0.15), 1, 5),
]


@pytest.mark.parametrize("dim,shape,lmin,lmax", GRADED)
def test_second_derivatives_exact_for_quadratics(dim, shape, lmin, lmax):
    _, tab, xyz = grid(dim, shape, lmin, lmax)
    hanging = np.any(tab.d_tr[:, 0] * tab.d_tr[:, 1] != 0, axis=0)
    assert hanging.sum() > 10, "test needs hanging arms"
    for seed in range(3):
        f, exact = quadratic(xyz, dim, seed)
        dd = O.second_derivatives(tab, f)
        err = np.max(np.abs(dd - exact[:, None]))
        assert err < 1e-7, err  # exact up to rounding of O(1/h^2) cancellations


@pytest.mark.parametrize("dim,shape,lmin,lmax", GRADED)
def test_quadratic_neighbor_interpolation_exact(dim, shape, lmin, lmax):
    _, tab, xyz = grid(dim, shape, lmin, lmax)
    f, _ = quadratic(xyz, dim, 5)
    q = O.ngbd_quadratic(tab, f)
    assert np.array_equal(q[-1], f)
    # neighbour along face direction: x -/+ d_face on axis a (negative d on a
    # wall arm points back into the domain)
    for face in range(2 * dim):
        a, s = face >> 1, face & 1
        pos = xyz.copy()
        pos[:, a] += (1 if s else -1) * tab.d[face]
        fexact, _ = quadratic(pos, dim, 5)
        assert np.max(np.abs(q[face] - fexact)) < 1e-12


def test_face_linear_exact_for_linear_fields():
    _, tab, xyz = grid(3, ls.Sphere((0.45, 0.5, 0.55), 0.25), 2, 5)
    g = 0.3 + 1.5 * xyz[:, 0] - 0.7 * xyz[:, 1] + 2.1 * xyz[:, 2]
    fl = O.face_linear(tab, g)
    for face in range(6):
        a, s = face >> 1, face & 1
        pos = xyz.copy()
        pos[:, a] += (1 if s else -1) * tab.d[face]
        ge = 0.3 + 1.5 * pos[:, 0] - 0.7 * pos[:, 1] + 2.1 * pos[:, 2]
        assert np.max(np.abs(fl[face] - ge)) < 1e-14


# ---------------------------------------------------------------- reinit
def test_linear_signed_distance_is_fixed_point():
    for dim in (2, 3):
        shape = ls.Plane([1.0] + [0.0] * (dim - 1), -0.37)
        _, tab, xyz = grid(dim, shape, 2, 6 if dim == 2 else 4, lip=1.2)
        phi0 = xyz[:, 0] - 0.37
        for os_, ot in ((2, 2), (1, 1), (2, 1), (1, 2)):
            out = O.reinit(tab, phi0, os_, ot, 10)
            assert np.max(np.abs(out - phi0)) < 1e-13


def test_steep_linear_field_converges_to_distance():
    shape = ls.Plane([1.0, 0.0], -0.43)
    _, tab, xyz = grid(2, shape, 3, 7, lip=1.2)
    phi0 = 3.0 * (xyz[:, 0] - 0.43)
    out = O.reinit(tab, phi0, 2, 2, 200)
    band = np.abs(xyz[:, 0] - 0.43) < 0.1
    assert np.max(np.abs(out[band] - (xyz[band, 0] - 0.43))) < 1e-10


def star_problem(lmin=3, lmax=7):
    f, tab, xyz = grid(2, ls.Star(), lmin, lmax)
    phi0 = f.sample(ls.Star())
    return f, tab, xyz, phi0


def test_sign_preserved_and_anchor_and_limits():
    f, tab, xyz, phi0 = star_problem()
    out = O.reinit(tab, phi0, 2, 2, 20)
    nz = phi0 != 0
    assert np.all(np.sign(out[nz]) == np.sign(phi0[nz]))
    # anchoring: tiny |phi0| -> stage value 0; Euler stores it, RK2 averages
    # it with phi_n each iteration (phi_n = (phi_n + 0)/2)
    p = phi0.copy()
    p[100] = 3e-14
    assert O.reinit(tab, p, 2, 1, 3)[100] == 0.0
    assert O.reinit(tab, p, 2, 2, 3)[100] == 3e-14 / 8
    # outside (limit_low, limit_high) nodes keep phi0
    lim = 0.05
    out = O.reinit(tab, phi0, 2, 2, 5, limit_high=lim, limit_low=-lim)
    outside = np.abs(phi0) >= lim * (1 + 1e-12) + 1e-13
    assert np.array_equal(out[outside], phi0[outside])
    assert np.any(out[~outside] != phi0[~outside])
    # invalid orders
    with pytest.raises(ValueError):
        O.reinit(tab, phi0, 3, 2, 1)


def test_mask_keeps_masked_nodes():
    f, tab, xyz, phi0 = star_problem()
    mask = (xyz[:, 0] < 0.5).astype(np.uint8)
    out = O.reinit(tab, phi0, 2, 2, 10, mask=mask)
    assert np.array_equal(out[mask == 0], phi0[mask == 0])
    assert np.any(out[mask == 1] != phi0[mask == 1])


def test_mirror_symmetry_is_exact():
    # the star is symmetric under y -> 1 - y; the forest and every arithmetic
    # step are too, so the result must be exactly mirror symmetric
    f, tab, xyz, phi0 = star_problem(3, 8)
    ixyz, _ = f.nodes()
    R = 1 << 30
    key = {(int(a), int(b)): i for i, (a, b) in enumerate(ixyz)}
    mirror = np.array([key[(int(a), R - int(b))] for a, b in ixyz])
    assert np.array_equal(phi0, phi0[mirror])
    for os_, ot in ((2, 2), (1, 1)):
        out = O.reinit(tab, phi0, os_, ot, 20)
        assert np.array_equal(out, out[mirror])


def test_second_order_convergence_in_band():
    # circle signed distance perturbed as in examples/reinitialization/main_2d.cpp:112-116
    errs = []
    for lmax in (6, 7, 8):
        shape = ls.Modulated(center=(0.5, 0.5), radius=0.3, q=0.5)
        f = C.Forest(2).refine_levelset(ls.Sphere((0.5, 0.5), 0.3), lmax - 3, lmax, lip=1.2).build_nodes()
        tab = C.flat_stencil_from_forest(f)
        _, xyz = f.nodes()
        phi0 = f.sample(shape)
        exact = f.sample(ls.Sphere((0.5, 0.5), 0.3))
        out = O.reinit(tab, phi0, 2, 2, 2 ** (lmax - 3) * 4)
        diag = math.sqrt(2) / 2 ** lmax
        band = np.abs(out) < 3 * diag
        errs.append(np.max(np.abs(out[band] - exact[band])))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert all(o > 1.5 for o in orders), (errs, orders)


def test_thread_count_does_not_change_results():
    f, tab, xyz, phi0 = star_problem()
    n0 = O.num_threads()
    O.set_num_threads(1)
    a = O.reinit(tab, phi0, 2, 2, 5)
    O.set_num_threads(max(2, n0))
    b = O.reinit(tab, phi0, 2, 2, 5)
    O.set_num_threads(n0)
    assert np.array_equal(a, b)
