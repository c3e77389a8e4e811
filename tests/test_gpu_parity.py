"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs, the golden fixtures, and size-independent properties.

Bars: every FP64 reinitialization and advection result, second derivatives
included, equals the oracle's bit for bit (the kernels repeat the oracle's
operation order; np.array_equal); the first-derivative / normal / curvature
chain agrees within 1e-12 relative (max |gpu - oracle| <= 1e-12 max |oracle|);
the mixed path must stay within 1e-3 * h_min of the FP64 result in the
|phi| < 3 h_min band, and the nodes it never updates (outside the phi_0
limits, masked) keep their FP64 input bit for bit.
"""
import os

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL = 1e-12


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    scale = max(1e-300, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) / scale


def problem(dim, shape, lmin, lmax, lip=2.4):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=lip).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    return f, ngbd, f.sample(shape)


PROBLEMS = {
    "star2d": (2, ls.Star(), 3, 8),
    "spheres3d": (3, ls.random_spheres(6, seed=5, rmin=0.06, rmax=0.14), 2, 6),
    # 222K nodes, 29 % of the records with hanging faces (a quarter of those
    # with two), hanging faces at wall nodes
    "gyroid3d": (3, ls.Gyroid(), 4, 6),
}


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_second_derivatives_match_oracle(cuda, name):
    import torch
    say(f"[gpu] second derivatives {name}")
    dim, shape, lmin, lmax = PROBLEMS[name]
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    g = torch.from_numpy(phi0).to(cuda)
    dd = ngbd.second_derivatives_central(g).cpu().numpy()
    ref = O.second_derivatives(ngbd.tables, phi0)
    assert np.array_equal(dd, ref), rel_err(dd, ref)


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_tiled_encoding_built(cuda, name):
    dim, shape, lmin, lmax = PROBLEMS[name]
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    info = ngbd.device().tile_info()
    say(f"[gpu] tile info {name}: {info}")
    assert info["tiled"], info
    assert info["n_records"] >= ngbd.num_nodes and info["n_irregular"] > 0


@pytest.mark.parametrize("name", sorted(PROBLEMS))
@pytest.mark.parametrize("orders", [(2, 2), (2, 1), (1, 2), (1, 1)])
@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_reinit_fp64_matches_oracle(cuda, name, orders, path):
    import torch
    dim, shape, lmin, lmax = PROBLEMS[name]
    say(f"[gpu] reinit fp64 {name} orders={orders} path={path}")
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    it = 12
    ref = O.reinit(ngbd.tables, phi0, orders[0], orders[1], it)
    ls_ = C.LevelSet(ngbd)
    ngbd.device().set_path(path)
    phi = torch.from_numpy(phi0).to(cuda)
    ls_.reinitialize_within_range_of_phi_0(phi, orders[0], orders[1], C.DBL_MAX, -C.DBL_MAX, it)
    out = phi.cpu().numpy()
    say(f"    max |gpu - oracle| = {np.max(np.abs(out - ref)):.3e}")
    assert np.array_equal(out, ref), rel_err(out, ref)


@pytest.mark.parametrize("dim,box_max", [(2, [2.0, 1.0]), (3, [2.0, 1.0, 0.5]), (2, [3.0, 3.0]), (2, [3.0, 1.0]),
                                          (3, [1.5, 1.0, 0.75])],
                         ids=["rect2d", "brick3d", "square_3", "rect2d_3to1", "brick3d_nonpow2"])
@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_noncubic_box_matches_oracle(cuda, dim, box_max, path):
    """A single tree over a box whose edges differ: cells are not cubes, so a
    hanging arm's transverse sums differ from its length.  Edges in power-of-
    two ratios are tiled (one scale per axis); a cubic box of any edge is
    tiled too; non-cubic cells on other edges (3:1, 1.5:1:0.75) keep the
    grid-wide tables.  FP64 gives the oracle's bits on every path (on edges
    that are not powers of two it runs the grid-wide kernels)."""
    import torch
    say(f"[gpu] non-cubic box {box_max} path={path}")
    shape = ls.Sphere([0.45 * e for e in box_max], 0.3 * min(box_max))
    f = C.Forest(dim, [0.0] * dim, box_max).refine_levelset(shape, 2, 7 if dim == 2 else 6, lip=2.4).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    phi0 = f.sample(shape)
    info = ngbd.device().tile_info()
    dyadic = all(float(np.log2(e)).is_integer() for e in box_max)
    assert info["tiled"] == (dyadic or len(set(box_max)) == 1), info
    ngbd.device().set_path(path)
    assert ngbd.device().path == (path if info["tiled"] else "two_pass")
    it = 10
    ref = O.reinit(ngbd.tables, phi0, 2, 2, it)
    phi = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd).reinitialize_within_range_of_phi_0(phi, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    out = phi.cpu().numpy()
    say(f"    max |gpu - oracle| = {np.max(np.abs(out - ref)):.3e}")
    assert np.array_equal(out, ref), rel_err(out, ref)
    h = float(np.min(f.dxyz_min()))
    mixed = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd, precision="mixed").reinitialize_within_range_of_phi_0(mixed, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    band = np.abs(out) < 3 * h
    assert np.max(np.abs(mixed.cpu().numpy() - out)[band]) <= 1e-3 * h


NONDYADIC = {
    "2pi_square": (2, [0.0, 0.0], [2 * np.pi, 2 * np.pi], ls.Sphere((np.pi, 3.0), 2.0), 7),
    "rect_0.3x0.1": (2, [-0.1, 0.0], [0.2, 0.1], ls.Sphere((0.05, 0.05), 0.04), 7),
    "cube_0.3": (3, [0.0, 0.0, 0.0], [0.3, 0.3, 0.3], ls.Sphere((0.14, 0.15, 0.16), 0.08), 5),
}


@pytest.mark.parametrize("name", sorted(NONDYADIC))
def test_nondyadic_box_matches_oracle(cuda, name):
    """Boxes whose cell sizes FP32 does not hold exactly ([0, 2 pi]^2 and
    the like): the stencil keeps an FP64 copy of the distances and no tiled
    encoding; second derivatives, reinitialization and advection in FP64
    give the oracle's bits, the mixed path stays within 1e-3 h."""
    import torch
    dim, lo, hi, shape, lmax = NONDYADIC[name]
    say(f"[gpu] non-dyadic box {name}")
    f = C.Forest(dim, lo, hi).refine_levelset(shape, 2, lmax, lip=2.4).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    tab = ngbd.tables
    assert np.any(tab.d.astype(np.float32).astype(np.float64) != tab.d)
    phi0 = f.sample(shape)
    assert not ngbd.device().tile_info()["tiled"]
    g = torch.from_numpy(phi0).to(cuda)
    dd = ngbd.second_derivatives_central(g).cpu().numpy()
    assert np.array_equal(dd, O.second_derivatives(tab, phi0)), rel_err(dd, O.second_derivatives(tab, phi0))
    it = 10
    ref = O.reinit(tab, phi0, 2, 2, it)
    phi = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd).reinitialize_within_range_of_phi_0(phi, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    out = phi.cpu().numpy()
    say(f"    {tab.n_nodes} nodes, max |gpu - oracle| = {np.max(np.abs(out - ref)):.3e}")
    assert np.array_equal(out, ref), rel_err(out, ref)
    h = float(np.min(f.dxyz_min()))
    mixed = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd, precision="mixed").reinitialize_within_range_of_phi_0(mixed, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    band = np.abs(out) < 3 * h
    assert np.max(np.abs(mixed.cpu().numpy() - out)[band]) <= 1e-3 * h
    _, xyz = f.nodes()
    vn = 0.5 + np.sin(4 * xyz[:, 0]) * np.cos(3 * xyz[:, 1])
    p = torch.from_numpy(phi0).to(cuda)
    dt = C.LevelSet(ngbd).advect_in_normal_direction(torch.from_numpy(vn).to(cuda), p)
    assert dt == O.advect_cfl_dt(tab, vn)
    assert np.array_equal(p.cpu().numpy(), O.advect(tab, phi0, vn, dt))


BRICKS = {
    "2x1_periodic_x": (2, ([0.0, 0.0], [2.0, 1.0]), [2, 1], [1, 0], ls.Sphere((0.1, 0.5), 0.3), 2, 7),
    "periodic_xy": (2, ([0.0, 0.0], [1.0, 1.0]), [1, 1], [1, 1], ls.Star(center=(0.05, 0.95), r0=0.3), 2, 7),
    "3x2_box": (2, ([-1.0, 0.0], [2.0, 2.0]), [3, 2], [0, 0], ls.Sphere((0.0, 1.0), 0.6), 2, 6),
    "3d_2x1x1_periodic_y": (3, ([0.0, 0.0, 0.0], [2.0, 1.0, 1.0]), [2, 1, 1], [0, 1, 0],
                            ls.Sphere((1.0, 0.1, 0.5), 0.3), 1, 5),
    "3d_periodic_xyz": (3, ([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]), [1, 2, 2], [1, 1, 1],
                        ls.Sphere((0.05, 0.5, 0.95), 0.3), 1, 5),
}


@pytest.mark.parametrize("name", sorted(BRICKS))
@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_brick_forest_matches_oracle(cuda, name, path):
    """Tables of bricks of trees, periodic or not (arms that wrap around the
    domain, neighbours across tree faces): FP64 on both paths gives the
    oracle's bits, mixed stays within 1e-3 h of it."""
    import torch
    dim, box, n_trees, periodic, shape, lmin, lmax = BRICKS[name]
    say(f"[gpu] brick {name} path={path}")
    f = C.Forest(dim, box[0], box[1], n_trees=n_trees, periodic=periodic)
    f.refine_levelset(shape, lmin, lmax, lip=2.4).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    phi0 = f.sample(shape)
    info = ngbd.device().tile_info()
    assert info["tiled"], info
    ngbd.device().set_path(path)
    it = 10
    ref = O.reinit(ngbd.tables, phi0, 2, 2, it)
    phi = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd).reinitialize_within_range_of_phi_0(phi, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    out = phi.cpu().numpy()
    say(f"    {ngbd.tables.n_nodes} nodes, max |gpu - oracle| = {np.max(np.abs(out - ref)):.3e}")
    assert np.array_equal(out, ref), rel_err(out, ref)
    h = float(np.min(f.dxyz_min()))
    mixed = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd, precision="mixed").reinitialize_within_range_of_phi_0(mixed, 2, 2, C.DBL_MAX, -C.DBL_MAX, it)
    band = np.abs(out) < 3 * h
    assert np.max(np.abs(mixed.cpu().numpy() - out)[band]) <= 1e-3 * h


@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_reinit_limits_and_mask(cuda, path):
    import torch
    dim, shape, lmin, lmax = PROBLEMS["star2d"]
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    ngbd.device().set_path(path)
    _, xyz = f.nodes()
    lsr = C.LevelSet(ngbd)
    say("[gpu] limits")
    phi = torch.from_numpy(phi0).to(cuda)
    lsr.reinitialize_2nd_order_above_threshold(phi, -0.02, 7, 0.05)
    ref = O.reinit(ngbd.tables, phi0, 2, 2, 7, limit_high=0.05, limit_low=-0.02)
    assert np.array_equal(phi.cpu().numpy(), ref)
    say("[gpu] mask")
    mask = (xyz[:, 1] > 0.45).astype(np.uint8)
    phi = torch.from_numpy(phi0).to(cuda)
    lsr.reinitialize_2nd_order_with_mask(phi, torch.from_numpy(mask).to(cuda), int((mask == 0).sum()), 6)
    ref = O.reinit(ngbd.tables, phi0, 2, 2, 6, mask=mask)
    assert np.array_equal(phi.cpu().numpy(), ref)
    # numpy in / numpy out (host buffers copied through the device)
    p = phi0.copy()
    lsr.reinitialize_1st_order_time_2nd_order_space(p, 4)
    assert np.array_equal(p, O.reinit(ngbd.tables, phi0, 2, 1, 4))


@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_mixed_limits_and_mask_keep_fp64_input(cuda, path):
    """The mixed path computes in FP32 but nodes outside the phi_0 limits end
    at phi_0 and masked nodes at their input, in FP64, bit for bit (as the
    reference leaves them); updated nodes stay within 1e-3 h of FP64."""
    import torch
    dim, shape, lmin, lmax = PROBLEMS["star2d"]
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    ngbd.device().set_path(path)
    h = float(np.min(f.dxyz_min()))
    _, xyz = f.nodes()
    lsr = C.LevelSet(ngbd)
    say(f"[gpu] mixed limits {path}")
    lo, hi = -0.02, 0.05
    out = {}
    for prec in ("fp64", "mixed"):
        p = torch.from_numpy(phi0).to(cuda)
        lsr.reinitialize_within_range_of_phi_0(p, 2, 2, hi, lo, 7, precision=prec)
        out[prec] = p.cpu().numpy()
    # the widened limits of reinitialize_one_iteration (src/my_p4est_level_set.cpp:75-76)
    zdt = ngbd.tables.zero_distance_threshold
    lim_lo, lim_hi = lo - max(zdt, 1e-13 * abs(lo)), hi + max(zdt, 1e-13 * abs(hi))
    outside = ~((lim_lo < phi0) & (phi0 < lim_hi))
    assert outside.sum() > 100
    assert np.array_equal(out["mixed"][outside], phi0[outside])
    band = (np.abs(out["fp64"]) < 3 * h) & ~outside
    assert np.max(np.abs(out["mixed"][band] - out["fp64"][band])) <= 1e-3 * h
    say(f"[gpu] mixed mask {path}")
    mask = (xyz[:, 1] > 0.45).astype(np.uint8)
    p = torch.from_numpy(phi0).to(cuda)
    lsr.reinitialize_2nd_order_with_mask(p, torch.from_numpy(mask).to(cuda), int((mask == 0).sum()), 6,
                                         precision="mixed")
    pm = p.cpu().numpy()
    assert np.array_equal(pm[mask == 0], phi0[mask == 0])
    ref = O.reinit(ngbd.tables, phi0, 2, 2, 6, mask=mask)
    band = (np.abs(ref) < 3 * h) & (mask != 0)
    assert np.max(np.abs(pm[band] - ref[band])) <= 1e-3 * h
    # a mask comes with orders (2, 2) only
    with pytest.raises(ValueError):
        ngbd.device().reinit(p, 1, 2, 3, C.DBL_MAX, -C.DBL_MAX, 0, mask=torch.from_numpy(mask).to(cuda))


def test_invalid_orders_raise(cuda):
    import torch
    f, ngbd, phi0 = problem(2, ls.Star(), 2, 5)
    phi = torch.from_numpy(phi0).to(cuda)
    with pytest.raises(ValueError):
        C.LevelSet(ngbd).reinitialize_within_range_of_phi_0(phi, 3, 1, 1.0, -1.0, 2)


@pytest.mark.parametrize("name", ["star2d.npz", "spheres3d.npz"])
@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_golden_fixtures(cuda, name, path):
    import torch
    say(f"[gpu] golden {name} {path}")
    z = np.load(os.path.join(G, name), allow_pickle=False)
    tab = C.FlatStencil(int(z["dim"]), z["nbr"], z["d"], z["d_tr"], z["wall"], float(z["zdt"]))
    dev = C.DeviceStencil(tab)
    dev.set_path(path)
    phi0 = z["phi0"]
    g = torch.from_numpy(phi0).to(cuda)
    assert np.array_equal(dev.second_derivatives(g).cpu().numpy(), z["dd0"])
    for it in (1, 2, 20):
        p = torch.from_numpy(phi0).to(cuda)
        dev.reinit(p, 2, 2, it, C.DBL_MAX, -C.DBL_MAX, 0)
        assert np.array_equal(p.cpu().numpy(), z[f"rk2_{it}"]), it
    for os_, ot in ((1, 1), (2, 1), (1, 2)):
        p = torch.from_numpy(phi0).to(cuda)
        dev.reinit(p, os_, ot, 5, C.DBL_MAX, -C.DBL_MAX, 0)
        assert np.array_equal(p.cpu().numpy(), z[f"order{os_}{ot}_5"]), (os_, ot)


def test_star2d_lmax9_fp64_matches_oracle(cuda):
    """2D star lmin 4 / lmax 9, 20 RK2 iterations, FP64 (small enough for a one-core oracle run)."""
    import torch
    say("[gpu] 2D star lmin 4 / lmax 9, 20 iterations")
    f, ngbd, phi0 = problem(2, ls.Star(), 4, 9)
    ref = O.reinit(ngbd.tables, phi0, 2, 2, 20)
    phi = torch.from_numpy(phi0).to(cuda)
    C.LevelSet(ngbd).reinitialize_2nd_order(phi, 20)
    out = phi.cpu().numpy()
    assert np.array_equal(out, ref), rel_err(out, ref)


@pytest.mark.parametrize("name", sorted(PROBLEMS))
@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_mixed_precision_error_band(cuda, name, path):
    import torch
    dim, shape, lmin, lmax = PROBLEMS[name]
    say(f"[gpu] mixed vs fp64 {name} {path}")
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    ngbd.device().set_path(path)
    h = float(np.min(f.dxyz_min()))
    lsr = C.LevelSet(ngbd)
    a = torch.from_numpy(phi0).to(cuda)
    b = torch.from_numpy(phi0).to(cuda)
    lsr.reinitialize_2nd_order(a, 20, precision="fp64")
    lsr.reinitialize_2nd_order(b, 20, precision="mixed")
    err = ngbd.device().band_linf(b, a, a, 3 * h)
    an, bn = a.cpu().numpy(), b.cpu().numpy()
    band = np.abs(an) < 3 * h
    assert abs(err - np.max(np.abs(an[band] - bn[band]))) <= 1e-18
    say(f"    Linf(mixed - fp64, |phi|<3h) = {err:.3e} = {err / h:.3e} h")
    assert err <= 1e-3 * h
    nz = phi0 != 0
    assert np.all(np.sign(bn[nz]) == np.sign(phi0[nz]))


def batch_trees(n, seed=17):
    rng = np.random.default_rng(seed)
    tabs, phis = [], []
    for _ in range(n):
        shape = ls.random_star(rng)
        lmax = int(rng.integers(5, 8))
        f = C.Forest(2).refine_levelset(shape, 3, lmax, lip=2.4).build_nodes()
        tabs.append(C.flat_stencil_from_forest(f))
        phis.append(f.sample(shape))
    return tabs, phis


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("schedule", [(256, 1), (512, 0), (1024, 0)],
                         ids=["table_kernel", "tile_kernel", "tile_kernel_1024"])
def test_batch_matches_single_tree_path(cuda, precision, schedule):
    import torch
    say(f"[gpu] batch vs single {precision} schedule {schedule}")
    tabs, phis = batch_trees(24)
    batch = C.TreeBatch(tabs)
    batch.set_schedule(*schedule)
    phi = torch.from_numpy(np.concatenate(phis)).to(cuda)
    batch.reinitialize_2nd_order(phi, 9, precision=precision)
    out = phi.cpu().numpy()
    for k, (tab, p0) in enumerate(zip(tabs, phis)):
        dev = C.DeviceStencil(tab)
        p = torch.from_numpy(p0).to(cuda)
        dev.reinit(p, 2, 2, 9, C.DBL_MAX, -C.DBL_MAX, 0 if precision == "fp64" else 1)
        seg = out[batch.offsets[k]:batch.offsets[k + 1]]
        assert np.array_equal(seg, p.cpu().numpy()), k
        if precision == "fp64" and k < 4:
            assert np.array_equal(seg, O.reinit(tab, p0, 2, 2, 9)), k


def test_batch_graph_replay_and_autotune(cuda):
    import torch
    say("[gpu] batch graph replay + autotune")
    tabs, phis = batch_trees(40, seed=3)
    batch = C.TreeBatch(tabs)
    base = torch.from_numpy(np.concatenate(phis)).to(cuda)
    eager = base.clone()
    batch.reinitialize_2nd_order(eager, 6, precision="mixed")
    torch.cuda.synchronize()
    phi = base.clone()
    for rep in range(3):
        phi.copy_(base)
        batch.reinitialize_2nd_order(phi, 6, precision="mixed", graph=True)
        torch.cuda.synchronize()
        assert torch.equal(phi, eager), rep
        # eager work between replays
        tmp = base.clone()
        batch.reinitialize_2nd_order(tmp, 2, precision="fp64")
        torch.cuda.synchronize()
    bt, npt, ms = batch.autotune(base, 6, "mixed")
    say(f"    autotune: block={bt} nodes/thread={npt} {ms:.3f} ms")
    assert batch.schedule() == (bt, npt)
    assert batch.tile_info()["tiled"]
    for bt_, npt_ in ((64, 2), (512, 1), (512, 0), (256, 0)):
        batch.set_schedule(bt_, npt_)
        phi.copy_(base)
        batch.reinitialize_2nd_order(phi, 6, precision="mixed")
        assert torch.equal(phi, eager)


@pytest.mark.parametrize("name", ["star2d", "spheres3d"])
def test_first_derivatives_normals_curvature_match_oracle(cuda, name):
    """first_derivatives_central, compute_normals and compute_mean_curvature
    through the C ABI against the oracle (FP64, same operation order)."""
    import torch
    dim, shape, lmin, lmax = PROBLEMS[name]
    say(f"[gpu] curvature chain {name}")
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    tab = ngbd.tables
    phi = ngbd.device()
    p = torch.from_numpy(phi0).to(cuda)
    fd = phi.first_derivatives(p).cpu().numpy()
    assert rel_err(fd, O.first_derivatives(tab, phi0)) <= REL
    nrm = C.compute_normals(ngbd, p)
    nrm_o = O.normals(tab, phi0)
    assert rel_err(nrm.cpu().numpy(), nrm_o) <= REL
    kap = C.compute_mean_curvature(ngbd, nrm).cpu().numpy()
    assert rel_err(kap, O.mean_curvature(tab, nrm_o)) <= REL


@pytest.mark.parametrize("name", ["star2d", "spheres3d"])
def test_advect_in_normal_direction_matches_oracle(cuda, name):
    """advect_in_normal_direction (both stage velocities) and its CFL time
    step through the C ABI against the oracle (FP64), on the tiled and the
    grid-wide paths."""
    import torch
    dim, shape, lmin, lmax = PROBLEMS[name]
    say(f"[gpu] advect {name}")
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    tab = ngbd.tables
    _, xyz = f.nodes()
    rng = np.random.default_rng(3)
    vn = 0.5 + np.sin(4 * xyz[:, 0]) * np.cos(3 * xyz[:, 1]) + 0.1 * rng.standard_normal(tab.n_nodes)
    vn1 = vn * 1.3 - 0.2
    lsr = C.LevelSet(ngbd)
    v = torch.from_numpy(vn).to(cuda)
    assert ngbd.device().tile_info()["tiled"]
    out = {}
    for path in ("tiled", "two_pass"):
        ngbd.device().set_path(path)
        p = torch.from_numpy(phi0).to(cuda)
        dt = lsr.advect_in_normal_direction(v, p)
        assert dt == O.advect_cfl_dt(tab, vn)
        assert np.array_equal(p.cpu().numpy(), O.advect(tab, phi0, vn, dt)), path
        p2 = torch.from_numpy(phi0).to(cuda)
        lsr.advect_in_normal_direction(v, p2, dt=0.5 * dt, vn_np1=torch.from_numpy(vn1).to(cuda))
        assert np.array_equal(p2.cpu().numpy(), O.advect(tab, phi0, vn, 0.5 * dt, vn_np1=vn1)), path
        out[path] = (p.cpu().numpy(), p2.cpu().numpy())
    ngbd.device().set_path("tiled")
    # the fused tiled stages and the grid-wide kernels: the same bits
    assert np.array_equal(out["tiled"][0], out["two_pass"][0])
    assert np.array_equal(out["tiled"][1], out["two_pass"][1])


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_reinit_graph_replays_equal_eager(cuda, name):
    """casl_reinit_graph: every replay equals the eager call bit for bit, with
    eager work of the same stencil in between; other arguments or another
    stage path record again."""
    import torch
    dim, shape, lmin, lmax = PROBLEMS[name]
    say(f"[gpu] graph replay {name}")
    f, ngbd, phi0 = problem(dim, shape, lmin, lmax)
    ds = ngbd.device()
    p0 = torch.from_numpy(phi0).to(cuda)
    for prec in (0, 1):
        for n_iter in (7, 3):
            eager = p0.clone()
            ds.reinit(eager, 2, 2, n_iter, C.DBL_MAX, -C.DBL_MAX, prec)
            g = torch.empty_like(p0)
            other = p0.clone()
            for rep in range(3):
                g.copy_(p0)
                ds.reinit(g, 2, 2, n_iter, C.DBL_MAX, -C.DBL_MAX, prec, graph=True)
                torch.cuda.synchronize()
                assert torch.equal(g, eager), (prec, n_iter, rep)
                # eager work on the same stencil (same scratch) between replays
                other.copy_(p0)
                ds.reinit(other, 2, 1, 2, C.DBL_MAX, -C.DBL_MAX, 1 - prec)
                torch.cuda.synchronize()
    ds.set_path("two_pass")
    eager = p0.clone()
    ds.reinit(eager, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, 0)
    g = p0.clone()
    ds.reinit(g, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, 0, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(g, eager)
    ds.set_path("tiled")
    with pytest.raises(C.CaslError):
        ds.reinit(g, 3, 2, 4, C.DBL_MAX, -C.DBL_MAX, 0, graph=True)
