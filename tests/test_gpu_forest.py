"""Grid construction on the device (casl_device_grid_*) against the host
forest (Forest.refine_levelset + build_nodes + construct_neighbors):

  * the same number of leaves, the same finest level, the same nodes in the
    same order (unclamped logical coordinates);
  * the construct_neighbors tables (neighbour ids, face and transverse
    distances, wall bits, zero-distance threshold) bit for bit;
  * sample_cf_on_nodes: bit for bit for the sphere and plane level sets
    (correctly rounded arithmetic only), within 4 ulp of |phi| scale for the
    level sets that call sin, cos or atan2;
  * reinitialize_2nd_order on the device-built stencil equals the host-built
    one bit for bit;
  * the tiled encoding of the device-built stencil, made on the device from
    tables that never leave it, equals the host builder's encoding of the
    host tables byte for byte (cubic and non-cubic cells);
  * the level sets refine the measured workloads (star2d, spheres3d at full
    size, gyroid3d at lmax 9) to the same grids.
"""
import time

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from casl_p4est_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def host_grid(dim, phi, lmin, lmax, lip, band, box_min=None, box_max=None):
    f = C.Forest(dim, box_min, box_max).refine_levelset(phi, lmin, lmax, lip=lip, band=band).build_nodes()
    return f, C.flat_stencil_from_forest(f)


def assert_same_grid(f, tab, g):
    assert g.sizes() == f.sizes()
    assert np.array_equal(g.node_ixyz(), f.nodes()[0])
    dt = g.tables()
    assert dt.zero_distance_threshold == tab.zero_distance_threshold
    assert np.array_equal(dt.nbr, tab.nbr)
    assert np.array_equal(dt.d.view(np.uint64), tab.d.view(np.uint64))
    assert np.array_equal(dt.d_tr.view(np.uint64), tab.d_tr.view(np.uint64))
    assert np.array_equal(dt.wall, tab.wall)


EXACT = [
    # name, dim, level set, lmin, lmax, lip, band, box_min, box_max
    ("circle2d", 2, ls.Sphere([0.5, 0.5], 0.3), 2, 8, 1.2, 0.0, None, None),
    ("circle2d_box", 2, ls.Sphere([0.1, -0.2], 0.45), 1, 7, 2.4, 1.0, [-1.0, -1.0], [1.0, 1.0]),
    ("circle2d_offcentre_box", 2, ls.Sphere([0.0, 0.0], 0.7), 3, 9, 1.2, 0.0, [-0.5, -1.0], [1.5, 1.0]),
    ("plane2d", 2, ls.Plane([0.6, -0.8], 0.05), 2, 10, 1.2, 0.0, None, None),
    ("uniform2d", 2, ls.Sphere([0.5, 0.5], 0.3), 5, 5, 1.2, 0.0, None, None),
    ("lmin_above_lmax2d", 2, ls.Sphere([0.5, 0.5], 0.3), 6, 3, 1.2, 0.0, None, None),
    ("sphere3d", 3, ls.Sphere([0.45, 0.5, 0.55], 0.3), 2, 7, 1.2, 0.0, None, None),
    ("sphere3d_box", 3, ls.Sphere([0.2, 0.1, -0.1], 0.6), 1, 6, 2.4, 0.0, [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]),
    ("plane3d", 3, ls.Plane([0.48, 0.6, -0.64], -0.1), 1, 7, 1.2, 0.0, None, None),
    ("spheres3d_unmodulated", 3, ls.Spheres([[0.3, 0.3, 0.3], [0.62, 0.55, 0.6]], [0.12, 0.2], 0.0),
     2, 7, 2.4, 0.0, None, None),
]


@pytest.mark.parametrize("case", EXACT, ids=[c[0] for c in EXACT])
def test_device_grid_equals_host(cuda, case):
    import torch
    name, dim, phi, lmin, lmax, lip, band, bmin, bmax = case
    say(f"[grid] {name}")
    f, tab = host_grid(dim, phi, lmin, lmax, lip, band, bmin, bmax)
    g = C.DeviceGrid(dim, phi, lmin, lmax, lip=lip, band=band, box_min=bmin, box_max=bmax)
    assert_same_grid(f, tab, g)
    hs = f.sample(phi)
    ds = g.sample(phi).cpu().numpy()
    assert np.array_equal(ds.view(np.uint64), hs.view(np.uint64))
    # the device-built stencil runs the same reinitialization
    a = torch.from_numpy(hs).to(cuda)
    b = a.clone()
    C.DeviceStencil(tab).reinit(a, 2, 2, 5, C.DBL_MAX, -C.DBL_MAX, 0)
    g.device().reinit(b, 2, 2, 5, C.DBL_MAX, -C.DBL_MAX, 0)
    assert torch.equal(a, b)
    ok, why = g.device().tile_check(tab)
    assert ok, why


@pytest.mark.parametrize("dim,box_max", [(2, [2.0, 1.0]), (3, [2.0, 1.0, 0.5]), (3, [1.5, 1.0, 0.75])],
                         ids=["rect2d", "brick3d", "brick3d_nonpow2"])
def test_device_tiling_noncubic_cells(cuda, dim, box_max):
    """Cells that are not cubes: on edges in power-of-two ratios the device
    builds the same tiled encoding as the host (one scale per axis); on
    others (1.5:1:0.75) neither builds one.  The device-built stencil gives
    the host-built one's results on both paths."""
    import torch
    phi = ls.Sphere([0.45 * e for e in box_max], 0.3 * min(box_max))
    f, tab = host_grid(dim, phi, 2, 7 if dim == 2 else 6, 2.4, 0.0, [0.0] * dim, box_max)
    g = C.DeviceGrid(dim, phi, 2, 7 if dim == 2 else 6, lip=2.4, box_min=[0.0] * dim, box_max=box_max)
    ds = g.device()
    info = ds.tile_info()
    dyadic = all(float(np.log2(e)).is_integer() for e in box_max)
    assert info["tiled"] == dyadic, info
    if dyadic:
        assert info["n_irregular"] > 0, info
    else:
        assert "not cubes" in info["reason"], info
    ok, why = ds.tile_check(tab)
    assert ok, why
    hs = C.DeviceStencil(tab)
    for path in ("tiled", "two_pass"):
        hs.set_path(path)
        ds.set_path(path)
        assert ds.path == (path if dyadic else "two_pass")
        a = torch.from_numpy(f.sample(phi)).to(cuda)
        b = a.clone()
        hs.reinit(a, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, 0)
        ds.reinit(b, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, 0)
        assert torch.equal(a, b), path


def test_device_grid_root_only_fails_like_host(cuda):
    """A grid of the root alone has a wall on both sides of every node: the
    host and the device construction both refuse it."""
    phi = ls.Sphere([0.5, 0.5], 10.0)  # far from every corner: no refinement
    f = C.Forest(2).refine_levelset(phi, 0, 4, lip=1.2).build_nodes()
    with pytest.raises(C.CaslError, match="second-degree"):
        C.flat_stencil_from_forest(f)
    with pytest.raises(C.CaslError, match="second-degree"):
        C.DeviceGrid(2, phi, 0, 4, lip=1.2)


def test_device_grid_rejects_bad_arguments(cuda):
    phi = ls.Sphere([0.5, 0.5], 0.3)
    with pytest.raises(C.CaslError):
        C.DeviceGrid(2, phi, 2, 29)  # beyond P4EST_MAXLEVEL - 2 in 2D
    with pytest.raises(C.CaslError):
        C.DeviceGrid(3, ls.Sphere([0.5, 0.5, 0.5], 0.3), 2, 18)
    with pytest.raises(C.CaslError):
        C.DeviceGrid(2, ls.Gyroid(), 2, 6)  # a 3D level set
    with pytest.raises(C.CaslError):
        C.DeviceGrid(2, ls.Callback(lambda x: x[0], 2), 2, 6)  # host callbacks stay on the host
    with pytest.raises(C.CaslError):
        C.DeviceGrid(2, phi, 2, 6, box_min=[0.0, 0.0], box_max=[1.0, 0.0])


TRANSCENDENTAL = [
    ("modulated2d", 2, ls.Modulated((0.0, 0.0), 0.5, 0.5), 3, 9, 1.2, 0.0, [-1.0, -1.0], [1.0, 1.0]),
    ("star2d_small", 2, ls.Star(), 4, 8, 2.4, 0.0, None, None),
    ("spheres3d_small", 3, ls.random_spheres(16, seed=7), 3, 7, 2.4, 0.0, None, None),
    ("gyroid3d_small", 3, ls.Gyroid(), 3, 7, 2.4, 0.0, None, None),
]


@pytest.mark.parametrize("case", TRANSCENDENTAL, ids=[c[0] for c in TRANSCENDENTAL])
def test_device_grid_transcendental_level_sets(cuda, case):
    name, dim, phi, lmin, lmax, lip, band, bmin, bmax = case
    say(f"[grid] {name}")
    f, tab = host_grid(dim, phi, lmin, lmax, lip, band, bmin, bmax)
    g = C.DeviceGrid(dim, phi, lmin, lmax, lip=lip, band=band, box_min=bmin, box_max=bmax)
    assert_same_grid(f, tab, g)
    hs = f.sample(phi)
    ds = g.sample(phi).cpu().numpy()
    scale = float(np.max(np.abs(hs)))
    assert float(np.max(np.abs(ds - hs))) <= 4 * np.finfo(np.float64).eps * scale


@pytest.mark.parametrize("name,lmax", [("star2d", None), ("spheres3d", None), ("gyroid3d", 9)])
def test_device_grid_workloads(cuda, name, lmax):
    cfg = dict(workloads.TREES[name])
    if lmax:
        cfg["lmax"] = lmax
    phi = workloads.tree_shape(name)
    say(f"[grid] workload {name}, lmin {cfg['lmin']} / lmax {cfg['lmax']}")
    t0 = time.time()
    f, tab = host_grid(cfg["dim"], phi, cfg["lmin"], cfg["lmax"], 2.4, 0.0)
    t1 = time.time()
    g = C.DeviceGrid(cfg["dim"], phi, cfg["lmin"], cfg["lmax"], lip=2.4)
    t2 = time.time()
    say(f"    {f.num_nodes} nodes: host {t1 - t0:.2f} s, device {t2 - t1:.2f} s")
    assert_same_grid(f, tab, g)
    t3 = time.time()
    ds = g.device()
    t4 = time.time()
    info = ds.tile_info()
    say(f"    device tiling {t4 - t3:.3f} s: {info}")
    assert info["tiled"]
    ok, why = ds.tile_check(tab)
    assert ok, why


def test_device_cache_regrid_loop(cuda):
    """A regrid loop with the device block cache on: grids built from blocks
    that dropped grids released (old contents and all) equal a grid built
    with the cache off, table for table, and reinitialize to the same bits on
    both stage paths; the cache holds the released blocks until trimmed."""
    import torch
    dev = torch.cuda.current_device()
    phi = ls.Sphere([0.45, 0.5, 0.55], 0.3)
    f, tab = host_grid(3, phi, 2, 7, 1.2, 0.0)
    phi0 = torch.from_numpy(f.sample(phi)).to(cuda)
    ref = {}
    g = C.DeviceGrid(3, phi, 2, 7, lip=1.2)
    for path in ("tiled", "two_pass"):
        g.device().set_path(path)
        a = phi0.clone()
        g.device().reinit(a, 2, 2, 5, C.DBL_MAX, -C.DBL_MAX, 1)
        ref[path] = a
    del g
    assert C.device_cache_held(dev) == 0
    C.device_cache_set_limit(16 << 30)
    try:
        for it in range(3):
            say(f"[cache] regrid {it}, held {C.device_cache_held(dev)} bytes")
            g = C.DeviceGrid(3, phi, 2, 7, lip=1.2)
            assert_same_grid(f, tab, g)
            for path in ("tiled", "two_pass"):
                g.device().set_path(path)
                b = phi0.clone()
                g.device().reinit(b, 2, 2, 5, C.DBL_MAX, -C.DBL_MAX, 1)
                assert torch.equal(ref[path], b), (it, path)
            del g
            assert C.device_cache_held(dev) > 0
        C.device_cache_trim()
        assert C.device_cache_held(dev) == 0
    finally:
        C.device_cache_set_limit(0)
        C.device_cache_trim()
