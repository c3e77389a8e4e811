"""Synthetic reinitialization workloads (host side): the grids, stencils and
initial level sets that bench.py measures and the size-independent GPU
tests check.  All refine with refine_levelset_cf at lip = 2.4, i.e. where
|phi0(corner)| <= 1.2 diag(cell) or phi0 changes sign
(src/my_p4est_refine_coarsen.cpp:17-89), seed 20261202.

    star2d     2D, lmin 4 / lmax 9, star r = 0.3 + 0.05 cos 5t, phi0 scaled by
               (1 + 0.8 |x - c|^2); 20 iterations
    spheres3d  3D, lmin 5 / lmax 10, min over 64 spheres (centres U[0.2,0.8]^3,
               radii U[0.03,0.08]) scaled by (1 + 0.5 sin 2pi x cos 2pi y); 50 iterations
    gyroid3d   3D, lmin 6 / lmax 11, sin 8pi x cos 8pi y + sin 8pi y cos 8pi z +
               sin 8pi z cos 8pi x - 0.3 scaled by (1 + 0.5 x y z); 200 iterations
    batch2d    4096 independent quadtrees, lmin 3 / lmax U{6..9}, random stars
               (centre U[0.35,0.65]^2, r0 U[0.15,0.3], amplitude U[0,0.08],
               lobes U{3..8}) scaled by (1 + 0.8 |x - c|^2); 20 iterations
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import levelset as ls
from .forest import Forest, NodeNeighbors, flat_stencil_from_forest

SEED = 20261202

TREES = {
    "star2d": dict(dim=2, lmin=4, lmax=9, iters=20,
                   desc="2D quadtree unit square, lmin 4 / lmax 9, star r=0.3+0.05cos5t x (1+0.8|x-c|^2), "
                        "reinitialize_2nd_order 20 iterations"),
    "spheres3d": dict(dim=3, lmin=5, lmax=10, iters=50,
                      desc="3D octree unit cube, lmin 5 / lmax 10, |phi0|<=1.2 diag refinement, 64 spheres "
                           "(seed 20261202) x (1+0.5 sin 2pi x cos 2pi y), reinitialize_2nd_order 50 iterations"),
    "gyroid3d": dict(dim=3, lmin=6, lmax=11, iters=200,
                     desc="3D octree unit cube, lmin 6 / lmax 11, |phi0|<=1.2 diag refinement, gyroid "
                          "sin 8pi x cos 8pi y + sin 8pi y cos 8pi z + sin 8pi z cos 8pi x = 0.3 scaled by "
                          "(1 + 0.5 x y z), reinitialize_2nd_order 200 iterations"),
}


def tree_shape(name):
    if name == "star2d":
        return ls.Star()
    if name == "spheres3d":
        return ls.random_spheres(64, seed=SEED)
    if name == "gyroid3d":
        return ls.Gyroid()
    raise ValueError(f"unknown workload {name!r}")


def build_tree(name, lmax=None, log=None):
    """One tree: dict(dim, lmin, lmax, iters, forest, ngbd, phi0, desc, h)."""
    cfg = dict(TREES[name], name=name)
    shape = tree_shape(name)
    if lmax:
        cfg["lmax"] = lmax
        cfg["desc"] += f" [lmax reduced to {lmax}]"
    t0 = time.time()
    forest = Forest(cfg["dim"]).refine_levelset(shape, cfg["lmin"], cfg["lmax"], lip=2.4)
    t1 = time.time()
    forest.build_nodes()
    t2 = time.time()
    ngbd = NodeNeighbors(forest).init_neighbors()
    t3 = time.time()
    phi0 = forest.sample(shape)
    if log:
        log(f"[workload] {name}: refine {t1 - t0:.1f}s, nodes {t2 - t1:.1f}s, stencils {t3 - t2:.1f}s; "
            f"{forest.num_nodes} nodes, {forest.num_leaves} leaves")
    cfg.update(forest=forest, ngbd=ngbd, phi0=phi0, h=float(np.min(forest.dxyz_min())), shape=shape,
               build_s=dict(refine=t1 - t0, nodes=t2 - t1, stencils=t3 - t2))
    return cfg


def build_batch(n_trees=4096, threads=16, log=None):
    """The batched workload: dict(tabs, phis, h, lmax, iters, desc)."""
    rng = np.random.default_rng(SEED)
    specs = [(ls.random_star(rng), int(rng.integers(6, 10))) for _ in range(n_trees)]

    def one(spec):
        shape, lmax = spec
        f = Forest(2).refine_levelset(shape, 3, lmax, lip=2.4).build_nodes()
        return flat_stencil_from_forest(f), f.sample(shape), float(np.min(f.dxyz_min()))

    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(one, specs))
    if log:
        log(f"[workload] batch2d: {n_trees} trees in {time.time() - t0:.1f}s")
    desc = (f"{n_trees} independent 2D quadtrees (unit square, lmin 3, lmax U{{6..9}}, random stars, "
            f"phi0 x (1+0.8|x-c|^2), seed 20261202), reinitialize_2nd_order 20 iterations, one persistent launch")
    return dict(tabs=[r[0] for r in res], phis=[r[1] for r in res], h=np.array([r[2] for r in res]),
                lmax=np.array([s[1] for s in specs]), iters=20, desc=desc)
