"""casl_p4est_amd: MI355X-native level-set reinitialization (the hot path of
parCASL's my_p4est_level_set_t) with HIP kernels for gfx950.

Host side (C++ in libcasl_reinit.so, driven from here):
    Forest, NodeNeighbors      grid + flattened node stencils (p4est's role)
    LevelSet                   my_p4est_level_set_t reinitialize_* entry points
    TreeBatch                  many small trees in one persistent kernel
Level-set functions for refinement and sampling live in .levelset.
"""
from . import levelset
from ._lib import CaslError
from .batch import TreeBatch
from .forest import (DeviceGrid, FlatStencil, Forest, NodeNeighbors, device_cache_held, device_cache_set_limit,
                     device_cache_trim, flat_stencil_from_forest)
from .level_set import (DBL_MAX, DeviceStencil, LevelSet, compute_mean_curvature, compute_normals,
                        my_p4est_level_set_t, tile_plan)

my_p4est_node_neighbors_t = NodeNeighbors

__all__ = [
    "CaslError", "DBL_MAX", "DeviceGrid", "DeviceStencil", "FlatStencil", "Forest", "LevelSet", "NodeNeighbors", "TreeBatch",
    "compute_mean_curvature", "compute_normals", "device_cache_held", "device_cache_set_limit", "device_cache_trim",
    "flat_stencil_from_forest", "levelset", "my_p4est_level_set_t", "my_p4est_node_neighbors_t", "tile_plan",
]
