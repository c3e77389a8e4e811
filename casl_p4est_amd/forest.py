"""Host-side grid objects: the role p4est, my_p4est_nodes_new and
my_p4est_node_neighbors_t play in the reference.

    Forest            a brick of trees on a box, periodic or not per axis
                      (my_p4est_brick_new, src/my_p4est_tools.c:37-170, +
                      my_p4est_new + my_p4est_refine)
    Forest.refine_levelset
                      refine_levelset_cf applied max_lvl times
                      (src/my_p4est_refine_coarsen.cpp:17-89,
                      examples/reinitialization/main_2d.cpp:165-173)
    Forest.build_nodes my_p4est_nodes_new (src/my_p4est_nodes.c:383-446, 614-671)
    NodeNeighbors     my_p4est_node_neighbors_t: init_neighbors builds the
                      flattened quad_neighbor_nodes_of_node_t tables
                      (src/my_p4est_node_neighbors.cpp:33-61, 114-1201) and
                      second_derivatives_central runs on the GPU
                      (src/my_p4est_node_neighbors.cpp:1422-1641)
"""
import ctypes

import numpy as np

from . import _lib as L
from .levelset import LevelSet as LevelSetFunction

P4EST_MAXLEVEL = {2: 30, 3: 19}


def device_cache_set_limit(bytes_per_device):
    """Keep device blocks the library releases, up to bytes_per_device per
    GPU, for reuse by its later allocations of the same size
    (casl_device_cache_set_limit); 0, the default, hands them back at once.
    For regrid loops: build a grid on the device, drop the old one, repeat."""
    if int(bytes_per_device) < 0:
        raise ValueError("bytes_per_device must be >= 0")
    L.call("casl_device_cache_set_limit", ctypes.c_size_t(int(bytes_per_device)))


def device_cache_trim():
    """Free every block the device cache holds (casl_device_cache_trim)."""
    L.call("casl_device_cache_trim")


def device_cache_held(device=0):
    """Bytes the device cache holds on one GPU (casl_device_cache_held)."""
    b = ctypes.c_size_t()
    L.call("casl_device_cache_held", int(device), ctypes.byref(b))
    return b.value


class Forest:
    """A forest of quadtrees (dim=2) or octrees (dim=3) on [box_min, box_max]:
    n_trees[a] trees along axis a (default one tree), periodic[a] wrapping
    axis a (default none), as my_p4est_brick_new(n_xyz, xyz_min, xyz_max,
    &brick, periodic)."""

    def __init__(self, dim, box_min=None, box_max=None, n_trees=None, periodic=None):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.box_min = np.asarray(box_min if box_min is not None else [0.0] * dim, dtype=np.float64)
        self.box_max = np.asarray(box_max if box_max is not None else [1.0] * dim, dtype=np.float64)
        self.n_trees = np.asarray(n_trees if n_trees is not None else [1] * dim, dtype=np.int32)
        self.periodic = np.asarray([int(bool(p)) for p in periodic] if periodic is not None else [0] * dim,
                                   dtype=np.int32)
        for v in (self.box_min, self.box_max, self.n_trees, self.periodic):
            if v.shape != (dim,):
                raise ValueError(f"box, trees and periodicity need {dim} entries")
        h = ctypes.c_void_p()
        L.call("casl_forest_new_brick", dim, self.box_min.ctypes.data_as(L.c_double_p),
               self.box_max.ctypes.data_as(L.c_double_p), self.n_trees.ctypes.data_as(L.c_int32_p),
               self.periodic.ctypes.data_as(L.c_int32_p), ctypes.byref(h))
        L.own(self, h, "casl_forest_destroy")
        self._nodes_built = False

    @property
    def handle(self):
        return self._h

    def refine_levelset(self, phi, min_lvl, max_lvl, lip=1.2, band=0.0):
        """refine_levelset_cf with splitting_criteria_cf_t(min_lvl, max_lvl, phi, lip, band)."""
        spec = phi.spec()
        L.call("casl_forest_refine_levelset", self._h, int(min_lvl), int(max_lvl), float(lip), float(band),
               ctypes.byref(spec))
        self._nodes_built = False
        return self

    def build_nodes(self):
        L.call("casl_forest_build_nodes", self._h)
        self._nodes_built = True
        return self

    def sizes(self):
        nl, nn, ml = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        L.call("casl_forest_sizes", self._h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(ml))
        return nl.value, nn.value, ml.value

    @property
    def num_leaves(self):
        return self.sizes()[0]

    @property
    def num_nodes(self):
        return self.sizes()[1]

    @property
    def max_level(self):
        return self.sizes()[2]

    def leaves(self):
        """(xyz [L][dim] logical coords, level [L], local_nodes [L][2^dim] or None)."""
        nl = self.num_leaves
        xyz = np.empty((nl, self.dim), dtype=np.int32)
        lvl = np.empty(nl, dtype=np.int8)
        local = np.empty((nl, 2 ** self.dim), dtype=np.int32) if self._nodes_built else None
        L.call("casl_forest_get_leaves", self._h, xyz.ctypes.data_as(L.c_int32_p),
               lvl.ctypes.data_as(L.c_int8_p), local.ctypes.data_as(L.c_int32_p) if local is not None else None)
        return xyz, lvl, local

    def trees(self):
        """(brick position of every tree [T][dim], tree of every leaf [L],
        canonical tree of every node [N] or None before build_nodes)."""
        nt = ctypes.c_int32()
        L.call("casl_forest_trees", self._h, ctypes.byref(nt), None, None, None)
        pos = np.empty((nt.value, 3), dtype=np.int32)
        leaf = np.empty(self.num_leaves, dtype=np.int32)
        node = np.empty(self.num_nodes, dtype=np.int32) if self._nodes_built else None
        L.call("casl_forest_trees", self._h, None, pos.ctypes.data_as(L.c_int32_p), leaf.ctypes.data_as(L.c_int32_p),
               node.ctypes.data_as(L.c_int32_p) if node is not None else None)
        return pos[:, :self.dim], leaf, node

    def nodes(self):
        """(logical unclamped coords in the node's tree [N][dim] int32,
        physical coords [N][dim])."""
        n = self.num_nodes
        ixyz = np.empty((n, self.dim), dtype=np.int32)
        xyz = np.empty((n, self.dim), dtype=np.float64)
        L.call("casl_forest_get_nodes", self._h, ixyz.ctypes.data_as(L.c_int32_p), xyz.ctypes.data_as(L.c_double_p))
        return ixyz, xyz

    def sample(self, phi):
        """sample_cf_on_nodes: values of a level-set function at the nodes."""
        out = np.empty(self.num_nodes, dtype=np.float64)
        spec = phi.spec()
        L.call("casl_forest_sample", self._h, ctypes.byref(spec), out.ctypes.data_as(L.c_double_p))
        return out

    def dxyz_min(self):
        """Smallest cell sizes per axis (get_dxyz_min)."""
        ml = self.max_level
        return (self.box_max - self.box_min) / self.n_trees / float(2 ** ml)


class DeviceGrid:
    """The grid of Forest(dim).refine_levelset(phi, ...).build_nodes() and its
    construct_neighbors tables, built on the current GPU
    (casl_device_grid_build): refinement level by level, node numbering by a
    device radix sort, neighbours one thread per node.  Same leaves, node order
    and tables as the host Forest whenever phi evaluates to the same bits on
    both (sphere and plane level sets always do)."""

    def __init__(self, dim, phi, min_lvl, max_lvl, lip=1.2, band=0.0, box_min=None, box_max=None):
        from .level_set import _stream_ptr, _torch
        torch = _torch()
        torch.cuda.init()
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.box_min = np.asarray(box_min if box_min is not None else [0.0] * dim, dtype=np.float64)
        self.box_max = np.asarray(box_max if box_max is not None else [1.0] * dim, dtype=np.float64)
        h = ctypes.c_void_p()
        spec = phi.spec()
        L.call("casl_device_grid_build", dim, self.box_min.ctypes.data_as(L.c_double_p),
               self.box_max.ctypes.data_as(L.c_double_p), int(min_lvl), int(max_lvl), float(lip), float(band),
               ctypes.byref(spec), _stream_ptr(torch), ctypes.byref(h))
        L.own(self, h, "casl_device_grid_destroy")
        self._device = None

    def sizes(self):
        nl, nn, ml = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        L.call("casl_device_grid_sizes", self._h, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(ml))
        return nl.value, nn.value, ml.value

    @property
    def num_leaves(self):
        return self.sizes()[0]

    @property
    def num_nodes(self):
        return self.sizes()[1]

    @property
    def max_level(self):
        return self.sizes()[2]

    def dxyz_min(self):
        return (self.box_max - self.box_min) / float(2 ** self.max_level)

    def tables(self):
        """The tables copied to the host, as a FlatStencil."""
        n, D = self.num_nodes, self.dim
        F, C, T = 2 * D, 2 ** (D - 1), 2 * (D - 1)
        nbr = np.empty((F, C, n), dtype=np.int32)
        d = np.empty((F, n), dtype=np.float64)
        d_tr = np.empty((F, T, n), dtype=np.float64)
        wall = np.empty(n, dtype=np.uint8)
        L.call("casl_device_grid_download", self._h, nbr.ctypes.data_as(L.c_int32_p),
               d.ctypes.data_as(L.c_double_p), d_tr.ctypes.data_as(L.c_double_p),
               wall.ctypes.data_as(L.c_uint8_p), None)
        s = L.FlatStencil()
        L.call("casl_device_grid_tables", self._h, ctypes.byref(s))
        return FlatStencil(D, nbr, d, d_tr, wall, s.zero_distance_threshold)

    def node_ixyz(self):
        """Unclamped logical node coordinates [N][dim] (as Forest.nodes()[0])."""
        ixyz = np.empty((self.num_nodes, self.dim), dtype=np.int32)
        L.call("casl_device_grid_download", self._h, None, None, None, None, ixyz.ctypes.data_as(L.c_int32_p))
        return ixyz

    def sample(self, phi, out=None):
        """sample_cf_on_nodes on the device: a float64 tensor on the current GPU."""
        from .level_set import _stream_ptr, _torch
        torch = _torch()
        if out is None:
            out = torch.empty(self.num_nodes, dtype=torch.float64, device="cuda")
        if out.dtype != torch.float64 or not out.is_cuda or out.numel() != self.num_nodes or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 CUDA tensor with one value per node")
        spec = phi.spec()
        L.call("casl_device_grid_sample", self._h, ctypes.byref(spec), ctypes.c_void_p(out.data_ptr()),
               _stream_ptr(torch))
        return out

    def device(self):
        """A DeviceStencil for this grid's tables (casl_device_grid_stencil), created once."""
        if self._device is None:
            from .level_set import DeviceStencil
            h = ctypes.c_void_p()
            L.call("casl_device_grid_stencil", self._h, ctypes.byref(h))
            self._device = DeviceStencil(None, handle=h)
        return self._device


class FlatStencil:
    """Flattened quad_neighbor_nodes_of_node_t tables (numpy views, SoA).

    nbr   [2*dim][2^(dim-1)][N] int32   node_{face}_{mm,pm[,mp,pp]}
    d     [2*dim][N]                    d_{face}
    d_tr  [2*dim][2*(dim-1)][N]         d_{face}_{m0,p0[,0m,0p]}
    wall  [N] uint8                     bit f: node on wall face f
    """

    def __init__(self, dim, nbr, d, d_tr, wall, zero_distance_threshold, owner=None):
        self.dim = dim
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.d = np.ascontiguousarray(d, dtype=np.float64)
        self.d_tr = np.ascontiguousarray(d_tr, dtype=np.float64)
        self.wall = np.ascontiguousarray(wall, dtype=np.uint8)
        self.zero_distance_threshold = float(zero_distance_threshold)
        self.n_nodes = int(self.wall.shape[0])
        self._owner = owner

    def struct(self):
        s = L.FlatStencil()
        s.dim = self.dim
        s.n_nodes = self.n_nodes
        s.nbr = self.nbr.ctypes.data_as(L.c_int32_p)
        s.d = self.d.ctypes.data_as(L.c_double_p)
        s.d_tr = self.d_tr.ctypes.data_as(L.c_double_p)
        s.wall = self.wall.ctypes.data_as(L.c_uint8_p)
        s.zero_distance_threshold = self.zero_distance_threshold
        return s

    def copy(self):
        return FlatStencil(self.dim, self.nbr.copy(), self.d.copy(), self.d_tr.copy(), self.wall.copy(),
                           self.zero_distance_threshold)

    def save(self, path):
        np.savez_compressed(path, dim=self.dim, nbr=self.nbr, d=self.d, d_tr=self.d_tr, wall=self.wall,
                            zdt=self.zero_distance_threshold)

    @staticmethod
    def load(path):
        z = np.load(path, allow_pickle=False)
        return FlatStencil(int(z["dim"]), z["nbr"], z["d"], z["d_tr"], z["wall"], float(z["zdt"]))


def flat_stencil_from_forest(forest):
    """construct_neighbors for every node of a forest (host C++)."""
    s = L.FlatStencil()
    L.call("casl_forest_init_neighbors", forest.handle, ctypes.byref(s))
    dim, n = s.dim, s.n_nodes
    F, C, T = 2 * dim, 2 ** (dim - 1), 2 * (dim - 1)
    nbr = np.ctypeslib.as_array(s.nbr, shape=(F, C, n))
    d = np.ctypeslib.as_array(s.d, shape=(F, n))
    d_tr = np.ctypeslib.as_array(s.d_tr, shape=(F, T, n))
    wall = np.ctypeslib.as_array(s.wall, shape=(n,))
    # views of the forest's own storage; the FlatStencil keeps the forest alive
    return FlatStencil(dim, nbr, d, d_tr, wall, s.zero_distance_threshold, owner=forest)


class NodeNeighbors:
    """my_p4est_node_neighbors_t: flattened stencils + device tables."""

    def __init__(self, forest_or_tables):
        if isinstance(forest_or_tables, Forest):
            self.forest = forest_or_tables
            self.tables = None
        else:
            self.forest = None
            self.tables = forest_or_tables
        self._device = None

    def init_neighbors(self):
        if self.tables is None:
            self.tables = flat_stencil_from_forest(self.forest)
        return self

    @property
    def dim(self):
        return self.tables.dim

    @property
    def num_nodes(self):
        return self.tables.n_nodes

    def device(self):
        """The device copy of the tables (casl_stencil_create), created once."""
        if self._device is None:
            from .level_set import DeviceStencil
            if self.tables is None:
                self.init_neighbors()
            self._device = DeviceStencil(self.tables)
        return self._device

    def release_device(self):
        """Drop this object's reference to the device copy of the tables; its
        device memory is freed once no other reference holds it, and the next
        device() uploads the tables again."""
        self._device = None

    def second_derivatives_central(self, f, fdd=None):
        """fdd[a][n] = d^2f/dx_a^2 at node n, on the GPU (torch float64 tensors)."""
        return self.device().second_derivatives(f, fdd)

    def first_derivatives_central(self, f, fd=None):
        """fd[a][n] = df/dx_a at node n, on the GPU (torch float64 tensors)."""
        return self.device().first_derivatives(f, fd)
