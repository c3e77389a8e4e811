"""Python side of the oracle (TEST INFRASTRUCTURE ONLY: imported by the tests,
the smoke check and the benchmark's CPU-baseline leg, never by the product).

1. ctypes bindings of liboracle.so (oracle.c: the CPU restatement of the
   reference's reinitialization arithmetic).
2. An independent brute-force restatement of the grid construction the
   reference takes from p4est and my_p4est_node_neighbors_t, for small trees:
     - refinement check   refine_levelset_cf      src/my_p4est_refine_coarsen.cpp:17-89
     - node numbering     my_p4est_nodes_new      src/my_p4est_nodes.c:383-446, 614-671
                          with p4est_quadrant_compare (p4est 2.8, p4est_bits.c;
                          not vendored in the reference): the axis whose
                          coordinates differ in the highest bit decides, z
                          before y before x on ties -- restated here as an
                          explicit per-bit comparison, not a key interleave
     - cell search        find_neighbor_cell_of_node src/my_p4est_hierarchy.cpp:569-601
                          (restated as a containment test over all leaves)
     - stencils           construct_neighbors     src/my_p4est_node_neighbors.cpp:114-1201
     - bricks, periodicity my_p4est_brick_new     src/my_p4est_tools.c:37-170, and the
                          periodic shifts of construct_neighbors (:345-357 and repeats)
   Pure Python loops: meant for trees of at most a few thousand nodes.
"""
import ctypes
import functools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
sys.path.insert(0, os.path.dirname(HERE))

from casl_p4est_amd import _lib as L  # noqa: E402  (FlatStencil struct layout only)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        h = ctypes.CDLL(LIB_PATH)
        d = ctypes.c_double
        h.oracle_minmod.restype = d
        h.oracle_minmod.argtypes = [d, d]
        h.oracle_interface_location.restype = d
        h.oracle_interface_location.argtypes = [d, d, d, d]
        h.oracle_interface_location_2nd.restype = d
        h.oracle_interface_location_2nd.argtypes = [d, d, d, d, d, d]
        sp = ctypes.POINTER(L.FlatStencil)
        for name in ("oracle_second_derivatives", "oracle_ngbd_quadratic", "oracle_face_linear",
                     "oracle_first_derivatives", "oracle_normals", "oracle_mean_curvature"):
            fn = getattr(h, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [sp, L.c_double_p, L.c_double_p]
        h.oracle_reinit.restype = ctypes.c_int
        h.oracle_reinit.argtypes = [sp, L.c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, L.c_uint8_p]
        h.oracle_advect.restype = ctypes.c_int
        h.oracle_advect.argtypes = [sp, L.c_double_p, L.c_double_p, L.c_double_p, d]
        h.oracle_advect_cfl_dt.restype = d
        h.oracle_advect_cfl_dt.argtypes = [sp, L.c_double_p]
        h.oracle_num_threads.restype = ctypes.c_int
        h.oracle_num_threads.argtypes = []
        h.oracle_set_num_threads.restype = None
        h.oracle_set_num_threads.argtypes = [ctypes.c_int]
        _lib = h
    return _lib


def _dp(a):
    return a.ctypes.data_as(L.c_double_p)


def minmod(a, b):
    return lib().oracle_minmod(a, b)


def interface_location(a, b, fa, fb):
    return lib().oracle_interface_location(a, b, fa, fb)


def interface_location_2nd(a, b, fa, fb, fxxa, fxxb):
    return lib().oracle_interface_location_2nd(a, b, fa, fb, fxxa, fxxb)


def second_derivatives(tables, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty((tables.dim, tables.n_nodes), dtype=np.float64)
    s = tables.struct()
    lib().oracle_second_derivatives(ctypes.byref(s), _dp(f), _dp(out))
    return out


def first_derivatives(tables, f):
    """[dim][N]: first_derivatives_central."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty((tables.dim, tables.n_nodes), dtype=np.float64)
    s = tables.struct()
    lib().oracle_first_derivatives(ctypes.byref(s), _dp(f), _dp(out))
    return out


def normals(tables, phi):
    """[dim][N]: compute_normals."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty((tables.dim, tables.n_nodes), dtype=np.float64)
    s = tables.struct()
    lib().oracle_normals(ctypes.byref(s), _dp(phi), _dp(out))
    return out


def mean_curvature(tables, nrm):
    """[N]: compute_mean_curvature(normals)."""
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    out = np.empty(tables.n_nodes, dtype=np.float64)
    s = tables.struct()
    lib().oracle_mean_curvature(ctypes.byref(s), _dp(nrm), _dp(out))
    return out


def advect(tables, phi, vn, dt, vn_np1=None):
    """advect_in_normal_direction: one RK2 step, returns the new phi."""
    out = np.array(phi, dtype=np.float64, copy=True)
    vn = np.ascontiguousarray(vn, dtype=np.float64)
    v1 = np.ascontiguousarray(vn_np1, dtype=np.float64) if vn_np1 is not None else None
    s = tables.struct()
    if lib().oracle_advect(ctypes.byref(s), _dp(out), _dp(vn), _dp(v1) if v1 is not None else None, dt) != 0:
        raise MemoryError("oracle_advect")
    return out


def advect_cfl_dt(tables, vn):
    vn = np.ascontiguousarray(vn, dtype=np.float64)
    s = tables.struct()
    return lib().oracle_advect_cfl_dt(ctypes.byref(s), _dp(vn))


def ngbd_quadratic(tables, f):
    """[2*dim + 1][N]: corrected face-neighbor values, then f_000."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty((2 * tables.dim + 1, tables.n_nodes), dtype=np.float64)
    s = tables.struct()
    lib().oracle_ngbd_quadratic(ctypes.byref(s), _dp(f), _dp(out))
    return out


def face_linear(tables, g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty((2 * tables.dim, tables.n_nodes), dtype=np.float64)
    s = tables.struct()
    lib().oracle_face_linear(ctypes.byref(s), _dp(g), _dp(out))
    return out


DBL_MAX = sys.float_info.max


def reinit(tables, phi, order_space=2, order_time=2, n_iter=20, limit_high=DBL_MAX, limit_low=-DBL_MAX, mask=None):
    """reinitialize_within_range_of_phi_0 on a copy of phi; returns it."""
    out = np.array(phi, dtype=np.float64, copy=True)
    s = tables.struct()
    m = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        m = mask.ctypes.data_as(L.c_uint8_p)
    rc = lib().oracle_reinit(ctypes.byref(s), _dp(out), int(order_space), int(order_time), int(n_iter),
                             float(limit_high), float(limit_low), m)
    if rc != 0:
        raise ValueError("valid orders of accuracy for space and pseudotime are 1 and 2 only!")
    return out


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()


# --------------------------------------------------------------------------
# Brute-force grid construction (independent restatement, small trees only)
# --------------------------------------------------------------------------
MAXLEVEL = {2: 30, 3: 19}


def quadrant_compare(a, b, D):
    """p4est_quadrant_compare for two points with non-negative coordinates:
    the axis with the highest differing bit decides; ties z > y > x."""
    ex = [a[i] ^ b[i] for i in range(D)]
    if not any(ex):
        return 0
    hb = [e.bit_length() for e in ex]
    if D == 3 and hb[2] >= max(hb[0], hb[1]):
        ax = 2
    elif hb[1] >= hb[0]:
        ax = 1
    else:
        ax = 0
    return -1 if a[ax] < b[ax] else 1


def brute_nodes(leaves_xyz, leaves_lvl, D):
    """Clamped corner nodes of all leaves, unique, in p4est order.
    Returns (list of unclamped integer coords, dict clamped-coord -> index)."""
    R = 1 << MAXLEVEL[D]
    pts = set()
    for q, lv in zip(leaves_xyz, leaves_lvl):
        h = 1 << (MAXLEVEL[D] - int(lv))
        for k in range(1 << D):
            p = tuple(min(int(q[a]) + ((k >> a) & 1) * h, R - 1) for a in range(D))
            pts.add(p)
    order = sorted(pts, key=functools.cmp_to_key(lambda a, b: quadrant_compare(a, b, D)))
    index = {p: i for i, p in enumerate(order)}
    unclamped = [tuple(R if c == R - 1 else c for c in p) for p in order]
    return unclamped, index


def brute_stencil(leaves_xyz, leaves_lvl, D, box_min=None, box_max=None):
    """construct_neighbors for every node of the given leaves, by direct
    geometric search.  Returns (nbr [F][C][N], d [F][N], d_tr [F][T][N], wall [N])."""
    bmin = [0.0] * D if box_min is None else list(box_min)
    bmax = [1.0] * D if box_max is None else list(box_max)
    size = [bmax[a] - bmin[a] for a in range(D)]
    ML = MAXLEVEL[D]
    R = 1 << ML
    leaves_xyz = np.asarray(leaves_xyz, dtype=np.int64)
    leaves_lvl = np.asarray(leaves_lvl, dtype=np.int64)
    lens = np.left_shift(np.int64(1), ML - leaves_lvl)
    nodes, index = brute_nodes(leaves_xyz, leaves_lvl, D)
    N = len(nodes)
    F, C, T = 2 * D, 1 << (D - 1), 2 * (D - 1)

    def corner_index(leaf, k):
        h = int(lens[leaf])
        p = tuple(min(int(leaves_xyz[leaf][a]) + ((k >> a) & 1) * h, R - 1) for a in range(D))
        return index[p]

    def cell_at(p):
        """Leaf containing logical point p (strictly inside the domain), or -1."""
        for a in range(D):
            if p[a] < 0 or p[a] > R:
                return -1
        inside = np.ones(len(leaves_xyz), dtype=bool)
        for a in range(D):
            inside &= (leaves_xyz[:, a] <= p[a]) & (p[a] < leaves_xyz[:, a] + lens)
        hits = np.nonzero(inside)[0]
        assert len(hits) == 1, f"point {p} in {len(hits)} leaves"
        return int(hits[0])

    nbr = np.full((F, C, N), -1, dtype=np.int32)
    d = np.zeros((F, N))
    dtr = np.zeros((F, T, N))
    wall = np.zeros(N, dtype=np.uint8)

    def transverse(a):
        if D == 2:
            return [1 - a]
        return [b for b in range(3) if b != a]

    def coord(v, a):
        return size[a] * (v / float(R)) + bmin[a]

    def fill(n, f, leaf, side):
        a = f >> 1
        tr = transverse(a)
        for c in range(C):
            k = side << a
            k |= (c & 1) << tr[0]
            if D == 3:
                k |= ((c >> 1) & 1) << tr[1]
            nbr[f, c, n] = corner_index(leaf, k)
        qh = int(lens[leaf]) / float(R)
        for t, ax in enumerate(tr):
            thr = size[ax] * 1.0 / float(R)
            X = coord(nodes[n][ax], ax)
            QX = size[ax] * (int(leaves_xyz[leaf][ax]) / float(R)) + bmin[ax]
            m0 = (X - QX) if abs(X - QX) > thr else 0.0
            p0 = (size[ax] * qh - m0) if abs(size[ax] * qh - m0) > thr else 0.0
            dtr[f, 2 * t, n] = m0
            dtr[f, 2 * t + 1, n] = p0

    for n, p in enumerate(nodes):
        cells = {}
        for dirv in range(1 << D):
            q = tuple(p[a] + (1 if (dirv >> a) & 1 else -1) for a in range(D))
            cells[dirv] = cell_at(q)
        lvl = {k: (-1 if v < 0 else int(leaves_lvl[v])) for k, v in cells.items()}
        walls = []
        for f in range(F):
            a, s = f >> 1, f & 1
            cands = [k for k in range(1 << D) if ((k >> a) & 1) == s]
            if all(cells[k] < 0 for k in cands):
                walls.append(f)
                continue
            best = cands[0]
            for k in cands[1:]:
                if lvl[best] < lvl[k]:
                    best = k
            leaf = cells[best]
            d[f, n] = size[a] * (int(lens[leaf]) / float(R))
            fill(n, f, leaf, s)
        for f in walls:
            a, s = f >> 1, f & 1
            so = 1 - s
            cands = [k for k in range(1 << D) if ((k >> a) & 1) == so]
            best, lbest = None, -1
            for k in cands:
                if lbest < lvl[k]:
                    best, lbest = k, lvl[k]
            qmin = cells[best]
            corner = so << a
            for b in range(D):
                if b != a:
                    corner |= (0 if (best >> b) & 1 else 1) << b
            ntmp = corner_index(qmin, corner)
            pt = nodes[ntmp]
            q = tuple(pt[b] + (1 if (best >> b) & 1 else -1) for b in range(D))
            qtmp = cell_at(q)
            assert qtmp >= 0
            d[f, n] = -d[f ^ 1, n] - size[a] * (int(lens[qtmp]) / float(R))
            fill(n, f, qtmp, so)
        w = 0
        for a in range(D):
            if p[a] == 0:
                w |= 1 << (2 * a)
            if p[a] in (R - 1, R):
                w |= 1 << (2 * a + 1)
        wall[n] = w
    return nbr, d, dtr, wall, nodes


def brick_geometry(D, box_min, box_max, n_trees):
    """Tree vertex coordinates of my_p4est_brick_new (src/my_p4est_tools.c:
    138-168): tree i spans [xyz_min + i dx, xyz_min + (i + 1) dx] with
    dx = (xyz_max - xyz_min) / n.  Returns per-axis lists (vertex min,
    extent = vertex max - vertex min) by brick position, and the domain
    corners as the vertices give them."""
    vmin, ext, dom = [], [], []
    for a in range(D):
        dx = (float(box_max[a]) - float(box_min[a])) / float(n_trees[a])
        lo = [float(box_min[a]) + i * dx for i in range(n_trees[a])]
        hi = [float(box_min[a]) + (i + 1) * dx for i in range(n_trees[a])]
        vmin.append(lo)
        ext.append([h - l for h, l in zip(hi, lo)])
        dom.append((float(box_min[a]) + 0 * dx, float(box_min[a]) + n_trees[a] * dx))
    return vmin, ext, dom


def brute_stencil_brick(leaves_xyz, leaves_lvl, leaf_tree, tree_pos, D, box_min, box_max, n_trees, periodic):
    """construct_neighbors on a brick of trees, periodic or not per axis
    (src/my_p4est_node_neighbors.cpp:114-1201, the periodic shifts :345-357
    and their repeats per face), by direct geometric search in brick
    coordinates (tree position * ROOT_LEN + logical coordinate):

      - a node's owner is found by listing every (tree, coordinate) that
        holds it, periodic images included, and keeping the least by tree
        number then p4est_quadrant_compare of the clamped coordinates
        (p4est_node_canonicalize's rule);
      - nodes are ordered by owner tree, then p4est_quadrant_compare;
      - the cell of a perturbed point is the one leaf whose brick box holds
        it (off a non-periodic boundary: none; across a periodic one the
        point wraps).

    Returns (nbr, d, d_tr, wall, nodes [(tree, unclamped coords)])."""
    ML = MAXLEVEL[D]
    R = 1 << ML
    F, C, T = 2 * D, 1 << (D - 1), 2 * (D - 1)
    leaves_xyz = [tuple(int(v) for v in q) for q in leaves_xyz]
    leaves_lvl = [int(v) for v in leaves_lvl]
    leaf_tree = [int(v) for v in leaf_tree]
    tree_pos = [tuple(int(v) for v in p[:D]) for p in tree_pos]
    tree_of = {p: t for t, p in enumerate(tree_pos)}
    span = [n_trees[a] * R for a in range(D)]
    vmin, ext, dom = brick_geometry(D, box_min, box_max, n_trees)
    lens = [1 << (ML - lv) for lv in leaves_lvl]
    gleaf = [tuple(tree_pos[t][a] * R + q[a] for a in range(D)) for q, t in zip(leaves_xyz, leaf_tree)]

    def fold(g):
        return tuple(g[a] % span[a] if periodic[a] else g[a] for a in range(D))

    def owner(g):
        per_axis = []
        for a in range(D):
            images = [g[a]] + ([g[a] + span[a]] if periodic[a] and g[a] == 0 else [])
            cands = [(i, v - i * R) for v in images for i in range(n_trees[a]) if i * R <= v <= (i + 1) * R]
            per_axis.append(cands)
        best = None
        for combo in __import__("itertools").product(*per_axis):
            t = tree_of[tuple(c[0] for c in combo)]
            loc = tuple(c[1] for c in combo)
            clamped = tuple(min(v, R - 1) for v in loc)
            if best is None or t < best[0] or (t == best[0] and quadrant_compare(clamped, best[2], D) < 0):
                best = (t, loc, clamped)
        return best

    pts = {}
    for gq, ln in zip(gleaf, lens):
        for k in range(1 << D):
            g = fold(tuple(gq[a] + ((k >> a) & 1) * ln for a in range(D)))
            if g not in pts:
                pts[g] = owner(g)
    order = sorted(pts.items(), key=functools.cmp_to_key(
        lambda x, y: (x[1][0] - y[1][0]) if x[1][0] != y[1][0] else quadrant_compare(x[1][2], y[1][2], D)))
    index = {g: i for i, (g, _) in enumerate(order)}
    nodes = [(own[0], own[1]) for _, own in order]
    gnode = [g for g, _ in order]
    N = len(order)

    def corner_index(leaf, k):
        return index[fold(tuple(gleaf[leaf][a] + ((k >> a) & 1) * lens[leaf] for a in range(D)))]

    def cell_at(p):
        p = list(p)
        for a in range(D):
            if periodic[a]:
                p[a] %= span[a]
            elif p[a] < 0 or p[a] > span[a]:
                return -1
        hits = [l for l in range(len(gleaf))
                if all(gleaf[l][a] <= p[a] < gleaf[l][a] + lens[l] for a in range(D))]
        assert len(hits) == 1, f"point {p} in {len(hits)} leaves"
        return hits[0]

    def coord(n, a):
        t, loc = nodes[n]
        i = tree_pos[t][a]
        return ext[a][i] * (loc[a] / float(R)) + vmin[a][i]

    def transverse(a):
        return [1 - a] if D == 2 else [b for b in range(3) if b != a]

    nbr = np.full((F, C, N), -1, dtype=np.int32)
    d = np.zeros((F, N))
    dtr = np.zeros((F, T, N))
    wall = np.zeros(N, dtype=np.uint8)

    def arm(leaf, a):
        return ext[a][tree_pos[leaf_tree[leaf]][a]] * (lens[leaf] / float(R))

    def fill(n, f, leaf, side):
        a = f >> 1
        tr = transverse(a)
        for c in range(C):
            k = side << a
            k |= (c & 1) << tr[0]
            if D == 3:
                k |= ((c >> 1) & 1) << tr[1]
            nbr[f, c, n] = corner_index(leaf, k)
        qh = lens[leaf] / float(R)
        i = tree_pos[leaf_tree[leaf]]
        for t, ax in enumerate(tr):
            e = ext[ax][i[ax]]
            thr = e * 1.0 / float(R)
            X = coord(n, ax)
            QX = e * (leaves_xyz[leaf][ax] / float(R)) + vmin[ax][i[ax]]
            m0 = (X - QX) if abs(X - QX) > thr else 0.0
            if periodic[ax] and X < QX - e * qh / 4.0:
                w = m0 + (dom[ax][1] - dom[ax][0])
                m0 = w if abs(w) > thr else 0.0
            p0 = (e * qh - m0) if abs(e * qh - m0) > thr else 0.0
            dtr[f, 2 * t, n] = m0
            dtr[f, 2 * t + 1, n] = p0

    for n in range(N):
        g = gnode[n]
        cells = {}
        for dirv in range(1 << D):
            cells[dirv] = cell_at(tuple(g[a] + (1 if (dirv >> a) & 1 else -1) for a in range(D)))
        lvl = {k: (-1 if v < 0 else leaves_lvl[v]) for k, v in cells.items()}
        walls = []
        for f in range(F):
            a, s_ = f >> 1, f & 1
            cands = [k for k in range(1 << D) if ((k >> a) & 1) == s_]
            if all(cells[k] < 0 for k in cands):
                walls.append(f)
                continue
            best = cands[0]
            for k in cands[1:]:
                if lvl[best] < lvl[k]:
                    best = k
            leaf = cells[best]
            d[f, n] = arm(leaf, a)
            fill(n, f, leaf, s_)
        for f in walls:
            a, s_ = f >> 1, f & 1
            so = 1 - s_
            cands = [k for k in range(1 << D) if ((k >> a) & 1) == so]
            best, lbest = None, -1
            for k in cands:
                if lbest < lvl[k]:
                    best, lbest = k, lvl[k]
            qmin = cells[best]
            corner = so << a
            for b in range(D):
                if b != a:
                    corner |= (0 if (best >> b) & 1 else 1) << b
            pt = gnode[corner_index(qmin, corner)]
            qtmp = cell_at(tuple(pt[b] + (1 if (best >> b) & 1 else -1) for b in range(D)))
            assert qtmp >= 0
            d[f, n] = -d[f ^ 1, n] - arm(qtmp, a)
            fill(n, f, qtmp, so)
        w = 0
        for a in range(D):
            if periodic[a]:
                continue
            if g[a] == 0:
                w |= 1 << (2 * a)
            if g[a] == span[a]:
                w |= 1 << (2 * a + 1)
        wall[n] = w
    return nbr, d, dtr, wall, nodes


def refine_decision(q_xyz, level, D, min_lvl, max_lvl, lip, band_factor, phi, box_min=None, box_max=None,
                    tree_min=None, tree_ext=None):
    """refine_levelset_cf for one quadrant (src/my_p4est_refine_coarsen.cpp:17-89);
    tree_min / tree_ext: the quadrant's tree vertex minimum and extent (a
    brick), else the box."""
    bmin = [0.0] * D if box_min is None else list(box_min)
    bmax = [1.0] * D if box_max is None else list(box_max)
    size = [bmax[a] - bmin[a] for a in range(D)]
    if tree_min is not None:
        bmin, size = list(tree_min), list(tree_ext)
    ML = MAXLEVEL[D]
    R = float(1 << ML)
    if level < min_lvl:
        return True
    if level >= max_lvl:
        return False
    dmin = float(1 << (ML - level)) / R
    dx = [size[a] * dmin for a in range(D)]
    smallest = max(size[0], size[1]) * float(1 << (ML - max_lvl)) / R
    if D == 3:
        smallest = max(smallest, size[2] * float(1 << (ML - max_lvl)) / R)
    d2 = dx[0] * dx[0] + dx[1] * dx[1]
    if D == 3:
        d2 = d2 + dx[2] * dx[2]
    dd = float(np.sqrt(d2))
    x = [size[a] * float(q_xyz[a]) / R + bmin[a] for a in range(D)]
    band = band_factor * smallest
    f = []
    for k in range(1 << D):
        xc = [x[a] + ((k >> a) & 1) * dx[a] for a in range(D)]
        v = float(phi(np.array([xc]))[0])
        f.append(v)
        if abs(v) - band <= 0.5 * lip * dd:
            return True
    if D == 3:
        pairs = [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4), (4, 5), (5, 6), (6, 7)]
    else:
        pairs = [(0, 1), (0, 2), (1, 3), (2, 3)]
    return any(f[i] * f[j] < 0 for i, j in pairs)
