"""CPU oracle for the learned-curvature path: TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg may import this
module, and only as the checker; the product path (casl_p4est_amd) never
calls it.

Restates in numpy:
  * numerical curvature on a uniform patch: normals from central differences
    (compute_normals, src/my_p4est_utils.cpp:1698-1740; central_derivative,
    src/my_p4est_quad_neighbor_nodes_of_node.h:200-206), kappa = div n
    (qnnn::get_curvature, src/my_p4est_quad_neighbor_nodes_of_node.h:2167-2171),
    h*kappa bilinearly interpolated at the projection x - phi n
    (src/my_p4est_curvature_ml.cpp:414-480), and the 3x3 phi/h stencil
    features (src/my_p4est_curvature_ml.h:49-60);
  * the network-input preparation and hybrid blend of
    kml::Curvature::_computeHybridHK (src/my_p4est_curvature_ml.cpp:327-398,
    483-556);
  * the fully connected ReLU network in BF16 with FP32 accumulation: forward,
    backward of the mean squared error and Adam, as the HIP kernels compute
    them (the reference has inference only, src/my_p4est_curvature_ml.cpp:214-286).

How each part is pinned (tests/test_curvature_oracle.py): the float64
backward against central finite differences of the forward; BF16 rounding
against hand-worked bit patterns; normalisation against the rotation and
reflection of linear fields (gradient lands in the first quadrant, four
quarter turns are the identity); the hybrid blend at its end points; the
numerical curvature against circles (h*kappa -> h/R at second order) and
against numpy.gradient for the central differences.
"""
import numpy as np

EPS = 1e-13


# ---------------------------------------------------------------- BF16
def bf16_bits(x):
    """float32 -> bfloat16 bit patterns, round to nearest even (NaN -> 0x7fc0)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, 0x7FC0, r)
    return r.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    """Round float32 values to the nearest bfloat16 (returned as float32)."""
    return bf16_value(bf16_bits(x))


# ---------------------------------------------------------------- network
class Net:
    """Weights W[l] [in][out] and biases b[l] (FP32 masters)."""

    def __init__(self, W, b, mean=None, std=None):
        self.W = [np.asarray(w, dtype=np.float32) for w in W]
        self.b = [np.asarray(v, dtype=np.float32) for v in b]
        d0 = self.W[0].shape[0]
        self.mean = np.zeros(d0, np.float32) if mean is None else np.asarray(mean, np.float32)
        self.std = np.ones(d0, np.float32) if std is None else np.asarray(std, np.float32)

    @staticmethod
    def from_flat(dims, w_flat, b_flat, mean=None, std=None):
        W, b, ow, ob = [], [], 0, 0
        for i, o in zip(dims[:-1], dims[1:]):
            W.append(np.asarray(w_flat[ow:ow + i * o]).reshape(i, o))
            b.append(np.asarray(b_flat[ob:ob + o]))
            ow += i * o
            ob += o
        return Net(W, b, mean, std)


def forward(net, x, emulate_bf16=True):
    """y [n][out] and the stored activations.  emulate_bf16: inputs,
    activations and weights rounded to BF16, products summed in float64 and
    rounded to float32 (the kernels accumulate in FP32), bias added in FP32.
    Otherwise everything is float64."""
    L = len(net.W)
    if emulate_bf16:
        inv = np.float32(1.0) / net.std
        a = bf16((np.asarray(x, np.float32) - net.mean) * inv)
    else:
        a = (np.asarray(x, np.float64) - net.mean.astype(np.float64)) / net.std.astype(np.float64)
    acts = [a]
    for l in range(L):
        if emulate_bf16:
            z = (a.astype(np.float64) @ bf16(net.W[l]).astype(np.float64)).astype(np.float32) + net.b[l]
        else:
            z = a @ net.W[l].astype(np.float64) + net.b[l].astype(np.float64)
        if l < L - 1:
            z = np.maximum(z, 0)
            a = bf16(z) if emulate_bf16 else z
            acts.append(a)
        else:
            y = z
    return y, acts


def backward(net, x, t, emulate_bf16=True):
    """Gradients of mean((y - t)^2) over the n rows: (gW list, gb list, loss)."""
    y, acts = forward(net, x, emulate_bf16)
    n = y.shape[0]
    t = np.asarray(t).reshape(y.shape)
    L = len(net.W)
    if emulate_bf16:
        e = (y - t.astype(np.float32)).astype(np.float32)
        dz32 = np.float32(2.0) * e * np.float32(1.0 / n)
        loss = float(np.sum(e.astype(np.float64) ** 2) / n)
        dz16 = bf16(dz32)
    else:
        e = y - t
        dz32 = 2.0 * e / n
        loss = float(np.sum(e ** 2) / n)
        dz16 = dz32
    gW, gb = [None] * L, [None] * L
    for l in range(L - 1, -1, -1):
        gb[l] = dz32.astype(np.float64).sum(axis=0).astype(np.float32) if emulate_bf16 else dz32.sum(axis=0)
        a = acts[l]
        if emulate_bf16:
            gW[l] = (a.astype(np.float64).T @ dz16.astype(np.float64)).astype(np.float32)
        else:
            gW[l] = a.T @ dz16
        if l > 0:
            if emulate_bf16:
                d = (dz16.astype(np.float64) @ bf16(net.W[l]).astype(np.float64).T).astype(np.float32)
            else:
                d = dz16 @ net.W[l].T.astype(np.float64)
            d = np.where(acts[l] > 0, d, 0)
            dz32 = d.astype(np.float32) if emulate_bf16 else d
            dz16 = bf16(dz32) if emulate_bf16 else dz32
    return gW, gb, loss


def adam_update(p, g, m, v, step, lr, b1, b2, eps):
    """One Adam step in float32 (bias corrections as the kernels use them)."""
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    c1 = np.float32(1.0 / (1.0 - b1 ** step))
    c2 = np.float32(1.0 / (1.0 - b2 ** step))
    b1, b2, lr, eps = np.float32(b1), np.float32(b2), np.float32(lr), np.float32(eps)
    m = b1 * m + (np.float32(1) - b1) * g
    v = b2 * v + (np.float32(1) - b2) * g * g
    p = p - lr * (m * c1) / (np.sqrt(v * c2) + eps)
    return p, m, v


# ---------------------------------------------------------------- samples
ROT_CCW = np.array([2, 5, 8, 1, 4, 7, 0, 3, 6])   # rotateStencil90(dir >= 1)
ROT_CW = np.array([6, 3, 0, 7, 4, 1, 8, 5, 2])    # rotateStencil90(dir < 1)
ROT_HALF = ROT_CCW[ROT_CCW]
REFLECT = np.array([0, 3, 6, 1, 4, 7, 2, 5, 8])   # reflectStencil_yEqx


def normalize(feat):
    """[n][10] samples -> [2n][10] network inputs (negate if ihk > 0, rotate
    the centre gradient into the first quadrant, then as is / reflected)."""
    f = np.array(feat, dtype=np.float32, copy=True)
    f[f[:, 9] > 0] *= np.float32(-1)
    gx = f[:, 7] - f[:, 1]
    gy = f[:, 5] - f[:, 3]
    gx = np.where(gx == 0, np.float32(EPS), gx)
    gy = np.where(gy == 0, np.float32(EPS), gy)
    theta = np.arctan2(gy.astype(np.float64), gx.astype(np.float64))
    theta = np.where(theta < 0, 2 * np.pi + theta, theta)
    st = f[:, :9].copy()
    q2 = (theta > np.pi / 2) & (theta <= np.pi)
    q3 = (theta > np.pi) & (theta < 1.5 * np.pi)
    q4 = theta >= 1.5 * np.pi
    st[q2] = f[q2][:, ROT_CW]
    st[q3] = f[q3][:, ROT_HALF]
    st[q4] = f[q4][:, ROT_CCW]
    out = np.empty((2 * f.shape[0], 10), np.float32)
    out[0::2, :9] = st
    out[1::2, :9] = st[:, REFLECT]
    out[0::2, 9] = f[:, 9]
    out[1::2, 9] = f[:, 9]
    return out


def hybrid(feat, pred, lo=0.004, hi=0.007, residual=True):
    """Hybrid h*kappa from network outputs pred [2n] (float32 arithmetic)."""
    ihk = np.asarray(feat, np.float32)[:, 9]
    a = np.abs(ihk)
    pred = np.asarray(pred, np.float32).reshape(-1, 2)
    p = np.float32(0.5) * (pred[:, 0] + pred[:, 1])
    if residual:
        p = p + (-a)
    lo, hi = np.float32(lo), np.float32(hi)
    lam = (hi - a) / (hi - lo)
    pb = (np.float32(1) - lam) * p + lam * (-a)
    p = np.where(a <= hi, pb, p)
    out = np.where(ihk > 0, np.abs(p), -np.abs(p))
    return np.where(a < lo, ihk, out).astype(np.float32)


def central(fp, f0, fm, h):
    """central_derivative on a uniform grid (same operation order)."""
    return ((fp - f0) * h / h + (f0 - fm) * h / h) / (h + h)


def curvature_samples(phi, h, x0=0.0, y0=0.0):
    """phi [(M+1)][(M+1)] indexed [i][j] (x index i).  Returns (feat [n][10]
    float32, proj [n][2], ij [n][2]) for every node with 4 <= i, j <= M-4
    and an opposite-sign (or zero) face neighbour, in (i, j) order."""
    phi = np.asarray(phi, np.float64)
    M = phi.shape[0] - 1
    nx = np.zeros_like(phi)
    ny = np.zeros_like(phi)
    c = phi[1:M, 1:M]
    fx = central(phi[2:, 1:M], c, phi[:-2, 1:M], h)
    fy = central(phi[1:M, 2:], c, phi[1:M, :-2], h)
    a = np.sqrt(fx * fx + fy * fy)
    ok = a >= EPS
    with np.errstate(invalid="ignore", divide="ignore"):
        nx[1:M, 1:M] = np.where(ok, fx / a, 0.0)
        ny[1:M, 1:M] = np.where(ok, fy / a, 0.0)
    kap = np.zeros_like(phi)
    kap[2:M - 1, 2:M - 1] = (central(nx[3:M, 2:M - 1], nx[2:M - 1, 2:M - 1], nx[1:M - 2, 2:M - 1], h) +
                             central(ny[2:M - 1, 3:M], ny[2:M - 1, 2:M - 1], ny[2:M - 1, 1:M - 2], h))
    feats, projs, ijs = [], [], []
    for i in range(4, M - 3):
        for j in range(4, M - 3):
            f0 = phi[i, j]
            if not (f0 * phi[i + 1, j] <= 0 or f0 * phi[i - 1, j] <= 0 or f0 * phi[i, j + 1] <= 0 or
                    f0 * phi[i, j - 1] <= 0):
                continue
            x, y = x0 + i * h, y0 + j * h
            xg, yg = x - f0 * nx[i, j], y - f0 * ny[i, j]
            u, v = (xg - x0) / h, (yg - y0) / h
            ci, cj = int(np.floor(u)), int(np.floor(v))
            ci = min(max(ci, 2), M - 3)
            cj = min(max(cj, 2), M - 3)
            u -= ci
            v -= cj
            kg = ((1 - u) * (1 - v) * kap[ci, cj] + u * (1 - v) * kap[ci + 1, cj] + (1 - u) * v * kap[ci, cj + 1] +
                  u * v * kap[ci + 1, cj + 1])
            st = [phi[i + di, j + dj] / h for di in (-1, 0, 1) for dj in (-1, 0, 1)]
            feats.append(st + [h * kg])
            projs.append((xg, yg))
            ijs.append((i, j))
    return (np.asarray(feats, np.float32).reshape(-1, 10), np.asarray(projs).reshape(-1, 2),
            np.asarray(ijs, np.int64).reshape(-1, 2))


# ------------------------------------------------- samples on any grid (2D/3D)
def _block_offsets(dim):
    """Offsets of the 3^dim block, x slowest (the feature order)."""
    r = (-1, 0, 1)
    if dim == 2:
        return [(i, j) for i in r for j in r]
    return [(i, j, k) for i in r for j in r for k in r]


def _arm(tables, f, n, h):
    """Node across face f of node n if the arm is exactly h long and ends on
    a node: along every transverse axis one of the two transverse distances
    is zero, and the face corner on that side is the node."""
    if tables.d[f, n] != h:
        return None
    c = 0
    for t in range(tables.dim - 1):
        if tables.d_tr[f, 2 * t, n] == 0.0:
            bit = 0
        elif tables.d_tr[f, 2 * t + 1, n] == 0.0:
            bit = 1
        else:
            return None
        c |= bit << t
    return int(tables.nbr[f, c, n])


def grid_block(tables, n, h):
    """The 3^dim nodes around n reached by arms of length h that end on nodes, walking
    x, then y, then z (dict offset -> node), or None."""
    dim = tables.dim
    blk = {(0,) * dim: n}
    for s, f in ((-1, 0), (1, 1)):
        m = _arm(tables, f, n, h)
        if m is None:
            return None
        blk[(s,) + (0,) * (dim - 1)] = m
    for i in (-1, 0, 1):
        b = blk[(i,) + (0,) * (dim - 1)]
        for s, f in ((-1, 2), (1, 3)):
            m = _arm(tables, f, b, h)
            if m is None:
                return None
            blk[(i, s) + (0,) * (dim - 2)] = m
    if dim == 3:
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                b = blk[(i, j, 0)]
                for s, f in ((-1, 4), (1, 5)):
                    m = _arm(tables, f, b, h)
                    if m is None:
                        return None
                    blk[(i, j, s)] = m
    return blk


def grid_samples(tables, phi, normals, kappa, h, band, nodes=None):
    """Rows for every node with |phi| < band that has a uniform block, in node
    order: (feat [n][3^dim+1] float32, offset [n][dim], node [n], skipped).
    nodes: only these candidate nodes (in the given order), for grids too
    large to visit every band node in Python."""
    dim, N = tables.dim, tables.n_nodes
    offs = _block_offsets(dim)
    cand = np.nonzero(np.abs(phi) < band)[0] if nodes is None else \
        [int(n) for n in nodes if abs(phi[int(n)]) < band]
    feats, offsets, nodes, skipped = [], [], [], 0
    for n in cand:
        blk = grid_block(tables, int(n), h)
        if blk is None:
            skipped += 1
            continue
        f0 = phi[n]
        off = [-f0 * normals[a * N + n] for a in range(dim)]
        lo, w = [], []
        for a in range(dim):
            u = min(max(off[a] / h, -1.0), 1.0)
            lo.append(-1 if u < 0.0 else 0)
            w.append(u - lo[-1])
        k = 0.0
        for c in range(1 << dim):
            wt = 1.0
            o = []
            for a in range(dim):
                bit = (c >> a) & 1
                wt = wt * (w[a] if bit else 1.0 - w[a])
                o.append(lo[a] + bit)
            k = k + wt * kappa[blk[tuple(o)]]
        feats.append([np.float32(phi[blk[o]] / h) for o in offs] + [np.float32(h * k)])
        offsets.append(off)
        nodes.append(int(n))
    q = 3 ** dim + 1
    return (np.asarray(feats, dtype=np.float32).reshape(-1, q), np.asarray(offsets, dtype=np.float64).reshape(-1, dim),
            np.asarray(nodes, dtype=np.int64), skipped)


def normalize_3d(feat):
    """[n][28] -> [n][28]: negate when ihk > 0, reflect axes with a negative
    centred difference, order axes by decreasing difference (ties: lower
    axis first); then the centre value f0, the half centred differences g at
    the -e_a slots and, elsewhere, the deviation from the plane f0 + o.g
    (FP32, in the kernel's operation order)."""
    feat = np.asarray(feat, dtype=np.float32)
    out = np.empty_like(feat)
    offs = _block_offsets(3)
    pos = {o: q for q, o in enumerate(offs)}
    for s in range(feat.shape[0]):
        f = feat[s] * (np.float32(-1.0) if feat[s, 27] > 0 else np.float32(1.0))
        g = [f[pos[(1, 0, 0)]] - f[pos[(-1, 0, 0)]], f[pos[(0, 1, 0)]] - f[pos[(0, -1, 0)]],
             f[pos[(0, 0, 1)]] - f[pos[(0, 0, -1)]]]
        g = [np.float32(EPS) if v == 0 else v for v in g]
        r = [-1 if v < 0 else 1 for v in g]
        g = [abs(v) for v in g]
        p = sorted(range(3), key=lambda a: (-g[a], a))
        v = np.empty(27, np.float32)
        for q, (i, j, k) in enumerate(offs):
            c = [0, 0, 0]
            c[p[0]], c[p[1]], c[p[2]] = i * r[p[0]], j * r[p[1]], k * r[p[2]]
            v[q] = f[pos[tuple(c)]]
        half = np.float32(0.5)
        f0 = v[pos[(0, 0, 0)]]
        gx = half * (v[pos[(1, 0, 0)]] - v[pos[(-1, 0, 0)]])
        gy = half * (v[pos[(0, 1, 0)]] - v[pos[(0, -1, 0)]])
        gz = half * (v[pos[(0, 0, 1)]] - v[pos[(0, 0, -1)]])
        for q, (i, j, k) in enumerate(offs):
            plane = np.float32(np.float32(np.float32(i) * gx + np.float32(j) * gy) + np.float32(k) * gz)
            out[s, q] = v[q] - np.float32(f0 + plane)
        out[s, pos[(0, 0, 0)]] = f0
        out[s, pos[(-1, 0, 0)]] = gx
        out[s, pos[(0, -1, 0)]] = gy
        out[s, pos[(0, 0, -1)]] = gz
        out[s, 27] = f[27]
    return out


def denormalize_3d(x):
    """The canonical stencil values [n][27] back from normalize_3d's basis."""
    x = np.asarray(x, dtype=np.float64)
    offs = _block_offsets(3)
    pos = {o: q for q, o in enumerate(offs)}
    f0, g = x[:, pos[(0, 0, 0)]], x[:, [pos[(-1, 0, 0)], pos[(0, -1, 0)], pos[(0, 0, -1)]]]
    v = np.empty((x.shape[0], 27))
    for q, o in enumerate(offs):
        v[:, q] = x[:, q] + f0 + g @ np.asarray(o, dtype=np.float64)
    v[:, pos[(0, 0, 0)]] = f0
    # the -e_a slot held g_a; its plane deviation equals the +e_a slot's
    for a, o in enumerate(((-1, 0, 0), (0, -1, 0), (0, 0, -1))):
        v[:, pos[o]] = f0 - g[:, a] + x[:, pos[tuple(-c for c in o)]]
    return v


def hybrid_3d(feat, pred, lo=0.004, hi=0.007, residual=True, deadband=0.0):
    feat = np.asarray(feat, dtype=np.float32)
    pred = np.asarray(pred, dtype=np.float32).reshape(-1)
    ihk = feat[:, 27]
    a = np.abs(ihk)
    if residual:
        p = np.where(np.abs(pred) > np.float32(deadband), pred, np.float32(0)) + -a
    else:
        p = pred
    lam = (np.float32(hi) - a) / (np.float32(hi) - np.float32(lo))
    p = np.where(a <= hi, (np.float32(1) - lam) * p + lam * -a, p)
    p = np.where(ihk > 0, np.abs(p), -np.abs(p))
    return np.where(a < lo, ihk, p).astype(np.float32)
