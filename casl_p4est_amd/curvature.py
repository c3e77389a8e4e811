"""Learned h*kappa: numerical curvature samples, the BF16-MFMA network, its
training on GPU-reinitialized level sets, and the hybrid estimate.

The reference's counterpart is kml::Curvature (src/my_p4est_curvature_ml.cpp,
src/my_p4est_curvature_ml.h): it samples nodes next to the interface, feeds
their (sign-normalised, rotated, reflected) phi stencils and the numerical
h*kappa to a network, and blends with the numerical value where |h*kappa| is
small (_computeHybridHK, :483-556).  Here every step runs on the GPU through
the library's own kernels:

    UniformPatch            a (2^L)^2 uniform grid on [0, size]^2 as one tree
    make_dataset            circles and sinusoids -> phi0 -> reinitialize_2nd_order
                            (FP64, 20 iterations) -> casl_curvature_samples,
                            with exact h*kappa targets at the closest interface point
    CurvatureNet            casl_mlp_* (BF16 MFMA GEMMs, FP32 accumulation, Adam)
    train                   mini-batch Adam on normalised inputs
    hybrid_hk               casl_curvature_normalize -> network -> casl_curvature_hybrid

In 3D (no counterpart in the reference, which is 2D only) the inputs come
from any grid through the stencil tables (casl_curvature_grid_samples): the
3x3x3 phi/h block around every node with |phi| < band plus the interpolated
numerical h*kappa, put in one canonical orientation of the cube
(casl_curvature_normalize_3d), for a 28-128-128-128-128-1 network:

    UniformBox              a (2^L)^3 uniform grid on [0, size]^3 as one tree
    grid_samples            the samples of any DeviceStencil
    make_dataset_3d         spheres and wave surfaces w = A sin(k u + a) sin(k v + b)
                            in random frames -> phi0 -> reinitialize_2nd_order
                            (FP64, 20 iterations) -> grid_samples, exact h*kappa
                            (h * div n) at the closest surface point
    gyroid_closest_hk       exact h*kappa of the gyroid level set

The network predicts a correction to the numerical value (residual=True):
h*kappa = -|ihk| + net(x) for a normalised input x; the hybrid estimate uses
the numerical h*kappa alone where |ihk| < 0.004 and blends up to 0.007, as the
reference does.
"""
import ctypes
import math
import time

import numpy as np

from . import _lib as L
from .forest import Forest, NodeNeighbors
from .level_set import LevelSet, _stream_ptr, _torch
from .levelset import Plane

LO_MIN_HK = 0.004
UP_MIN_HK = 0.007
DIMS_2D = (10, 128, 128, 128, 128, 1)
DIMS_3D = (28, 128, 128, 128, 128, 1)
# 3D corrections no larger than this are dropped (the network's noise floor on
# well-resolved surfaces, where the numerical value is more accurate)
DEADBAND_3D = 1e-3


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class UniformPatch:
    """Uniform grid of 2^level cells per side on [0, size]^2 (one tree, built
    by refining to min = max level), with its device stencil and the map from
    grid position (i, j) to node index."""

    def __init__(self, level=8, size=2.0):
        torch = _torch()
        self.level, self.size = level, float(size)
        self.M = 2 ** level
        self.h = self.size / self.M
        f = Forest(2, [0.0, 0.0], [self.size, self.size]).refine_levelset(Plane([1.0, 0.0], -0.5 * self.size),
                                                                         level, level)
        f.build_nodes()
        self.forest = f
        self.ngbd = NodeNeighbors(f).init_neighbors()
        self.n_nodes = self.ngbd.num_nodes
        if self.n_nodes != (self.M + 1) ** 2:
            raise RuntimeError("uniform patch has an unexpected node count")
        _, xyz = f.nodes()
        ij = np.rint(xyz / self.h).astype(np.int64)
        grid = np.full((self.M + 1) * (self.M + 1), -1, dtype=np.int32)
        grid[ij[:, 0] * (self.M + 1) + ij[:, 1]] = np.arange(self.n_nodes, dtype=np.int32)
        if (grid < 0).any():
            raise RuntimeError("uniform patch nodes do not cover the grid")
        self.xy = xyz
        self.grid = torch.from_numpy(grid).cuda()
        self.solver = LevelSet(self.ngbd)

    def reinitialize(self, phi, iterations=20, precision="fp64"):
        """reinitialize_2nd_order on every row of phi [n_patch][n_nodes]."""
        for p in range(phi.shape[0]):
            self.solver.reinitialize_2nd_order(phi[p], iterations, precision=precision)
        return phi

    def samples(self, phi, max_samples=None):
        """casl_curvature_samples on phi [n_patch][n_nodes] (float64, device).
        Returns (feat [n][10] float32, proj [n][2] float64, node [n] int64) on
        the device, sorted by node (patch-major)."""
        torch = _torch()
        n_patch = phi.shape[0]
        cap = max_samples or max(1024, n_patch * 16 * self.M)
        while True:
            feat = torch.empty((cap, 10), dtype=torch.float32, device=phi.device)
            proj = torch.empty((cap, 2), dtype=torch.float64, device=phi.device)
            node = torch.empty(cap, dtype=torch.int64, device=phi.device)
            cnt = ctypes.c_int64()
            L.call("casl_curvature_samples", _ptr(phi), n_patch, self.n_nodes, self.M, _ptr(self.grid), self.h, 0.0,
                   0.0, _ptr(feat), _ptr(proj), _ptr(node), cap, ctypes.byref(cnt), _stream_ptr(torch))
            if cnt.value <= cap:
                break
            cap = cnt.value
        n = cnt.value
        order = torch.argsort(node[:n])
        return feat[:n][order], proj[:n][order], node[:n][order]


# ------------------------------------------------------------------ shapes
def circle_phi0(xy, c, R):
    """(|x - c|^2 - R^2) / (2 R): negative inside, not a signed distance."""
    d2 = (xy[:, 0] - c[0]) ** 2 + (xy[:, 1] - c[1]) ** 2
    return (d2 - R * R) / (2.0 * R)


def sinusoid_frame(xy, c, theta):
    ct, st = math.cos(theta), math.sin(theta)
    dx, dy = xy[..., 0] - c[0], xy[..., 1] - c[1]
    return ct * dx + st * dy, -st * dx + ct * dy


def sinusoid_phi0(xy, c, theta, A, omega, phase):
    """v - A sin(omega u + phase) in the frame rotated by theta about c."""
    u, v = sinusoid_frame(xy, c, theta)
    return v - A * np.sin(omega * u + phase)


def sinusoid_closest_hk(points, c, theta, A, omega, phase, h):
    """Exact h*kappa at the point of the curve closest to each point (brute
    force over +-3h along the curve parameter, then Newton)."""
    u, v = sinusoid_frame(points, c, theta)
    s = u[:, None] + np.linspace(-3 * h, 3 * h, 97)[None, :]
    d2 = (s - u[:, None]) ** 2 + (A * np.sin(omega * s + phase) - v[:, None]) ** 2
    s = s[np.arange(len(u)), np.argmin(d2, axis=1)]
    for _ in range(8):
        ph = omega * s + phase
        f, f1, f2 = A * np.sin(ph), A * omega * np.cos(ph), -A * omega * omega * np.sin(ph)
        g1 = (s - u) + (f - v) * f1
        g2 = 1.0 + f1 * f1 + (f - v) * f2
        s = s - g1 / np.where(g2 > 0.1, g2, 0.1)
    ph = omega * s + phase
    f1, f2 = A * omega * np.cos(ph), -A * omega * omega * np.sin(ph)
    return h * (-f2 / (1.0 + f1 * f1) ** 1.5)


def make_dataset(patch, n_circles, n_sinusoids, seed=0, iterations=20, chunk=32, log=None):
    """Samples near random circles (radius U[1.6h, 0.5], centre jittered by
    U[-h/2, h/2]^2 about the patch centre) and sinusoids (A U[0.01, 0.2],
    omega U[2 pi, 12 pi], random phase, rotation and offset), each phi0
    reinitialized by the GPU solver.  Returns dict(feat (device float32
    [n][10]), target (numpy float64 exact h*kappa), kind (0 circle, 1 sinusoid),
    timings)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    h, mid = patch.h, 0.5 * patch.size
    shapes = []
    for _ in range(n_circles):
        shapes.append(("circle", (mid + rng.uniform(-h / 2, h / 2), mid + rng.uniform(-h / 2, h / 2)),
                       rng.uniform(1.6 * h, 0.5)))
    for _ in range(n_sinusoids):
        omega = rng.uniform(2 * math.pi, 12 * math.pi)
        shapes.append(("sinusoid", (mid + rng.uniform(-h / 2, h / 2), mid + rng.uniform(-h / 2, h / 2)),
                       rng.uniform(0, 2 * math.pi), rng.uniform(0.01, 0.2), omega, rng.uniform(0, 2 * math.pi)))
    feats, targets, kinds = [], [], []
    t_reinit = t_samples = 0.0
    for c0 in range(0, len(shapes), chunk):
        group = shapes[c0:c0 + chunk]
        phi_np = np.empty((len(group), patch.n_nodes))
        for k, s in enumerate(group):
            phi_np[k] = circle_phi0(patch.xy, s[1], s[2]) if s[0] == "circle" else sinusoid_phi0(patch.xy, *s[1:])
        phi = torch.from_numpy(phi_np).cuda()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        patch.reinitialize(phi, iterations)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        feat, proj, node = patch.samples(phi)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_reinit += t1 - t0
        t_samples += t2 - t1
        which = (node // patch.n_nodes).cpu().numpy()
        proj_np = proj.cpu().numpy()
        tgt = np.empty(len(which))
        for k, s in enumerate(group):
            sel = which == k
            if s[0] == "circle":
                tgt[sel] = h / s[2]
            else:
                tgt[sel] = sinusoid_closest_hk(proj_np[sel], *s[1:], h)
        feats.append(feat)
        targets.append(tgt)
        kinds.append(np.array([0 if group[k][0] == "circle" else 1 for k in which], dtype=np.int8))
        if log:
            log(f"[curvature] shapes {c0 + len(group)}/{len(shapes)}: {sum(len(t) for t in targets)} samples")
    return dict(feat=torch.cat(feats), target=np.concatenate(targets), kind=np.concatenate(kinds),
                reinit_s=t_reinit, samples_s=t_samples)


# ------------------------------------------------------------------ network
class CurvatureNet:
    """casl_mlp: fully connected ReLU network, BF16 MFMA GEMMs with FP32
    accumulation, FP32 master weights, Adam."""

    def __init__(self, dims=DIMS_2D, seed=0, zero_last_layer=True):
        _torch().cuda.init()
        self.dims = tuple(int(d) for d in dims)
        d = np.asarray(self.dims, dtype=np.int32)
        h = ctypes.c_void_p()
        L.call("casl_mlp_create", d.ctypes.data_as(L.c_int32_p), len(self.dims) - 1, int(seed), ctypes.byref(h))
        L.own(self, h, "casl_mlp_destroy")
        nw, nb = ctypes.c_int64(), ctypes.c_int64()
        L.call("casl_mlp_num_params", h, ctypes.byref(nw), ctypes.byref(nb))
        self.n_weights, self.n_biases = nw.value, nb.value
        self.mean = np.zeros(self.dims[0], np.float32)
        self.std = np.ones(self.dims[0], np.float32)
        # 3D hybrid: predicted corrections up to this size are dropped (chosen
        # after training on validation shapes, stored with the weights)
        self.deadband = DEADBAND_3D
        if zero_last_layer:
            # a correction network then starts at zero: the numerical value
            w, b = self.params()
            w[-self.dims[-2] * self.dims[-1]:] = 0.0
            self.set_params(w, b)

    def scale_output(self, factor):
        """Multiply the network output by factor (folded into the last layer)."""
        w, b = self.params()
        w[-self.dims[-2] * self.dims[-1]:] *= np.float32(factor)
        b[-self.dims[-1]:] *= np.float32(factor)
        self.set_params(w, b)

    def set_scaler(self, mean, std):
        self.mean = np.ascontiguousarray(mean, dtype=np.float32)
        self.std = np.ascontiguousarray(std, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        L.call("casl_mlp_set_scaler", self._h, self.mean.ctypes.data_as(fp), self.std.ctypes.data_as(fp))

    def params(self):
        w = np.empty(self.n_weights, np.float32)
        b = np.empty(self.n_biases, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        L.call("casl_mlp_get_params", self._h, w.ctypes.data_as(fp), b.ctypes.data_as(fp))
        return w, b

    def set_params(self, w, b):
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        if w.size != self.n_weights or b.size != self.n_biases:
            raise ValueError("parameter sizes do not match the network")
        fp = ctypes.POINTER(ctypes.c_float)
        L.call("casl_mlp_set_params", self._h, w.ctypes.data_as(fp), b.ctypes.data_as(fp))

    def save(self, path, note=""):
        """Weights, biases, input scaler and widths as an .npz (no pickles)."""
        w, b = self.params()
        np.savez(path, dims=np.asarray(self.dims, np.int32), w=w, b=b, mean=self.mean, std=self.std,
                 deadband=np.float32(self.deadband), note=np.asarray(note))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        net = cls(tuple(int(d) for d in z["dims"]), zero_last_layer=False)
        net.set_params(z["w"], z["b"])
        net.set_scaler(z["mean"], z["std"])
        if "deadband" in z.files:
            net.deadband = float(z["deadband"])
        return net

    def grads(self):
        w = np.empty(self.n_weights, np.float32)
        b = np.empty(self.n_biases, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        L.call("casl_mlp_get_grads", self._h, w.ctypes.data_as(fp), b.ctypes.data_as(fp))
        return w, b

    def _check(self, x):
        torch = _torch()
        if not isinstance(x, torch.Tensor) or x.device.type != "cuda" or x.dtype != torch.float32:
            raise TypeError("inputs must be a float32 CUDA tensor")
        if x.dim() != 2 or x.shape[1] != self.dims[0] or not x.is_contiguous():
            raise ValueError(f"inputs must be contiguous [n][{self.dims[0]}]")

    def forward(self, x, out=None):
        torch = _torch()
        self._check(x)
        if out is None:
            out = torch.empty((x.shape[0], self.dims[-1]), dtype=torch.float32, device=x.device)
        L.call("casl_mlp_forward", self._h, _ptr(x), int(x.shape[0]), _ptr(out), _stream_ptr(torch))
        return out

    def backward(self, x, t):
        """Forward + backward of mean((y - t)^2); returns the loss."""
        torch = _torch()
        self._check(x)
        if t.dtype != torch.float32 or t.numel() != x.shape[0] or not t.is_contiguous():
            raise ValueError("targets must be a contiguous float32 tensor with one value per row")
        loss = ctypes.c_float()
        L.call("casl_mlp_backward", self._h, _ptr(x), _ptr(t), int(x.shape[0]), ctypes.byref(loss),
               _stream_ptr(torch))
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        L.call("casl_mlp_adam_step", self._h, float(lr), float(beta1), float(beta2), float(eps),
               _stream_ptr(_torch()))


def normalize(feat):
    """casl_curvature_normalize: [n][10] samples -> [2n][10] network inputs."""
    torch = _torch()
    out = torch.empty((2 * feat.shape[0], 10), dtype=torch.float32, device=feat.device)
    L.call("casl_curvature_normalize", _ptr(feat), int(feat.shape[0]), _ptr(out), _stream_ptr(torch))
    return out


def hybrid(feat, pred, lo=LO_MIN_HK, hi=UP_MIN_HK, residual=True):
    torch = _torch()
    hk = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
    L.call("casl_curvature_hybrid", _ptr(feat), _ptr(pred), int(feat.shape[0]), float(lo), float(hi),
           int(bool(residual)), _ptr(hk), _stream_ptr(torch))
    return hk


def training_pairs(feat, target):
    """Normalised inputs ([2n][10] in 2D, [n][28] in 3D) and their targets:
    the exact h*kappa with the sign convention of the normalised stencil,
    minus the normalised numerical value -|ihk| (the network learns the
    correction)."""
    torch = _torch()
    three = feat.shape[1] == 28
    x = normalize_3d(feat) if three else normalize(feat)
    t = torch.as_tensor(target, dtype=torch.float64, device=feat.device)
    ihk = feat[:, -1].double()
    tn = (torch.where(ihk > 0, -t, t) + ihk.abs()).float()
    return x, (tn if three else tn.repeat_interleave(2)).contiguous()


def train(net, feat, target, epochs=10, batch=16384, lr=1e-3, lr_final=1e-5, seed=0, log=None):
    """Adam on normalised samples with a cosine learning-rate decay.  The
    input scaler comes from the data; the targets are trained at unit variance
    and that scale is folded into the last layer at the end.  Returns
    per-epoch mean losses (in the unit-variance scale) and timing."""
    torch = _torch()
    x, t = training_pairs(feat, target)
    mean = x.mean(dim=0).cpu().numpy()
    std = x.std(dim=0).clamp_min(1e-6).cpu().numpy()
    net.set_scaler(mean, std)
    sigma = float(t.std().clamp_min(1e-12))
    t = (t / sigma).contiguous()
    n = x.shape[0]
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    hist = []
    step = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.randperm(n, device=x.device, generator=g)
        acc = 0.0
        for k in range(steps_per_epoch):
            idx = perm[k * batch:(k + 1) * batch]
            xb = x.index_select(0, idx)
            tb = t.index_select(0, idx)
            acc += net.backward(xb, tb)
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / max(1, total - 1)))
            net.adam_step(cur)
            step += 1
        hist.append(acc / steps_per_epoch)
        if log:
            log(f"[curvature] epoch {ep + 1}/{epochs} loss {hist[-1]:.3e}")
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    net.scale_output(sigma)
    return dict(loss=hist, seconds=seconds, steps=step, rows=step * batch, target_scale=sigma)


def hybrid_hk(net, feat, residual=True, chunk=1 << 22, fused=True, deadband=None):
    """Learned h*kappa of samples feat ([n][10] or [n][28], device).

    fused (the default for the (10|28)-128x4-1 networks): one
    casl_curvature_learned launch that normalises, evaluates the network and
    blends, with the weights and activations in LDS.  Otherwise the three
    calls (normalisation, casl_mlp_forward, hybrid) in chunks of at most
    `chunk` samples (the network's activations are sized by the rows of one
    call).  Both give the same bits.  deadband (3D): the hybrid's deadband,
    the network's own (chosen after training) when None."""
    torch = _torch()
    dead = float(getattr(net, "deadband", DEADBAND_3D)) if deadband is None else float(deadband)
    if fused and net.dims in (DIMS_2D, DIMS_3D) and feat.shape[1] == net.dims[0]:
        if not feat.is_contiguous() or feat.dtype != torch.float32 or feat.device.type != "cuda":
            raise TypeError("feat must be a contiguous float32 CUDA tensor")
        hk = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
        L.call("casl_curvature_learned", net._h, _ptr(feat), int(feat.shape[0]), float(LO_MIN_HK),
               float(UP_MIN_HK), int(bool(residual)), dead, _ptr(hk), _stream_ptr(torch))
        return hk
    if feat.shape[0] > chunk:
        return torch.cat([hybrid_hk(net, feat[i:i + chunk], residual, chunk, False, deadband)
                          for i in range(0, feat.shape[0], chunk)])
    if feat.shape[1] == 28:
        return hybrid_3d(feat, net.forward(normalize_3d(feat)), residual=residual, deadband=dead)
    x = normalize(feat)
    y = net.forward(x)
    return hybrid(feat, y, residual=residual)


# ------------------------------------------------------------------ 3D
def normalize_3d(feat):
    """casl_curvature_normalize_3d: [n][28] samples -> [n][28] network inputs."""
    torch = _torch()
    out = torch.empty_like(feat)
    L.call("casl_curvature_normalize_3d", _ptr(feat), int(feat.shape[0]), _ptr(out), _stream_ptr(torch))
    return out


def hybrid_3d(feat, pred, lo=LO_MIN_HK, hi=UP_MIN_HK, residual=True, deadband=DEADBAND_3D):
    torch = _torch()
    hk = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
    L.call("casl_curvature_hybrid_3d", _ptr(feat), _ptr(pred), int(feat.shape[0]), float(lo), float(hi),
           int(bool(residual)), float(deadband), _ptr(hk), _stream_ptr(torch))
    return hk


def grid_samples(ds, phi, h, band, normals=None, kappa=None):
    """casl_curvature_grid_samples on a DeviceStencil: every node with
    |phi| < band that has a uniform 3^dim block at spacing h.  Returns
    (feat [n][3^dim+1] float32, offset [n][dim] float64, node [n] int64) on the
    device, sorted by node, and the count of band nodes without a block."""
    torch = _torch()
    if normals is None:
        normals = ds.normals(phi)
    if kappa is None:
        kappa = ds.mean_curvature(normals)
    nf = 3 ** ds.dim + 1
    cap = max(1024, int((phi.abs() < band).sum()))  # every sample is a band node
    while True:
        feat = torch.empty((cap, nf), dtype=torch.float32, device=phi.device)
        off = torch.empty((cap, ds.dim), dtype=torch.float64, device=phi.device)
        node = torch.empty(cap, dtype=torch.int64, device=phi.device)
        cnt, skipped = ctypes.c_int64(), ctypes.c_int64()
        L.call("casl_curvature_grid_samples", ds.handle, _ptr(phi), _ptr(normals), _ptr(kappa), float(h),
               float(band), _ptr(feat), _ptr(off), _ptr(node), cap, ctypes.byref(cnt), ctypes.byref(skipped),
               _stream_ptr(torch))
        if cnt.value <= cap:
            break
        cap = cnt.value
    n = cnt.value
    order = torch.argsort(node[:n])
    return feat[:n][order], off[:n][order], node[:n][order], skipped.value


class UniformBox:
    """Uniform grid of 2^level cells per side on [0, size]^3 (one tree refined
    to min = max level), with its device stencil and solver."""

    def __init__(self, level=7, size=1.0):
        self.level, self.size = level, float(size)
        self.M = 2 ** level
        self.h = self.size / self.M
        f = Forest(3, [0.0] * 3, [self.size] * 3).refine_levelset(Plane([1.0, 0.0, 0.0], -0.5 * self.size),
                                                                level, level)
        f.build_nodes()
        self.forest = f
        self.ngbd = NodeNeighbors(f).init_neighbors()
        self.n_nodes = self.ngbd.num_nodes
        if self.n_nodes != (self.M + 1) ** 3:
            raise RuntimeError("uniform box has an unexpected node count")
        _, self.xyz = f.nodes()
        self.ds = self.ngbd.device()
        self.solver = LevelSet(self.ngbd)


def random_frame(rng):
    """A random rotation (columns e_u, e_v, e_w) from the QR of a Gaussian matrix."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.diag(r))


def sphere3_phi0(xyz, c, R):
    """(|x - c|^2 - R^2) / (2 R): negative inside, not a signed distance."""
    d2 = ((xyz - np.asarray(c)) ** 2).sum(axis=1)
    return (d2 - R * R) / (2.0 * R)


def wave_phi0(xyz, c, Q, A, omega, a, b):
    """w - A sin(omega u + a) sin(omega v + b) in the frame Q about c."""
    uvw = (xyz - np.asarray(c)) @ Q
    return uvw[:, 2] - A * np.sin(omega * uvw[:, 0] + a) * np.sin(omega * uvw[:, 1] + b)


def wave_closest_hk(points, c, Q, A, omega, a, b, h):
    """Exact h * div(n) at the surface point closest to each point (Newton on
    the surface parameters (u, v), started at the point's own (u, v))."""
    p = (np.asarray(points) - np.asarray(c)) @ Q
    u, v = p[:, 0].copy(), p[:, 1].copy()

    def derivs(u, v):
        su, cu = np.sin(omega * u + a), np.cos(omega * u + a)
        sv, cv = np.sin(omega * v + b), np.cos(omega * v + b)
        f = A * su * sv
        fu, fv = A * omega * cu * sv, A * omega * su * cv
        fuu, fvv = -A * omega * omega * su * sv, -A * omega * omega * su * sv
        fuv = A * omega * omega * cu * cv
        return f, fu, fv, fuu, fvv, fuv

    for _ in range(12):
        f, fu, fv, fuu, fvv, fuv = derivs(u, v)
        r = f - p[:, 2]
        g1 = (u - p[:, 0]) + r * fu
        g2 = (v - p[:, 1]) + r * fv
        h11 = 1.0 + fu * fu + r * fuu
        h22 = 1.0 + fv * fv + r * fvv
        h12 = fu * fv + r * fuv
        det = h11 * h22 - h12 * h12
        ok = det > 1e-3
        du = np.where(ok, (h22 * g1 - h12 * g2) / np.where(ok, det, 1.0), 0.1 * g1)
        dv = np.where(ok, (h11 * g2 - h12 * g1) / np.where(ok, det, 1.0), 0.1 * g2)
        u, v = u - du, v - dv
    f, fu, fv, fuu, fvv, fuv = derivs(u, v)
    s2 = 1.0 + fu * fu + fv * fv
    return h * -((1.0 + fv * fv) * fuu - 2.0 * fu * fv * fuv + (1.0 + fu * fu) * fvv) / s2 ** 1.5


def implicit_closest_hk(phi_fn, grad_hess_fn, points, h, iterations=6):
    """Exact h * div(grad phi / |grad phi|) at the point of {phi = 0} reached
    from each point by Newton steps along the gradient (the foot point for
    points this close to the surface)."""
    x = np.array(points, dtype=np.float64)
    for _ in range(iterations):
        g, _ = grad_hess_fn(x)
        x = x - (phi_fn(x) / (g * g).sum(axis=1))[:, None] * g
    g, H = grad_hess_fn(x)
    gn2 = (g * g).sum(axis=1)
    tr = H[:, 0, 0] + H[:, 1, 1] + H[:, 2, 2]
    gHg = np.einsum("ni,nij,nj->n", g, H, g)
    return h * (gn2 * tr - gHg) / gn2 ** 1.5


def gyroid_fields(w=8 * math.pi, c=0.3, e=0.5):
    """phi(x) = G(x) (1 + e x y z), G = sin(wx)cos(wy) + sin(wy)cos(wz) + sin(wz)cos(wx) - c,
    with its gradient and Hessian (the level set of levelset.Gyroid)."""

    def parts(x):
        X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
        sx, cx, sy, cy, sz, cz = (np.sin(w * X), np.cos(w * X), np.sin(w * Y), np.cos(w * Y),
                                  np.sin(w * Z), np.cos(w * Z))
        G = sx * cy + sy * cz + sz * cx - c
        Gx = w * (cx * cy - sz * sx)
        Gy = w * (-sx * sy + cy * cz)
        Gz = w * (-sy * sz + cz * cx)
        Gxx = w * w * (-sx * cy - sz * cx)
        Gyy = w * w * (-sx * cy - sy * cz)
        Gzz = w * w * (-sy * cz - sz * cx)
        Gxy = -w * w * cx * sy
        Gyz = -w * w * cy * sz
        Gxz = -w * w * cz * sx
        m = 1.0 + e * X * Y * Z
        mx, my, mz = e * Y * Z, e * X * Z, e * X * Y
        return G, (Gx, Gy, Gz), (Gxx, Gyy, Gzz, Gxy, Gyz, Gxz), m, (mx, my, mz), (e * Z, e * X, e * Y)

    def phi(x):
        G, _, _, m, _, _ = parts(x)
        return G * m

    def grad_hess(x):
        G, (Gx, Gy, Gz), (Gxx, Gyy, Gzz, Gxy, Gyz, Gxz), m, (mx, my, mz), (mxy, myz, mxz) = parts(x)
        g = np.stack([Gx * m + G * mx, Gy * m + G * my, Gz * m + G * mz], axis=1)
        H = np.empty((x.shape[0], 3, 3))
        H[:, 0, 0] = Gxx * m + 2 * Gx * mx
        H[:, 1, 1] = Gyy * m + 2 * Gy * my
        H[:, 2, 2] = Gzz * m + 2 * Gz * mz
        H[:, 0, 1] = H[:, 1, 0] = Gxy * m + Gx * my + Gy * mx + G * mxy
        H[:, 1, 2] = H[:, 2, 1] = Gyz * m + Gy * mz + Gz * my + G * myz
        H[:, 0, 2] = H[:, 2, 0] = Gxz * m + Gx * mz + Gz * mx + G * mxz
        return g, H

    return phi, grad_hess


def make_dataset_3d(box, n_spheres, n_waves, seed=0, iterations=20, per_shape=4096, log=None):
    """Samples near random spheres (radius log-uniform in [2h, 8], the surface
    passing within 0.3 of the box centre) and wave surfaces (A U[0.02, 0.2],
    omega U[pi, 6 pi], random phases, frame and offset), each phi0
    reinitialized by the GPU solver (FP64), at most per_shape samples per
    shape.  Returns dict(feat (device float32 [n][28]), target (numpy exact
    h*kappa), kind (0 sphere, 1 wave), timings)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    h, mid = box.h, 0.5 * box.size
    shapes = []
    for _ in range(n_spheres):
        R = float(np.exp(rng.uniform(np.log(2 * h), np.log(8.0))))
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        d = max(0.0, R - 0.3 * rng.uniform(0.3, 1.0))
        c = mid - d * u + rng.uniform(-h / 2, h / 2, 3)
        shapes.append(("sphere", c, R))
    for _ in range(n_waves):
        c = mid + rng.uniform(-h / 2, h / 2, 3)
        shapes.append(("wave", c, random_frame(rng), rng.uniform(0.02, 0.2), rng.uniform(math.pi, 6 * math.pi),
                       rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi)))
    feats, targets, kinds = [], [], []
    t_reinit = t_samples = 0.0
    phi = torch.empty(box.n_nodes, dtype=torch.float64, device="cuda")
    for k, s in enumerate(shapes):
        p0 = sphere3_phi0(box.xyz, s[1], s[2]) if s[0] == "sphere" else wave_phi0(box.xyz, *s[1:])
        phi.copy_(torch.from_numpy(p0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        box.solver.reinitialize_2nd_order(phi, iterations, precision="fp64")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        feat, off, node, _ = grid_samples(box.ds, phi, h, 1.5 * h)
        torch.cuda.synchronize()
        t_reinit += t1 - t0
        t_samples += time.perf_counter() - t1
        if feat.shape[0] == 0:
            continue
        if feat.shape[0] > per_shape:
            pick = torch.from_numpy(np.sort(rng.choice(feat.shape[0], per_shape, replace=False))).cuda()
            feat, off, node = feat[pick], off[pick], node[pick]
        pts = box.xyz[node.cpu().numpy()] + off.cpu().numpy()
        if s[0] == "sphere":
            tgt = np.full(feat.shape[0], 2.0 * h / s[2])
        else:
            tgt = wave_closest_hk(pts, *s[1:], h)
        feats.append(feat)
        targets.append(tgt)
        kinds.append(np.full(feat.shape[0], 0 if s[0] == "sphere" else 1, dtype=np.int8))
        if log and (k + 1) % 50 == 0:
            log(f"[curvature3d] shapes {k + 1}/{len(shapes)}: {sum(len(t) for t in targets)} samples")
    return dict(feat=torch.cat(feats), target=np.concatenate(targets), kind=np.concatenate(kinds),
                reinit_s=t_reinit, samples_s=t_samples)
