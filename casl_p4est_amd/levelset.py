"""Level-set functions used to refine forests and sample phi_0 (CF_2/CF_3
continuous functions in the reference).  Each maps to a native evaluator so
refinement of multi-million-cell trees does not call back into Python.
"""
import ctypes

import numpy as np

from . import _lib as L


class LevelSet:
    kind = None
    dim = None

    def __init__(self, params, n_items=0):
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.n_items = int(n_items)
        self._keep = None

    def spec(self):
        s = L.LevelSetSpec()
        s.kind = self.kind
        s.n_items = self.n_items
        s.params = self.params.ctypes.data_as(L.c_double_p)
        s.fn = L.LEVELSET_FN()
        s.ctx = None
        return s

    def __call__(self, xyz):
        """Evaluates at points xyz [n][dim] (native evaluator)."""
        xyz = np.ascontiguousarray(np.atleast_2d(xyz), dtype=np.float64)
        n = xyz.shape[0]
        out = np.empty(n, dtype=np.float64)
        spec = self.spec()
        L.call("casl_levelset_eval", ctypes.byref(spec), int(xyz.shape[1]), n, xyz.ctypes.data_as(L.c_double_p),
               out.ctypes.data_as(L.c_double_p))
        return out


class Star(LevelSet):
    """(|x-c| - (r0 + A cos(p theta))) * (1 + k |x-c|^2), theta = atan2(y-cy, x-cx)."""
    kind = L.CASL_LS_STAR
    dim = 2

    def __init__(self, center=(0.5, 0.5), r0=0.3, amplitude=0.05, lobes=5, k=0.8):
        super().__init__([center[0], center[1], r0, amplitude, lobes, k])


class Spheres(LevelSet):
    """min_i(|x-c_i| - r_i) * (1 + m sin(2 pi x) cos(2 pi y))."""
    kind = L.CASL_LS_SPHERES
    dim = 3

    def __init__(self, centers, radii, modulation=0.5):
        centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
        radii = np.asarray(radii, dtype=np.float64).reshape(-1)
        p = np.concatenate([[modulation], np.column_stack([centers, radii]).reshape(-1)])
        super().__init__(p, n_items=len(radii))


class Gyroid(LevelSet):
    """(sin(wx)cos(wy) + sin(wy)cos(wz) + sin(wz)cos(wx) - c) * (1 + e x y z)."""
    kind = L.CASL_LS_GYROID
    dim = 3

    def __init__(self, w=8 * np.pi, c=0.3, e=0.5):
        super().__init__([w, c, e])


class Sphere(LevelSet):
    """Signed distance |x-c| - r (circle in 2D)."""
    kind = L.CASL_LS_SPHERE

    def __init__(self, center, radius):
        self.dim = len(center)
        super().__init__(list(center) + [radius])


class Plane(LevelSet):
    """a . x + b."""
    kind = L.CASL_LS_PLANE

    def __init__(self, normal, offset):
        self.dim = len(normal)
        super().__init__(list(normal) + [offset])


class Modulated(LevelSet):
    """(|x-c| - r) * (1 + q sin(5x) cos(4y)), the 2D reinitialization
    example's phi_0 (examples/reinitialization/main_2d.cpp:112-116)."""
    kind = L.CASL_LS_MODULATED
    dim = 2

    def __init__(self, center=(0.0, 0.0), radius=0.5, q=0.5):
        super().__init__([center[0], center[1], radius, q])


class Callback(LevelSet):
    """Arbitrary Python callable f(xyz) -> float (slow; for small trees)."""
    kind = L.CASL_LS_CALLBACK

    def __init__(self, fn, dim):
        self.dim = dim
        super().__init__([0.0])

        def c_fn(xyz_p, _ctx):
            return float(fn(np.ctypeslib.as_array(xyz_p, shape=(dim,)).copy()))

        self._cfn = L.LEVELSET_FN(c_fn)

    def spec(self):
        s = super().spec()
        s.fn = self._cfn
        return s


def random_spheres(n, seed, lo=0.2, hi=0.8, rmin=0.03, rmax=0.08, modulation=0.5):
    """n spheres with centres uniform in [lo, hi]^3 and radii uniform in
    [rmin, rmax], drawn from numpy's PCG64 stream of `seed`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centers = rng.uniform(lo, hi, size=(n, 3))
    radii = rng.uniform(rmin, rmax, size=n)
    return Spheres(centers, radii, modulation)


def random_star(rng):
    """A random star: centre U[0.35,0.65]^2, r0 U[0.15,0.3], amplitude U[0,0.08],
    lobes U{3..8}, scale factor (1 + 0.8 |x-c|^2)."""
    c = rng.uniform(0.35, 0.65, size=2)
    r0 = rng.uniform(0.15, 0.3)
    a = rng.uniform(0.0, 0.08)
    p = int(rng.integers(3, 9))
    return Star(center=(c[0], c[1]), r0=r0, amplitude=a, lobes=p, k=0.8)
