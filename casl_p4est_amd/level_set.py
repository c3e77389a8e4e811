"""my_p4est_level_set_t reinitialization entry points on the GPU.

Mirrors src/my_p4est_level_set.h:385-613 (names, argument meaning and error
behaviour): every reinitialize_* method forwards to
reinitialize_within_range_of_phi_0 (src/my_p4est_level_set.cpp:290-557),
which here is one call into the native library (casl_reinit).  phi is
updated in place.  It may be a torch float64 CUDA tensor (kept on the
device) or a float64 numpy array (staged through pinned memory to the
device and back, casl_reinit_host).

Errors: invalid orders raise ValueError (the reference throws
std::invalid_argument); any other native failure raises CaslError.
"""
import ctypes
import sys

import numpy as np

from . import _lib as L

DBL_MAX = sys.float_info.max


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("casl_p4est_amd needs a HIP device for reinitialization")
    return torch


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def tile_plan(tables):
    """The tiled encoding of host tables, built and checked on the host only
    (casl_tile_plan; no device needed).  Returns counts of tiles, records
    (T + H1 over all tiles), records with hanging faces, tile nodes with
    hanging faces and records with two hanging faces; raises CaslError with
    the reason when the tables cannot be tiled (their stencil then runs the
    two-pass path)."""
    vals = [ctypes.c_int64() for _ in range(5)]
    s = tables.struct()
    L.call("casl_tile_plan", ctypes.byref(s), *[ctypes.byref(v) for v in vals])
    keys = ("n_tiles", "n_records", "n_irregular", "n_irregular_nodes", "n_two_hanging")
    return {k: v.value for k, v in zip(keys, vals)}


class DeviceStencil:
    """Device copy of the flattened stencils (casl_stencil_create)."""

    def __init__(self, tables, handle=None):
        torch = _torch()
        torch.cuda.init()
        self.tables = tables
        if handle is not None:  # a handle created by the library (casl_device_grid_stencil)
            dim, n = ctypes.c_int32(), ctypes.c_int64()
            L.call("casl_stencil_info", handle, ctypes.byref(dim), ctypes.byref(n))
            self.dim, self.n_nodes = dim.value, n.value
            L.own(self, handle, "casl_stencil_destroy")
            return
        self.dim = tables.dim
        self.n_nodes = tables.n_nodes
        h = ctypes.c_void_p()
        s = tables.struct()
        L.call("casl_stencil_create", ctypes.byref(s), ctypes.byref(h))
        L.own(self, h, "casl_stencil_destroy")

    @property
    def handle(self):
        return self._h

    def set_path(self, path):
        """"tiled" (fused stage kernel over LDS-staged Morton tiles; the default
        when the stencil makes at least one tile per compute unit) or
        "two_pass" (grid-wide pass A + pass B; the default for smaller grids)."""
        code = {"tiled": L.CASL_PATH_TILED, "two_pass": L.CASL_PATH_TWO_PASS}[path]
        L.call("casl_stencil_set_path", self._h, code)

    @property
    def path(self):
        v = ctypes.c_int32()
        L.call("casl_stencil_get_path", self._h, ctypes.byref(v))
        return "tiled" if v.value == L.CASL_PATH_TILED else "two_pass"

    def tile_info(self):
        vals = [ctypes.c_int32()] + [ctypes.c_int64() for _ in range(4)]
        L.call("casl_stencil_tile_info", self._h, *[ctypes.byref(v) for v in vals])
        keys = ("tiled", "n_tiles", "n_halo", "n_records", "n_irregular")
        out = {k: v.value for k, v in zip(keys, vals)}
        out["tiled"] = bool(out["tiled"])
        if not out["tiled"]:
            msg = L.lib().casl_last_error()
            out["reason"] = msg.decode() if msg else ""
        return out

    def tile_check(self, tables):
        """casl_stencil_tile_check: (True, "") when this stencil's tiled
        encoding equals the host builder's encoding of `tables` (host tables
        of the same grid) byte for byte, else (False, the first difference)."""
        eq = ctypes.c_int32()
        s = tables.struct()
        L.call("casl_stencil_tile_check", self._h, ctypes.byref(s), ctypes.byref(eq))
        msg = L.lib().casl_last_error()
        return bool(eq.value), ("" if eq.value else (msg.decode() if msg else ""))

    def _check_field(self, t, name):
        torch = _torch()
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float64:
            raise TypeError(f"{name} must be a float64 CUDA tensor")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def second_derivatives(self, f, fdd=None):
        torch = _torch()
        self._check_field(f, "f")
        if f.numel() != self.n_nodes:
            raise ValueError("f must hold one value per node")
        if fdd is None:
            fdd = torch.empty((self.dim, self.n_nodes), dtype=torch.float64, device=f.device)
        self._check_field(fdd, "fdd")
        if fdd.numel() != self.dim * self.n_nodes:
            raise ValueError("fdd must hold dim values per node")
        L.call("casl_second_derivatives", self._h, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(fdd.data_ptr()),
               _stream_ptr(torch))
        return fdd

    def reinit(self, phi, order_space, order_time, n_iter, limit_high, limit_low, precision, mask=None,
               graph=False):
        """reinitialize_within_range_of_phi_0 on a float64 CUDA tensor, in place.
        graph=True: record the call into a hipGraph once and replay it
        (casl_reinit_graph); not with a mask."""
        torch = _torch()
        self._check_field(phi, "phi")
        if phi.numel() != self.n_nodes:
            raise ValueError("phi must hold one value per node")
        st = _stream_ptr(torch)
        if graph and mask is not None:
            raise ValueError("graph replay is not available with a mask")
        if mask is not None and (int(order_space), int(order_time)) != (2, 2):
            # reinitialize_2nd_order_with_mask is second order in space and
            # pseudo-time only (src/my_p4est_level_set.h:487-530)
            raise ValueError("a mask is supported with orders (2, 2) only")
        if mask is None:
            L.call("casl_reinit_graph" if graph else "casl_reinit", self._h, ctypes.c_void_p(phi.data_ptr()),
                   int(order_space), int(order_time), int(n_iter), float(limit_high), float(limit_low),
                   int(precision), st)
        else:
            if mask.dtype != torch.uint8 or mask.device.type != "cuda" or mask.numel() != self.n_nodes:
                raise TypeError("mask must be a uint8 CUDA tensor with one entry per node")
            L.call("casl_reinit_masked", self._h, ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(mask.data_ptr()),
                   int(n_iter), float(limit_high), float(limit_low), int(precision), st)
        return phi

    def reinit_host(self, phi, order_space, order_time, n_iter, limit_high, limit_low, precision, mask=None):
        """reinitialize_within_range_of_phi_0 on a float64 numpy array in host
        memory, in place (casl_reinit_host: pinned, chunked staging); mask: a
        numpy array, nodes with mask == 0 keep their value (orders 2/2)."""
        torch = _torch()
        if not isinstance(phi, np.ndarray) or phi.dtype != np.float64 or phi.shape != (self.n_nodes,) \
                or not phi.flags.c_contiguous or not phi.flags.writeable:
            raise TypeError("phi must be a writeable contiguous float64 array with one value per node")
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
            if m.shape != (self.n_nodes,):
                raise TypeError("mask must hold one entry per node")
        L.call("casl_reinit_host", self._h, ctypes.c_void_p(phi.ctypes.data),
               ctypes.c_void_p(m.ctypes.data) if m is not None else None, int(order_space), int(order_time),
               int(n_iter), float(limit_high), float(limit_low), int(precision), _stream_ptr(torch))
        return phi

    def _field_out(self, f, out, comps):
        torch = _torch()
        self._check_field(f, "f")
        if f.numel() != comps[0] * self.n_nodes:
            raise ValueError("input has the wrong number of values")
        if out is None:
            out = torch.empty((comps[1], self.n_nodes), dtype=torch.float64, device=f.device)
        self._check_field(out, "out")
        if out.numel() != comps[1] * self.n_nodes:
            raise ValueError("output has the wrong number of values")
        return out

    def first_derivatives(self, f, fd=None):
        """first_derivatives_central: fd[a][n] = d f / dx_a (FP64)."""
        fd = self._field_out(f, fd, (1, self.dim))
        L.call("casl_first_derivatives", self._h, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(fd.data_ptr()),
               _stream_ptr(_torch()))
        return fd

    def normals(self, phi, out=None):
        """compute_normals: grad(phi) / |grad(phi)|, 0 where |grad(phi)| < EPS."""
        out = self._field_out(phi, out, (1, self.dim))
        L.call("casl_normals", self._h, ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(out.data_ptr()),
               _stream_ptr(_torch()))
        return out

    def mean_curvature(self, normals, out=None):
        """compute_mean_curvature(normals): div(normals) by central derivatives."""
        out = self._field_out(normals, out, (self.dim, 1))
        L.call("casl_mean_curvature", self._h, ctypes.c_void_p(normals.data_ptr()), ctypes.c_void_p(out.data_ptr()),
               _stream_ptr(_torch()))
        return out.view(-1)

    def advect(self, phi, vn, dt, vn_np1=None):
        """One advect_in_normal_direction step (FP64, in place)."""
        self._check_field(phi, "phi")
        self._check_field(vn, "vn")
        if vn_np1 is not None:
            self._check_field(vn_np1, "vn_np1")
        for t in (phi, vn) + ((vn_np1,) if vn_np1 is not None else ()):
            if t.numel() != self.n_nodes:
                raise ValueError("fields must hold one value per node")
        L.call("casl_advect_in_normal_direction", self._h, ctypes.c_void_p(phi.data_ptr()),
               ctypes.c_void_p(vn.data_ptr()),
               ctypes.c_void_p(vn_np1.data_ptr()) if vn_np1 is not None else None, float(dt),
               _stream_ptr(_torch()))
        return phi

    def cfl_dt(self, vn):
        self._check_field(vn, "vn")
        if vn.numel() != self.n_nodes:
            raise ValueError("vn must hold one value per node")
        out = ctypes.c_double()
        L.call("casl_advect_cfl_dt", self._h, ctypes.c_void_p(vn.data_ptr()), ctypes.byref(out),
               _stream_ptr(_torch()))
        return out.value

    def band_linf(self, a, b, ref, band):
        torch = _torch()
        for t, n in ((a, "a"), (b, "b"), (ref, "ref")):
            self._check_field(t, n)
        out = ctypes.c_double()
        L.call("casl_band_linf", self._h, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
               ctypes.c_void_p(ref.data_ptr()), float(band), ctypes.byref(out), _stream_ptr(torch))
        return out.value


class LevelSet:
    """my_p4est_level_set_t(ngbd): reinitialization of node-sampled level sets.

    precision: "fp64" (default; every operation in FP64) or "mixed" (FP32 bulk
    updates and second derivatives, FP64 subcell roots and norms).
    graph: replay unmasked calls on device tensors from a recorded hipGraph
    (casl_reinit_graph); the results are the same, the launches fewer.
    """

    def __init__(self, ngbd, precision="fp64", graph=False):
        self.ngbd = ngbd
        if ngbd.tables is None:
            ngbd.init_neighbors()
        self.precision = precision
        self.graph = bool(graph)

    def _precision_code(self, precision):
        p = self.precision if precision is None else precision
        if p in ("fp64", L.CASL_PRECISION_FP64):
            return L.CASL_PRECISION_FP64
        if p in ("mixed", L.CASL_PRECISION_MIXED):
            return L.CASL_PRECISION_MIXED
        raise ValueError(f"unknown precision {p!r}")

    def reinitialize_within_range_of_phi_0(self, phi, order_space, order_pseudotime, phi_0_limit_high,
                                           phi_0_limit_low, number_of_iterations, mask=None, precision=None):
        if order_space not in (1, 2) or order_pseudotime not in (1, 2):
            raise ValueError("my_p4est_level_set_t::reinitialize_within_range_of_phi_0: valid orders of "
                             "accuracy for space and pseudotime are 1 and 2 only!")
        code = self._precision_code(precision)
        dev = self.ngbd.device()
        if isinstance(phi, np.ndarray):
            if mask is not None and not isinstance(mask, np.ndarray):
                mask = mask.cpu().numpy()
            try:
                return dev.reinit_host(phi, order_space, order_pseudotime, number_of_iterations, phi_0_limit_high,
                                       phi_0_limit_low, code, mask)
            except L.CaslError as e:
                if e.code == L.CASL_ERR_INVALID_ARGUMENT:
                    raise ValueError(str(e)) from e
                raise
        try:
            return dev.reinit(phi, order_space, order_pseudotime, number_of_iterations, phi_0_limit_high,
                              phi_0_limit_low, code, mask, graph=self.graph and mask is None)
        except L.CaslError as e:
            if e.code == L.CASL_ERR_INVALID_ARGUMENT:
                raise ValueError(str(e)) from e
            raise

    def advect_in_normal_direction(self, vn, phi, dt=None, vn_np1=None):
        """advect_in_normal_direction (src/my_p4est_level_set.h:625-659): one
        RK2 step of phi_t + vn |grad phi| = 0 with node-sampled vn (and vn_np1
        for the second stage).  Without dt the step is the velocity-function
        form's CFL choice, min 0.8 |s_min / vn|.  phi is updated in place
        (numpy or float64 CUDA tensor); returns dt."""
        dev = self.ngbd.device()
        torch = _torch()
        if isinstance(phi, np.ndarray):
            t = torch.from_numpy(np.ascontiguousarray(phi, dtype=np.float64)).cuda()
            v = torch.from_numpy(np.ascontiguousarray(vn, dtype=np.float64)).cuda()
            v1 = torch.from_numpy(np.ascontiguousarray(vn_np1, dtype=np.float64)).cuda() if vn_np1 is not None else None
            dt = dev.cfl_dt(v) if dt is None else dt
            dev.advect(t, v, dt, v1)
            phi[:] = t.cpu().numpy()
            return dt
        dt = dev.cfl_dt(vn) if dt is None else dt
        dev.advect(phi, vn, dt, vn_np1)
        return dt

    # --- the public entry points of my_p4est_level_set.h:434-613
    def reinitialize_1st_order(self, phi, number_of_iterations=50, limit_absolute_value_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 1, 1, limit_absolute_value_phi_0,
                                                       -limit_absolute_value_phi_0, number_of_iterations)

    def reinitialize_1st_order_above_threshold(self, phi, threshold, number_of_iterations=50,
                                               upper_limit_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 1, 1, upper_limit_phi_0, threshold, number_of_iterations)

    def reinitialize_2nd_order(self, phi, number_of_iterations=50, limit_absolute_value_phi_0=DBL_MAX,
                               precision=None):
        return self.reinitialize_within_range_of_phi_0(phi, 2, 2, limit_absolute_value_phi_0,
                                                       -limit_absolute_value_phi_0, number_of_iterations,
                                                       precision=precision)

    def reinitialize_2nd_order_with_mask(self, phi, mask, number_of_masked=0, number_of_iterations=50,
                                         limit_absolute_value_phi_0=DBL_MAX, precision=None):
        """mask: nodes with mask == 0 keep their current value (number_of_masked is
        the reference's reservation hint and is not needed here)."""
        torch = _torch()
        if isinstance(mask, np.ndarray) or not isinstance(mask, torch.Tensor) or mask.dtype != torch.uint8:
            mask = torch.as_tensor(np.asarray(mask if not isinstance(mask, torch.Tensor) else mask.cpu()) != 0,
                                   dtype=torch.uint8)
        if isinstance(phi, torch.Tensor):
            mask = mask.to(phi.device)
        return self.reinitialize_within_range_of_phi_0(phi, 2, 2, limit_absolute_value_phi_0,
                                                       -limit_absolute_value_phi_0, number_of_iterations,
                                                       mask=mask, precision=precision)

    def reinitialize_2nd_order_above_threshold(self, phi, threshold, number_of_iterations=50,
                                               upper_limit_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 2, 2, upper_limit_phi_0, threshold, number_of_iterations)

    def reinitialize_2nd_order_time_1st_order_space(self, phi, number_of_iterations=50,
                                                    limit_absolute_value_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 1, 2, limit_absolute_value_phi_0,
                                                       -limit_absolute_value_phi_0, number_of_iterations)

    def reinitialize_2nd_order_time_1st_order_space_above_threshold(self, phi, threshold, number_of_iterations=50,
                                                                    upper_limit_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 1, 2, upper_limit_phi_0, threshold, number_of_iterations)

    def reinitialize_1st_order_time_2nd_order_space(self, phi, number_of_iterations=50,
                                                    limit_absolute_value_phi_0=DBL_MAX, precision=None):
        return self.reinitialize_within_range_of_phi_0(phi, 2, 1, limit_absolute_value_phi_0,
                                                       -limit_absolute_value_phi_0, number_of_iterations,
                                                       precision=precision)

    def reinitialize_1st_order_time_2nd_order_space_above_threshold(self, phi, threshold, number_of_iterations=50,
                                                                    upper_limit_phi_0=DBL_MAX):
        return self.reinitialize_within_range_of_phi_0(phi, 2, 1, upper_limit_phi_0, threshold, number_of_iterations)


def compute_normals(ngbd, phi, normals=None):
    """compute_normals(neighbors, phi, normals) (src/my_p4est_utils.cpp:1720-1740)."""
    return ngbd.device().normals(phi, normals)


def compute_mean_curvature(ngbd, normals, kappa=None):
    """compute_mean_curvature(neighbors, normals, kappa) (src/my_p4est_utils.cpp:1648-1690)."""
    return ngbd.device().mean_curvature(normals, kappa)


# the reference's class names
my_p4est_level_set_t = LevelSet
