"""Batched reinitialization of many independent small trees: one persistent
kernel for the whole batch (casl_batch_* in include/casl_reinit.h), applied
as reinitialize_2nd_order (src/my_p4est_level_set.h:468) to every tree.
"""
import ctypes

import numpy as np

from . import _lib as L
from .level_set import _stream_ptr, _torch

PRECISION = {"fp64": L.CASL_PRECISION_FP64, "mixed": L.CASL_PRECISION_MIXED}


class TreeBatch:
    """Device tables of a list of FlatStencil trees (same dimension)."""

    def __init__(self, tables):
        torch = _torch()
        torch.cuda.init()
        self.trees = list(tables)
        arr = (L.FlatStencil * len(self.trees))()
        for i, t in enumerate(self.trees):
            arr[i] = t.struct()
        h = ctypes.c_void_p()
        L.call("casl_batch_create", arr, len(self.trees), ctypes.byref(h))
        L.own(self, h, "casl_batch_destroy")
        self.sizes = np.array([t.n_nodes for t in self.trees], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.total_nodes = int(self.offsets[-1])

    def _check(self, phi):
        torch = _torch()
        if not isinstance(phi, torch.Tensor) or phi.dtype != torch.float64 or phi.device.type != "cuda":
            raise TypeError("phi must be a float64 CUDA tensor")
        if phi.numel() != self.total_nodes or not phi.is_contiguous():
            raise ValueError("phi must hold the concatenated node values of every tree")

    def set_schedule(self, block_threads, nodes_per_thread):
        """nodes_per_thread 1 or 2: the grid-table kernel; 0: the tiled kernel."""
        L.call("casl_batch_set_schedule", self._h, int(block_threads), int(nodes_per_thread))

    def tile_info(self):
        t, n = ctypes.c_int32(), ctypes.c_int64()
        L.call("casl_batch_tile_info", self._h, ctypes.byref(t), ctypes.byref(n))
        return {"tiled": bool(t.value), "n_tiles": n.value}

    def schedule(self):
        b, n = ctypes.c_int32(), ctypes.c_int32()
        L.call("casl_batch_get_schedule", self._h, ctypes.byref(b), ctypes.byref(n))
        return b.value, n.value

    def autotune(self, phi, number_of_iterations=20, precision="mixed"):
        """Times every launch shape on a copy of phi; keeps and returns the fastest."""
        torch = _torch()
        self._check(phi)
        b, n, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        L.call("casl_batch_autotune", self._h, ctypes.c_void_p(phi.data_ptr()), int(number_of_iterations),
               PRECISION[precision], _stream_ptr(torch), ctypes.byref(b), ctypes.byref(n), ctypes.byref(ms))
        return b.value, n.value, ms.value

    def reinitialize_2nd_order(self, phi, number_of_iterations=50, precision="mixed", graph=False):
        torch = _torch()
        self._check(phi)
        fn = "casl_batch_reinit_graph" if graph else "casl_batch_reinit"
        L.call(fn, self._h, ctypes.c_void_p(phi.data_ptr()), int(number_of_iterations), PRECISION[precision],
               _stream_ptr(torch))
        return phi
