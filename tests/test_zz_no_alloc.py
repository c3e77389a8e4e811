"""casl_reinit_noalloc (the entry point that allocates nothing: its buffers
come from a fixed device array) against casl_reinit through the C ABI: the
same bits for FP64 and mixed, on both stage paths, on two calls in a row the
second of which is smaller; a grid whose buffers do not fit is refused before
anything runs."""
import ctypes

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import _lib as L
from casl_p4est_amd import levelset as ls

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _grid(dim, lmax):
    shape = ls.Star() if dim == 2 else ls.random_spheres(5, seed=9, rmin=0.08, rmax=0.16)
    f = C.Forest(dim).refine_levelset(shape, 3, lmax, lip=2.4).build_nodes()
    return C.NodeNeighbors(f).init_neighbors(), f.sample(shape)


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _run(name, ds, phi, orders, it, prec, torch):
    L.call(name, ds.handle, ctypes.c_void_p(phi.data_ptr()), orders[0], orders[1], it, C.DBL_MAX, -C.DBL_MAX, prec,
           _stream(torch))
    torch.cuda.synchronize()
    return phi.cpu().numpy()


@pytest.mark.parametrize("path", ["tiled", "two_pass"])
def test_noalloc_equals_reinit(cuda, path):
    import torch
    # a large grid first, then a smaller one: the second call reuses the array
    # holding the first call's contents
    for dim, lmax, orders, it in ((3, 7, (2, 2), 6), (2, 8, (2, 1), 5), (3, 5, (1, 2), 4)):
        ngbd, phi0 = _grid(dim, lmax)
        ds = ngbd.device()
        ds.set_path(path)
        for prec in (0, 1):
            say(f"[noalloc] dim {dim} lmax {lmax} ({ngbd.num_nodes} nodes) {path} orders {orders} precision {prec}")
            ref = _run("casl_reinit", ds, torch.from_numpy(phi0).to(cuda), orders, it, prec, torch)
            out = _run("casl_reinit_noalloc", ds, torch.from_numpy(phi0).to(cuda), orders, it, prec, torch)
            assert np.array_equal(out, ref)


def test_noalloc_refuses_oversize(cuda):
    import torch
    # 3D two-pass FP64 needs 114 B per node: a grid of more than 589K nodes does not fit in 64 MiB
    ngbd, phi0 = _grid(3, 8)
    say(f"[noalloc] oversize grid: {ngbd.num_nodes} nodes")
    assert ngbd.num_nodes * 114 > (64 << 20)
    ds = ngbd.device()
    ds.set_path("two_pass")
    phi = torch.from_numpy(phi0).to(cuda)
    rc = L.lib().casl_reinit_noalloc(ds.handle, ctypes.c_void_p(phi.data_ptr()), 2, 2, 3, C.DBL_MAX, -C.DBL_MAX, 0,
                                     _stream(torch))
    assert rc == L.CASL_ERR_INVALID_ARGUMENT
    assert b"scratch" in L.lib().casl_last_error()
    torch.cuda.synchronize()
    assert np.array_equal(phi.cpu().numpy(), phi0)  # nothing ran
