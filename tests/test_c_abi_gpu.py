"""The library called from C++ through include/casl_reinit.h alone: the
program tests/c_abi/reinit_caller.cpp (compiled with g++ by build(), no HIP
headers) builds a 2D star forest and its stencils through the C ABI,
reinitializes phi held in host memory with casl_reinit_host (plain, masked,
mixed precision) and writes the results; they must equal the oracle's bit for
bit (FP64) or stay within 1e-3 h of it (mixed)."""
import os
import subprocess

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_caller_matches_oracle(cuda, tmp_path):
    exe = os.path.join(HERE, "c_abi", "reinit_caller")
    if not os.path.exists(exe):
        pytest.fail("tests/c_abi/reinit_caller is not built: run __graft_entry__.build()")
    out = tmp_path / "caller.bin"
    it = 12
    print("[c-abi] running the C++ caller", flush=True)
    r = subprocess.run([exe, str(out), str(it)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    # the same grid through the Python bindings of the same library: the oracle's tables
    f = C.Forest(2).refine_levelset(ls.Star(), 3, 8, lip=2.4).build_nodes()
    tab = C.flat_stencil_from_forest(f)
    n = tab.n_nodes
    raw = np.fromfile(out, dtype=np.uint8)
    assert raw.size == n * 33
    vals = raw[:n * 32].view(np.float64).reshape(n, 4)
    mask = raw[n * 32:]
    phi0 = vals[:, 0].copy()
    assert np.array_equal(phi0, f.sample(ls.Star()))
    assert 0 < int(mask.sum()) < n
    assert np.array_equal(vals[:, 1], O.reinit(tab, phi0, 2, 2, it))
    assert np.array_equal(vals[:, 2], O.reinit(tab, phi0, 2, 2, it, mask=mask))
    assert np.array_equal(vals[mask == 0, 2], phi0[mask == 0])
    h = float(np.min(f.dxyz_min()))
    ref = vals[:, 1]
    band = np.abs(ref) < 3 * h
    assert np.max(np.abs(vals[band, 3] - ref[band])) <= 1e-3 * h


@pytest.mark.parametrize("precision", [0, 1])
def test_host_staging_equals_device(cuda, precision):
    """casl_reinit_host on a grid of several staging chunks (one per 2^20
    nodes) equals casl_reinit on a device copy, with and without a mask."""
    import torch
    shape = ls.random_spheres(5, seed=9, rmin=0.08, rmax=0.16)
    f = C.Forest(3).refine_levelset(shape, 3, 9, lip=2.4).build_nodes()
    ngbd = C.NodeNeighbors(f).init_neighbors()
    ds = ngbd.device()
    phi0 = f.sample(shape)
    print(f"[c-abi] host staging, {ngbd.num_nodes} nodes, precision {precision}", flush=True)
    assert ngbd.num_nodes >= 3 << 20
    _, xyz = f.nodes()
    mask = (xyz[:, 2] > 0.4).astype(np.uint8)
    for m in (None, mask):
        host = phi0.copy()
        ds.reinit_host(host, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, precision, mask=m)
        dev = torch.from_numpy(phi0).to(cuda)
        ds.reinit(dev, 2, 2, 4, C.DBL_MAX, -C.DBL_MAX, precision,
                  mask=None if m is None else torch.from_numpy(m).to(cuda))
        assert np.array_equal(host, dev.cpu().numpy())
