"""The measured workloads at full size, against the oracle and through
properties that do not depend on size:

  * FP64 through the fused tiled kernel equals FP64 through the grid-wide
    two-pass kernels bit for bit (two independent encodings of the stencils);
  * the first FP64 iteration equals the CPU oracle's bit for bit; for the 3D
    spheres (34.2M nodes) all 50 FP64 iterations do (about 100 s of the
    multi-threaded oracle);
  * the mixed path stays within 1e-3 h_min of FP64 in the |phi| < 3 h_min band
    after the workload's full iteration count, and never flips a sign of phi0
    outside the anchored nodes;
  * batched trees: FP64 of one persistent launch equals the per-tree FP64 loop
    of the single-tree path, and the oracle's 20 iterations of all 4096 trees
    (run as one concatenated stencil: the trees share no node), bit for bit.

The gyroid runs at lmax 9 (17.5M nodes) and at its full lmax 11 (278M nodes,
about 130 GB of host memory for the host-built forest and tables).
"""
import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import workloads
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


def concat_tables(tabs):
    """One stencil of independent trees: their tables side by side, node
    indices offset by the nodes before each tree."""
    zdt = {t.zero_distance_threshold for t in tabs}
    assert len(zdt) == 1 and len({t.dim for t in tabs}) == 1
    offs = np.cumsum([0] + [t.n_nodes for t in tabs])
    nbr = np.concatenate([t.nbr + np.int32(o) for t, o in zip(tabs, offs)], axis=-1)
    d = np.concatenate([t.d for t in tabs], axis=-1)
    d_tr = np.concatenate([t.d_tr for t in tabs], axis=-1)
    wall = np.concatenate([t.wall for t in tabs])
    return C.FlatStencil(tabs[0].dim, nbr, d, d_tr, wall, zdt.pop())


def check_tree(W, torch, dev, iters_fp):
    ngbd, phi0, h = W["ngbd"], W["phi0"], W["h"]
    ds = ngbd.device()
    assert ds.tile_info()["tiled"]
    p0 = torch.from_numpy(phi0).to(dev)
    out = {}
    for path in ("tiled", "two_pass"):
        ds.set_path(path)
        p = p0.clone()
        ds.reinit(p, 2, 2, iters_fp, C.DBL_MAX, -C.DBL_MAX, 0)
        out[path] = p.cpu().numpy()
    ds.set_path("tiled")
    say(f"    tiled == two_pass after {iters_fp} FP64 iterations")
    assert np.array_equal(out["tiled"], out["two_pass"])
    one = p0.clone()
    ds.reinit(one, 2, 2, 1, C.DBL_MAX, -C.DBL_MAX, 0)
    ref = O.reinit(ngbd.tables, phi0, 2, 2, 1)
    e = rel_err(one.cpu().numpy(), ref)
    say(f"    one iteration vs oracle: rel {e:.2e}")
    assert np.array_equal(one.cpu().numpy(), ref)
    del ref
    fp64, mixed = p0.clone(), p0.clone()
    ds.reinit(fp64, 2, 2, W["iters"], C.DBL_MAX, -C.DBL_MAX, 0)
    ds.reinit(mixed, 2, 2, W["iters"], C.DBL_MAX, -C.DBL_MAX, 1)
    err = ds.band_linf(mixed, fp64, fp64, 3 * h)
    say(f"    mixed vs FP64 after {W['iters']} iterations: {err / h:.2e} h")
    assert err <= 1e-3 * h
    m = mixed.cpu().numpy()
    keep = np.abs(phi0) >= 1e-13  # anchored nodes (|phi0| < zdt) become 0
    assert np.all(np.sign(m[keep]) * np.sign(phi0[keep]) >= 0)
    return fp64


def check_gyroid_curvature(W, phi, cuda, n_check):
    """The 3x3x3 network-curvature stage on the reinitialized gyroid (the
    second half of the gyroid workload): samples of |phi| < 1.5 h against the
    oracle on n_check random band nodes (the same nodes have a uniform block,
    phi/h features bit for bit, interpolated h*kappa to FP32 rounding),
    normalisation bit for bit and the hybrid estimate to float rounding on
    those rows; over up to 2e5 samples the learned h*kappa is closer to the
    gyroid's exact h*kappa than the numerical one."""
    import os
    import torch
    from casl_p4est_amd import curvature as CV
    from oracle import curvature_oracle as CO
    ds, h = W["ngbd"].device(), W["h"]
    net = CV.CurvatureNet.load(os.path.join(os.path.dirname(CV.__file__), "weights", "curvature3d.npz"))
    nrm = ds.normals(phi)
    kap = ds.mean_curvature(nrm)
    feat, off, node, skipped = CV.grid_samples(ds, phi, h, 1.5 * h, nrm, kap)
    n = int(feat.shape[0])
    say(f"    curvature: {n} samples, {skipped} band nodes without a uniform block")
    assert n > 1000
    rng = np.random.default_rng(11)
    phi_h, kap_h, node_h = phi.cpu().numpy(), kap.cpu().numpy(), node.cpu().numpy()
    nrm_h = nrm.cpu().numpy().ravel()
    band = np.nonzero(np.abs(phi_h) < 1.5 * h)[0]
    assert len(band) == n + skipped
    cand = np.sort(rng.choice(band, min(len(band), n_check), replace=False))
    of, oo, on, osk = CO.grid_samples(W["ngbd"].tables, phi_h, nrm_h, kap_h, h, 1.5 * h, nodes=cand)
    del nrm_h, kap_h
    pos = np.minimum(np.searchsorted(node_h, cand), n - 1)
    found = node_h[pos] == cand
    say(f"    oracle on {len(cand)} band nodes: {len(on)} with a block")
    assert np.array_equal(cand[found], on) and osk == int((~found).sum())
    rows = pos[found]
    fd = feat[torch.from_numpy(rows).to(cuda)].cpu().numpy()
    assert np.array_equal(fd[:, :27], of[:, :27])
    np.testing.assert_allclose(fd[:, 27], of[:, 27], rtol=2e-7, atol=1e-12)
    assert np.array_equal(off.cpu().numpy()[rows], oo)
    fr = torch.from_numpy(fd).to(cuda)
    x = CV.normalize_3d(fr)
    assert np.array_equal(x.cpu().numpy(), CO.normalize_3d(fd))
    pred = net.forward(x)
    hk = CV.hybrid_3d(fr, pred).cpu().numpy()
    ohk = CO.hybrid_3d(fd, pred.cpu().numpy().ravel(), deadband=CV.DEADBAND_3D)
    np.testing.assert_allclose(hk, ohk, rtol=1e-6, atol=1e-9)
    # learned against numerical, both against the exact h*kappa at the projection
    pick = np.sort(rng.choice(n, min(n, 200000), replace=False))
    ml = CV.hybrid_hk(net, feat).cpu().numpy()[pick].astype(np.float64)
    ihk = feat[:, 27].cpu().numpy()[pick].astype(np.float64)
    _, xyz = W["forest"].nodes()
    pts = xyz[node_h[pick]] + off.cpu().numpy()[pick]
    del xyz
    exact = CV.implicit_closest_hk(*CV.gyroid_fields(), pts, h)
    e_num, e_ml = float(np.abs(ihk - exact).mean()), float(np.abs(ml - exact).mean())
    say(f"    mean |h kappa error| on {len(pick)} samples: learned {e_ml:.3e}, numerical {e_num:.3e}")
    assert e_ml < e_num


def test_spheres3d_full_size(cuda):
    import torch
    say("[config] spheres3d, full size")
    W = workloads.build_tree("spheres3d", log=say)
    check_tree(W, torch, cuda, 3)
    say(f"    all {W['iters']} FP64 iterations vs oracle ({O.num_threads()} threads)")
    p = torch.from_numpy(W["phi0"]).to(cuda)
    W["ngbd"].device().reinit(p, 2, 2, W["iters"], C.DBL_MAX, -C.DBL_MAX, 0)
    ref = O.reinit(W["ngbd"].tables, W["phi0"], 2, 2, W["iters"])
    out = p.cpu().numpy()
    say(f"    max |gpu - oracle| = {np.max(np.abs(out - ref)):.3e}")
    assert np.array_equal(out, ref)


def test_gyroid3d_lmax9(cuda):
    import torch
    say("[config] gyroid3d, lmax 9")
    W = workloads.build_tree("gyroid3d", lmax=9, log=say)
    phi = check_tree(W, torch, cuda, 3)
    check_gyroid_curvature(W, phi, cuda, 3000)


def test_gyroid3d_full_size(cuda):
    import torch
    say("[config] gyroid3d, full size")
    W = workloads.build_tree("gyroid3d", log=say)
    phi = check_tree(W, torch, cuda, 2)
    check_gyroid_curvature(W, phi, cuda, 3000)


def test_batch2d_full_size(cuda):
    import torch
    say("[config] batch2d, 4096 trees")
    B = workloads.build_batch(log=say)
    batch = C.TreeBatch(B["tabs"])
    p0 = torch.from_numpy(np.concatenate(B["phis"])).to(cuda)
    fp64 = p0.clone()
    batch.reinitialize_2nd_order(fp64, B["iters"], precision="fp64")
    out64 = fp64.cpu().numpy()
    say(f"    batched FP64 vs oracle, all {len(B['tabs'])} trees (lmax {sorted(set(B['lmax'].tolist()))})")
    ref = O.reinit(concat_tables(B["tabs"]), np.concatenate(B["phis"]), 2, 2, B["iters"])
    assert np.array_equal(out64, ref), float(np.max(np.abs(out64 - ref)))
    del ref
    say("    batched FP64 vs per-tree loop")
    for k, (tab, ph) in enumerate(zip(B["tabs"], B["phis"])):
        ds = C.DeviceStencil(tab)
        p = torch.from_numpy(ph).to(cuda)
        ds.reinit(p, 2, 2, B["iters"], C.DBL_MAX, -C.DBL_MAX, 0)
        assert np.array_equal(p.cpu().numpy(), out64[batch.offsets[k]:batch.offsets[k + 1]]), k
    batch.autotune(p0, B["iters"], "mixed")
    eager = p0.clone()
    batch.reinitialize_2nd_order(eager, B["iters"], precision="mixed")
    graph = p0.clone()
    batch.reinitialize_2nd_order(graph, B["iters"], precision="mixed", graph=True)
    assert torch.equal(graph, eager)
    mixed = eager.cpu().numpy()
    worst = 0.0
    for k in range(len(B["tabs"])):
        a, b = batch.offsets[k], batch.offsets[k + 1]
        band = np.abs(out64[a:b]) < 3 * B["h"][k]
        if band.any():
            worst = max(worst, float(np.max(np.abs(mixed[a:b][band] - out64[a:b][band]))) / B["h"][k])
    say(f"    mixed vs FP64, worst tree: {worst:.2e} h")
    assert worst <= 1e-3
