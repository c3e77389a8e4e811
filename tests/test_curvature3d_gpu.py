"""Learned curvature on any grid and in 3D, on the GPU:

  * casl_curvature_grid_samples equals the oracle's grid_samples (same nodes,
    phi/h features bit for bit, interpolated h*kappa to FP32 rounding) on
    adaptive 2D and 3D trees, with the nodes without a uniform block counted;
  * casl_curvature_normalize_3d equals the oracle bit for bit, and maps the 48
    symmetries of a stencil to one row; casl_curvature_hybrid_3d equals the
    oracle;
  * a 28-128x4-1 network trained on the GPU on spheres and wave surfaces beats
    the numerical h*kappa on held-out shapes, where the network is used
    (|ihk| >= 0.004), and the hybrid estimate is no worse overall.
"""
import itertools

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import curvature as CV
from casl_p4est_amd import levelset as ls
from oracle import curvature_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 4, 8), (3, 3, 6)])
def test_grid_samples_match_oracle(cuda, dim, lmin, lmax):
    import torch
    say(f"[curvature-grid] dim {dim}, lmin {lmin} / lmax {lmax}")
    shape = ls.Sphere([0.48] * dim, 0.27) if dim == 2 else ls.random_spheres(6, seed=3, rmin=0.08, rmax=0.15)
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=2.4).build_nodes()
    tab = C.flat_stencil_from_forest(f)
    ds = C.DeviceStencil(tab)
    h = float(f.dxyz_min()[0])
    phi = torch.from_numpy(f.sample(shape)).to(cuda)
    C.LevelSet(C.NodeNeighbors(tab)).reinitialize_2nd_order(phi, 10, precision="fp64")
    nrm = ds.normals(phi)
    kap = ds.mean_curvature(nrm)
    feat, off, node, skipped = CV.grid_samples(ds, phi, h, 1.5 * h, nrm, kap)
    of, oo, on, osk = O.grid_samples(tab, phi.cpu().numpy(), nrm.cpu().numpy().ravel(), kap.cpu().numpy(), h,
                                     1.5 * h)
    say(f"    {len(on)} samples, {osk} band nodes without a block")
    assert len(on) > 100
    assert skipped == osk
    assert np.array_equal(node.cpu().numpy(), on)
    fd = feat.cpu().numpy()
    q = 3 ** dim
    assert np.array_equal(fd[:, :q], of[:, :q])
    np.testing.assert_allclose(fd[:, q], of[:, q], rtol=2e-7, atol=1e-12)
    np.testing.assert_array_equal(off.cpu().numpy(), oo)


def test_normalize_and_hybrid_3d_match_oracle(cuda):
    import torch
    rng = np.random.default_rng(17)
    base = rng.standard_normal((40, 28)).astype(np.float32)
    base[:, 27] = (rng.choice([-1, 1], 40) * rng.uniform(0.0, 0.02, 40)).astype(np.float32)
    rows = [base]
    offs = O._block_offsets(3)
    pos = {o: q for q, o in enumerate(offs)}
    for perm in itertools.permutations(range(3)):
        for refl in itertools.product((-1, 1), repeat=3):
            sym = base.copy()
            for q, o in enumerate(offs):
                src = [0, 0, 0]
                for b in range(3):
                    src[perm[b]] = o[b] * refl[perm[b]]
                sym[:, q] = base[:, pos[tuple(src)]]
            rows.append(sym)
    feat = np.concatenate(rows)
    out = CV.normalize_3d(torch.from_numpy(feat).to(cuda)).cpu().numpy()
    assert np.array_equal(out, O.normalize_3d(feat))
    canon = out[:40]
    for k in range(1, len(rows)):
        assert np.array_equal(out[40 * k:40 * (k + 1)], canon)
    pred = rng.standard_normal(len(feat)).astype(np.float32) * 0.01
    for residual, dead in ((True, 0.0), (True, 0.005), (False, 0.0)):
        hk = CV.hybrid_3d(torch.from_numpy(feat).to(cuda), torch.from_numpy(pred).to(cuda), residual=residual,
                          deadband=dead)
        np.testing.assert_allclose(hk.cpu().numpy(), O.hybrid_3d(feat, pred, residual=residual, deadband=dead),
                                   rtol=1e-6, atol=1e-9)


def _errors(net, data):
    import torch
    hk = CV.hybrid_hk(net, data["feat"]).cpu().numpy().astype(np.float64)
    ihk = data["feat"][:, -1].cpu().numpy().astype(np.float64)
    t = data["target"]
    used = np.abs(ihk) >= CV.LO_MIN_HK
    return (float(np.mean(np.abs(hk[used] - t[used]))), float(np.mean(np.abs(ihk[used] - t[used]))),
            float(np.mean(np.abs(hk - t))), float(np.mean(np.abs(ihk - t))), int(used.sum()))


def test_learned_curvature_3d_beats_numerical(cuda):
    say("[curvature3d] building the box and the data")
    box = CV.UniformBox(level=6)
    train = CV.make_dataset_3d(box, 80, 80, seed=1, per_shape=2048, log=say)
    test = CV.make_dataset_3d(box, 20, 20, seed=2, per_shape=2048)
    say(f"    {train['feat'].shape[0]} training samples, {test['feat'].shape[0]} held out")
    net = CV.CurvatureNet(CV.DIMS_3D, seed=3)
    hist = CV.train(net, train["feat"], train["target"], epochs=12, batch=16384, lr=2e-3, log=say)
    assert hist["loss"][-1] < hist["loss"][0]
    le, ne, lall, nall, used = _errors(net, test)
    say(f"    held out, network used on {used}: learned {le:.3e} vs numerical {ne:.3e}; "
        f"all samples {lall:.3e} vs {nall:.3e}")
    assert le < ne
    assert lall <= nall


def test_shipped_3d_weights_beat_numerical(cuda):
    """The network under casl_p4est_amd/weights (trained by bench.py --config
    curvature3d) on shapes drawn from a seed it was not trained on."""
    import os
    path = os.path.join(os.path.dirname(CV.__file__), "weights", "curvature3d.npz")
    net = CV.CurvatureNet.load(path)
    assert net.dims == CV.DIMS_3D
    box = CV.UniformBox(level=6)
    data = CV.make_dataset_3d(box, 20, 20, seed=99, per_shape=2048)
    le, ne, lall, nall, used = _errors(net, data)
    say(f"    shipped weights, network used on {used}: learned {le:.3e} vs numerical {ne:.3e}")
    assert le < ne and lall < nall
