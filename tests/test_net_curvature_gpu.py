"""casl_curvature_learned (normalisation, network and hybrid blend fused in
one kernel) against the three separate calls it replaces
(casl_curvature_normalize(_3d), casl_mlp_forward, casl_curvature_hybrid(_3d)):
the same bits, in 2D and 3D, for random and trained networks, for sample
counts that leave partial tiles, with and without the residual form and the
3D deadband."""
import os

import numpy as np
import pytest

from casl_p4est_amd import curvature as CV

pytestmark = pytest.mark.gpu


def _features(dim, n, seed):
    rng = np.random.default_rng(seed)
    q = 9 if dim == 2 else 27
    f = rng.standard_normal((n, q + 1)).astype(np.float32)
    # |ihk| across the blend's range: below lo, inside [lo, hi], above hi
    f[:, q] = (rng.choice([-1, 1], n) * rng.uniform(0.0, 0.05, n)).astype(np.float32)
    return f


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [1, 63, 64, 129, 50000])
def test_fused_equals_separate_calls(cuda, dim, n):
    import torch
    dims = CV.DIMS_2D if dim == 2 else CV.DIMS_3D
    net = CV.CurvatureNet(dims, seed=5 + n, zero_last_layer=False)
    feat = torch.from_numpy(_features(dim, n, seed=n)).to(cuda)
    for residual in (True, False):
        for dead in ((0.0, 2e-3) if dim == 3 else (0.0,)):
            sep = CV.hybrid_hk(net, feat, residual=residual, fused=False, deadband=dead)
            fus = CV.hybrid_hk(net, feat, residual=residual, fused=True, deadband=dead)
            assert torch.equal(fus, sep), (residual, dead, float((fus - sep).abs().max()))


def test_fused_shipped_3d_network(cuda):
    import torch
    path = os.path.join(os.path.dirname(CV.__file__), "weights", "curvature3d.npz")
    net = CV.CurvatureNet.load(path)
    feat = torch.from_numpy(_features(3, 200003, seed=7)).to(cuda)
    assert torch.equal(CV.hybrid_hk(net, feat, fused=True), CV.hybrid_hk(net, feat, fused=False))
