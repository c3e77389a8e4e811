"""Learned curvature on the GPU: the BF16 MFMA GEMM against numpy, the
network's forward/backward/Adam against the CPU oracle, numerical-curvature
samples and input normalisation against the oracle, and a short training run
showing what the network gains over numerical curvature.

Bars: GEMM within FP32-accumulation error of the exact product of the BF16
operands; network outputs and gradients within BF16 re-rounding noise of the
oracle (relative Frobenius error per layer <= 2e-2, cosine >= 0.999), Adam to
float32 rounding; samples bit-exact (features) and 1e-12 (projections);
normalisation bit-exact; hybrid to 2e-6 relative + 1e-8 absolute (the blend
weight uses fast float division; cancellation near the blend end points).
"""
import numpy as np
import pytest

from oracle import curvature_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def dev16(torch, a):
    return torch.from_numpy(O.bf16_bits(a).view(np.int16)).cuda()


@pytest.mark.parametrize("M,N,K", [(37, 70, 45), (64, 64, 16), (130, 5, 300), (1, 129, 10)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_bf16_gemm_layout(cuda, M, N, K, ta, tb):
    import torch
    from casl_p4est_amd import _lib as L
    from casl_p4est_amd.level_set import _stream_ptr
    import ctypes
    say(f"[gemm] M={M} N={N} K={K} ta={ta} tb={tb}")
    rng = np.random.default_rng(M * 1000 + N + K + 7 * ta + 13 * tb)
    A = O.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = O.bf16(rng.standard_normal((K, N)).astype(np.float32))
    a_store = A.T.copy() if ta else A
    b_store = B.T.copy() if tb else B
    lda = M if ta else K
    ldb = K if tb else N
    bias = rng.standard_normal(N).astype(np.float32)
    dA, dB = dev16(torch, a_store), dev16(torch, b_store)
    dbias = torch.from_numpy(bias).cuda()
    C32 = torch.full((M, N), np.nan, dtype=torch.float32, device=cuda)
    C16 = torch.zeros((M, N), dtype=torch.int16, device=cuda)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    L.call("casl_bf16_gemm", M, N, K, p(dA), lda, ta, p(dB), ldb, tb, None, 0, p(C32), None, N, _stream_ptr(torch))
    exact = A.astype(np.float64) @ B.astype(np.float64)
    got = C32.cpu().numpy()
    tol = 4 * K * 2.0 ** -24 * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) + 1e-30
    assert np.all(np.abs(got - exact) <= tol)
    # fused bias + ReLU, BF16 output
    L.call("casl_bf16_gemm", M, N, K, p(dA), lda, ta, p(dB), ldb, tb, p(dbias), 1, p(C32), p(C16), N,
           _stream_ptr(torch))
    ref = np.maximum(exact + bias, 0)
    got = C32.cpu().numpy()
    assert np.all(np.abs(got - ref) <= tol + 1e-6 * np.abs(ref))
    assert np.array_equal(C16.cpu().numpy().view(np.uint16), O.bf16_bits(got))


@pytest.mark.parametrize("M,N,K", [(128, 128, 128), (256, 64, 192), (64, 128, 1024)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_bf16_gemm_lds_path_same_bits(cuda, M, N, K, ta, tb):
    """Shapes in multiples of 64 take the LDS-staged kernel; the same
    matrices stored with a padded leading dimension (not a multiple of 8)
    take the direct one: the same bits, with bias and ReLU too."""
    import torch
    from casl_p4est_amd import _lib as L
    from casl_p4est_amd.level_set import _stream_ptr
    import ctypes
    say(f"[gemm lds] M={M} N={N} K={K} ta={ta} tb={tb}")
    rng = np.random.default_rng(M + 3 * N + 5 * K + 7 * ta + 11 * tb)
    A = O.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = O.bf16(rng.standard_normal((K, N)).astype(np.float32))
    a_store = A.T.copy() if ta else A
    b_store = B.T.copy() if tb else B
    bias = torch.from_numpy(rng.standard_normal(N).astype(np.float32)).cuda()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    outs = []
    for pad in (0, 1):
        sa = np.pad(a_store, ((0, 0), (0, pad)))
        sb = np.pad(b_store, ((0, 0), (0, pad)))
        dA, dB = dev16(torch, sa), dev16(torch, sb)
        C32 = torch.full((M, N), np.nan, dtype=torch.float32, device=cuda)
        C16 = torch.zeros((M, N), dtype=torch.int16, device=cuda)
        L.call("casl_bf16_gemm", M, N, K, p(dA), sa.shape[1], ta, p(dB), sb.shape[1], tb, None, 0, p(C32), None, N,
               _stream_ptr(torch))
        plain = C32.cpu().numpy()
        L.call("casl_bf16_gemm", M, N, K, p(dA), sa.shape[1], ta, p(dB), sb.shape[1], tb, p(bias), 1, p(C32),
               p(C16), N, _stream_ptr(torch))
        outs.append((plain, C32.cpu().numpy(), C16.cpu().numpy()))
    for x, y in zip(outs[0], outs[1]):
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    tol = 4 * K * 2.0 ** -24 * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) + 1e-30
    assert np.all(np.abs(outs[0][0] - exact) <= tol)


def _net_and_data(seed, n, dims=(10, 128, 128, 128, 128, 1)):
    from casl_p4est_amd.curvature import CurvatureNet
    rng = np.random.default_rng(seed)
    net = CurvatureNet(dims, seed=seed, zero_last_layer=False)
    mean = rng.standard_normal(dims[0]).astype(np.float32) * 0.1
    std = (0.5 + rng.random(dims[0])).astype(np.float32)
    net.set_scaler(mean, std)
    w, b = net.params()
    b = (rng.standard_normal(b.size) * 0.05).astype(np.float32)  # nonzero biases exercise the bias path
    net.set_params(w, b)
    x = rng.standard_normal((n, dims[0])).astype(np.float32)
    t = (rng.standard_normal(n) * 0.1).astype(np.float32)
    onet = O.Net.from_flat(dims, w, b, mean, std)
    return net, onet, x, t


def _rel_fro(a, b):
    return float(np.linalg.norm(a - b) / max(1e-30, np.linalg.norm(b)))


def test_mlp_forward_matches_oracle(cuda):
    import torch
    say("[mlp] forward n=777")
    net, onet, x, _ = _net_and_data(11, 777)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    yo, _ = O.forward(onet, x)
    assert y.shape == yo.shape
    assert _rel_fro(y, yo) < 1e-2
    assert np.max(np.abs(y - yo)) < 3e-2 * np.max(np.abs(yo))


@pytest.mark.parametrize("n", [300, 16384])
def test_mlp_gradients_match_oracle(cuda, n):
    import torch
    say(f"[mlp] backward n={n}")
    net, onet, x, t = _net_and_data(5 + n, n)
    loss = net.backward(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    gw, gb = net.grads()
    ogW, ogb, oloss = O.backward(onet, x, t)
    assert loss == pytest.approx(oloss, rel=1e-3)
    dims = net.dims
    ow = ob = 0
    for l, (i, o) in enumerate(zip(dims[:-1], dims[1:])):
        g = gw[ow:ow + i * o].reshape(i, o)
        e = _rel_fro(g, ogW[l])
        cos = float(np.sum(g * ogW[l]) / (np.linalg.norm(g) * np.linalg.norm(ogW[l])))
        say(f"[mlp]   layer {l}: dW rel {e:.2e} cos {cos:.6f}")
        assert e < 2e-2 and cos > 0.999
        gbl = gb[ob:ob + o]
        assert _rel_fro(gbl, ogb[l]) < 2e-2
        ow += i * o
        ob += o


def test_mlp_adam_step_matches_oracle(cuda):
    import torch
    say("[mlp] adam")
    net, onet, x, t = _net_and_data(3, 512)
    w0, b0 = net.params()
    net.backward(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    gw, gb = net.grads()
    net.adam_step(1e-3, 0.9, 0.999, 1e-8)
    w1, b1 = net.params()
    z = np.zeros_like(w0)
    ew, _, _ = O.adam_update(w0, gw, z, z, 1, 1e-3, 0.9, 0.999, 1e-8)
    eb, _, _ = O.adam_update(b0, gb, np.zeros_like(b0), np.zeros_like(b0), 1, 1e-3, 0.9, 0.999, 1e-8)
    np.testing.assert_allclose(w1, ew, rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(b1, eb, rtol=2e-6, atol=2e-7)
    # the BF16 copy used by the GEMMs follows the masters: forward equals the oracle's with the new weights
    onet1 = O.Net.from_flat(net.dims, w1, b1, onet.mean, onet.std)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    assert _rel_fro(y, O.forward(onet1, x)[0]) < 1e-2


@pytest.fixture(scope="module")
def small_patch(cuda):
    from casl_p4est_amd.curvature import UniformPatch
    return UniformPatch(level=6, size=1.0)


def test_curvature_samples_match_oracle(small_patch):
    import torch
    from casl_p4est_amd.curvature import circle_phi0, sinusoid_phi0
    P = small_patch
    say(f"[curv] samples M={P.M}")
    phis = np.stack([circle_phi0(P.xy, (0.51, 0.49), 0.23),
                     sinusoid_phi0(P.xy, (0.5, 0.5), 0.4, 0.08, 9.0, 0.3)])
    phi = torch.from_numpy(phis).cuda()
    P.reinitialize(phi, 20)
    feat, proj, node = P.samples(phi)
    feat, proj, node = feat.cpu().numpy(), proj.cpu().numpy(), node.cpu().numpy()
    grid = P.grid.cpu().numpy().reshape(P.M + 1, P.M + 1)
    phin = phi.cpu().numpy()
    k0 = 0
    for p in range(2):
        of, op, oij = O.curvature_samples(phin[p][grid], P.h)
        onode = p * P.n_nodes + grid[oij[:, 0], oij[:, 1]]
        order = np.argsort(onode)
        sel = (node // P.n_nodes) == p
        assert sel.sum() == len(of) > 20
        np.testing.assert_array_equal(node[sel], onode[order])
        np.testing.assert_array_equal(feat[sel], of[order])
        np.testing.assert_allclose(proj[sel], op[order], rtol=0, atol=1e-12)
        k0 += len(of)


def test_normalize_and_hybrid_match_oracle(cuda):
    import torch
    from casl_p4est_amd import curvature as CV
    say("[curv] normalize + hybrid")
    rng = np.random.default_rng(9)
    n = 5000
    feat = rng.standard_normal((n, 10)).astype(np.float32)
    feat[:, 9] = rng.uniform(-0.02, 0.02, n).astype(np.float32)
    feat[:10, 7] = feat[:10, 1]  # zero x-gradient: the EPS rule
    x = CV.normalize(torch.from_numpy(feat).cuda()).cpu().numpy()
    np.testing.assert_array_equal(x, O.normalize(feat))
    pred = (rng.standard_normal(2 * n) * 0.01).astype(np.float32)
    for residual in (True, False):
        hk = CV.hybrid(torch.from_numpy(feat).cuda(), torch.from_numpy(pred).cuda(), residual=residual)
        np.testing.assert_allclose(hk.cpu().numpy(), O.hybrid(feat, pred, residual=residual), rtol=2e-6, atol=1e-8)


def test_learned_curvature_beats_numerical(cuda):
    import torch
    from casl_p4est_amd import curvature as CV
    say("[curv] dataset")
    patch = CV.UniformPatch(level=8, size=2.0)
    train_set = CV.make_dataset(patch, 400, 120, seed=1)
    test_set = CV.make_dataset(patch, 60, 20, seed=2)
    say(f"[curv] train {len(train_set['target'])} samples, test {len(test_set['target'])}")
    net = CV.CurvatureNet(seed=4)
    hist = CV.train(net, train_set["feat"], train_set["target"], epochs=40, batch=16384, lr=2e-3)
    say(f"[curv] loss {hist['loss'][0]:.3e} -> {hist['loss'][-1]:.3e} in {hist['seconds']:.1f} s")
    assert hist["loss"][-1] < 0.5 * hist["loss"][0]
    hk = CV.hybrid_hk(net, test_set["feat"]).cpu().numpy().astype(np.float64)
    ihk = test_set["feat"][:, 9].cpu().numpy().astype(np.float64)
    t = test_set["target"]
    e_num, e_ml = np.abs(ihk - t).mean(), np.abs(hk - t).mean()
    say(f"[curv] mean |h kappa error|: numerical {e_num:.3e}, learned {e_ml:.3e}")
    assert e_ml < 0.8 * e_num
