"""Pins of the learned-curvature CPU oracle (oracle/curvature_oracle.py)."""
import numpy as np
import pytest

from oracle import curvature_oracle as O


def test_bf16_rounding_bit_patterns():
    x = np.array([1.0, -2.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7 + 2.0 ** -9, 3.0e38, 0.0,
                  np.inf, 1e-45], dtype=np.float32)
    bits = O.bf16_bits(x)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6); 1+2^-7+2^-9 rounds down to 1+2^-7
    assert [int(b) for b in bits] == [0x3F80, 0xC000, 0x3F80, 0x3F82, 0x3F81, 0x7F62, 0x0000, 0x7F80, 0x0000]
    assert O.bf16_bits(np.array([np.nan], np.float32))[0] == 0x7FC0
    v = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    r = O.bf16(v)
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)
    assert np.array_equal(O.bf16(r), r)  # idempotent


def _tiny_net(rng, dims):
    W = [rng.standard_normal((i, o)) * np.sqrt(2.0 / i) for i, o in zip(dims[:-1], dims[1:])]
    b = [rng.standard_normal(o) * 0.1 for o in dims[1:]]
    net = O.Net(W, b, mean=rng.standard_normal(dims[0]) * 0.1, std=1.0 + rng.random(dims[0]))
    # float64 masters for the finite-difference check
    net.W = [w.astype(np.float64) for w in W]
    net.b = [v.astype(np.float64) for v in b]
    return net


def test_backward_matches_finite_differences_float64():
    rng = np.random.default_rng(7)
    dims = (3, 5, 4, 1)
    net = _tiny_net(rng, dims)
    x = rng.standard_normal((7, 3))
    t = rng.standard_normal(7)
    gW, gb, loss = O.backward(net, x, t, emulate_bf16=False)

    def loss_of():
        y, _ = O.forward(net, x, emulate_bf16=False)
        return float(np.mean((y.ravel() - t) ** 2))

    assert loss == pytest.approx(loss_of(), rel=1e-14)
    eps = 1e-6
    for l in range(len(dims) - 1):
        for arr, g in ((net.W[l], gW[l]), (net.b[l], gb[l])):
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + eps
                lp = loss_of()
                arr[idx] = old - eps
                lm = loss_of()
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * eps)
            np.testing.assert_allclose(g, num, rtol=1e-5, atol=1e-8)


def test_bf16_forward_close_to_float64():
    rng = np.random.default_rng(3)
    dims = (10, 64, 64, 1)
    net = _tiny_net(rng, dims)
    x = rng.standard_normal((50, 10))
    y64, _ = O.forward(net, x, emulate_bf16=False)
    net32 = O.Net([w.astype(np.float32) for w in net.W], [v.astype(np.float32) for v in net.b], net.mean, net.std)
    y16, acts = O.forward(net32, x.astype(np.float32), emulate_bf16=True)
    assert all(np.array_equal(O.bf16(a), a) for a in acts)
    err = np.abs(y16 - y64).max() / np.abs(y64).max()
    assert 1e-5 < err < 3e-2  # BF16 rounding is visible but small


def test_adam_update_first_step_is_sign_step():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -1e-3, 0.0], np.float32)
    z = np.zeros(3, np.float32)
    p1, m, v = O.adam_update(p, g, z, z, 1, 1e-2, 0.9, 0.999, 1e-8)
    # first Adam step: m_hat = g, v_hat = g^2 -> p - lr * sign(g) (g = 0 leaves p)
    np.testing.assert_allclose(p1, [1.0 - 1e-2, -2.0 + 1e-2, 0.5], rtol=1e-6)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=5e-5)  # 1 - 0.999 in float32


def _linear_stencil(gx, gy, ihk):
    """phi = gx x + gy y sampled at the 3x3 offsets (x slowest)."""
    st = [gx * di + gy * dj for di in (-1, 0, 1) for dj in (-1, 0, 1)]
    return np.array(st + [ihk], np.float32)


@pytest.mark.parametrize("gx,gy", [(0.6, 0.8), (-0.6, 0.8), (-0.6, -0.8), (0.6, -0.8), (0.0, 1.0), (-1.0, 0.0)])
def test_normalize_rotates_gradient_into_first_quadrant(gx, gy):
    s = _linear_stencil(gx, gy, -0.01)[None, :]
    out = O.normalize(s)
    for row in out:
        ngx, ngy = row[7] - row[1], row[5] - row[3]
        assert ngx >= 0 and ngy >= 0
        assert np.hypot(ngx, ngy) == pytest.approx(2.0, rel=1e-6)  # rotation keeps the magnitude
    # second row is the first reflected about y = x: gradient components swap
    assert out[1, 7] - out[1, 1] == pytest.approx(out[0, 5] - out[0, 3])
    assert out[0, 4] == s[0, 4] and out[0, 9] == s[0, 9]


def test_normalize_sign_flip_and_identity_cases():
    s = _linear_stencil(0.6, 0.8, 0.02)
    out = O.normalize(s[None, :])
    assert out[0, 9] == np.float32(-0.02)
    # negated field has gradient (-0.6, -0.8) -> rotated by 180 degrees back to (0.6, 0.8), negated values
    np.testing.assert_array_equal(out[0, :9], -s[O.ROT_HALF])
    q = _linear_stencil(0.6, 0.8, -0.02)
    np.testing.assert_array_equal(O.normalize(q[None, :])[0], q)  # already normalised
    # four quarter turns are the identity; CW undoes CCW
    idx = np.arange(9)
    for _ in range(4):
        idx = idx[O.ROT_CCW]
    assert np.array_equal(idx, np.arange(9))
    assert np.array_equal(np.arange(9)[O.ROT_CCW][O.ROT_CW], np.arange(9))
    assert np.array_equal(O.REFLECT[O.REFLECT], np.arange(9))


def test_hybrid_blend_end_points():
    feat = np.zeros((5, 10), np.float32)
    feat[:, 9] = [0.001, -0.003, 0.007, -0.0055, 0.2]
    pred = np.array([-0.5, -0.5, 9, 9, -0.03, -0.05, 0.0, 0.0, -0.21, -0.19], np.float32)
    hk = O.hybrid(feat, pred, residual=False)
    assert hk[0] == np.float32(0.001) and hk[1] == np.float32(-0.003)  # |ihk| < lo: numerical value
    assert hk[2] == pytest.approx(0.04)       # |ihk| = hi: network only, sign of ihk
    lam = (0.007 - 0.0055) / 0.003            # halfway: half network, half -|ihk|
    assert hk[3] == pytest.approx(-abs((1 - lam) * 0.0 + lam * -0.0055), rel=1e-5)
    assert hk[4] == pytest.approx(0.2)
    hr = O.hybrid(feat, pred, residual=True)
    assert hr[4] == pytest.approx(abs(-0.2 - 0.2))  # residual adds -|ihk|


def test_central_difference_matches_numpy_gradient():
    rng = np.random.default_rng(0)
    f = rng.standard_normal(20)
    h = 0.125
    np.testing.assert_allclose(O.central(f[2:], f[1:-1], f[:-2], h), np.gradient(f, h)[1:-1], rtol=1e-13)


@pytest.mark.parametrize("R", [0.15, 0.3])
def test_numerical_curvature_of_circle_converges(R):
    errs = []
    for M in (64, 128):
        h = 1.0 / M
        x = np.arange(M + 1) * h
        X, Y = np.meshgrid(x, x, indexing="ij")
        phi = np.hypot(X - 0.5 - 0.3 * h, Y - 0.5 - 0.1 * h) - R
        feat, proj, ij = O.curvature_samples(phi, h)
        assert len(feat) > 4 * R / h
        np.testing.assert_allclose(np.hypot(proj[:, 0] - 0.5 - 0.3 * h, proj[:, 1] - 0.5 - 0.1 * h), R, atol=h * h * 4)
        errs.append(np.abs(feat[:, 9].astype(np.float64) / h - 1.0 / R).max())
    assert errs[1] < errs[0] / 2.5   # second order in h for kappa
    assert errs[1] * R < 0.01


def test_samples_layout_x_slowest():
    M, h = 16, 0.25
    x = np.arange(M + 1) * h
    X, Y = np.meshgrid(x, x, indexing="ij")
    phi = Y - 2.05 + 0.01 * X  # a nearly horizontal line
    feat, proj, ij = O.curvature_samples(phi, h)
    assert len(feat) > 0
    i, j = ij[0]
    expect = [phi[i + di, j + dj] / h for di in (-1, 0, 1) for dj in (-1, 0, 1)]
    np.testing.assert_array_equal(feat[0, :9], np.float32(expect))
    assert np.abs(feat[:, 9]).max() < 1e-12  # straight line: zero curvature
