"""Host-side pieces of the learned-curvature data pipeline (no GPU)."""
import math

import numpy as np
import pytest

from casl_p4est_amd import curvature as CV


def test_circle_phi0_zero_set_and_sign():
    xy = np.array([[0.5 + 0.2, 0.5], [0.5, 0.5], [0.5, 0.9]])
    phi = CV.circle_phi0(xy, (0.5, 0.5), 0.2)
    assert abs(phi[0]) < 1e-15 and phi[1] < 0 and phi[2] > 0
    # gradient magnitude |x - c| / R: a signed distance only on the circle
    assert CV.circle_phi0(np.array([[0.9, 0.5]]), (0.5, 0.5), 0.2)[0] == pytest.approx((0.16 - 0.04) / 0.4, rel=1e-14)


def test_sinusoid_closest_point_curvature_against_dense_search():
    rng = np.random.default_rng(0)
    h = 2.0 ** -7
    for _ in range(5):
        c = (1.0 + rng.uniform(-h, h), 1.0 + rng.uniform(-h, h))
        theta, A = rng.uniform(0, 2 * math.pi), rng.uniform(0.01, 0.2)
        omega, phase = rng.uniform(2 * math.pi, 12 * math.pi), rng.uniform(0, 2 * math.pi)
        # points within ~1.2h of the curve
        s0 = rng.uniform(-0.5, 0.5, 200)
        ct, st = math.cos(theta), math.sin(theta)
        u = s0 + rng.uniform(-0.5, 0.5, 200) * h
        v = A * np.sin(omega * s0 + phase) + rng.uniform(-1.2, 1.2, 200) * h
        pts = np.stack([c[0] + ct * u - st * v, c[1] + st * u + ct * v], axis=1)
        hk = CV.sinusoid_closest_hk(pts, c, theta, A, omega, phase, h)
        # dense search for the closest curve parameter
        s = u[:, None] + np.linspace(-4 * h, 4 * h, 20001)[None, :]
        d2 = (s - u[:, None]) ** 2 + (A * np.sin(omega * s + phase) - v[:, None]) ** 2
        sb = s[np.arange(len(u)), np.argmin(d2, axis=1)]
        f1 = A * omega * np.cos(omega * sb + phase)
        f2 = -A * omega ** 2 * np.sin(omega * sb + phase)
        ref = h * (-f2 / (1 + f1 * f1) ** 1.5)
        # grid resolution of the dense search: 8h/20000 in s
        assert np.max(np.abs(hk - ref)) < 1e-3 * max(1e-3, np.abs(ref).max())


def test_sinusoid_phi0_sign_convention():
    # v - A sin(omega u + phase): negative below the curve, kappa = -f''/(1+f'^2)^1.5
    xy = np.array([[1.0, 0.5], [1.0, 1.5]])
    phi = CV.sinusoid_phi0(xy, (1.0, 1.0), 0.0, 0.1, 2 * math.pi, 0.0)
    assert phi[0] < 0 < phi[1]
    # at u = pi/(2 omega) the curve has its maximum: f'' < 0 -> kappa > 0 (convex side of phi < 0)
    top = np.array([[1.0 + 0.25, 1.0 + 0.1]])
    hk = CV.sinusoid_closest_hk(top, (1.0, 1.0), 0.0, 0.1, 2 * math.pi, 0.0, 1.0)
    assert hk[0] == np.float64(0.1 * (2 * math.pi) ** 2)


# ------------------------------------------------------------------ 3D targets
def _fd_grad_hess(phi, x, eps=1e-5):
    g = np.empty((x.shape[0], 3))
    H = np.empty((x.shape[0], 3, 3))
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        g[:, a] = (phi(x + e) - phi(x - e)) / (2 * eps)
        for b in range(3):
            f = np.zeros(3)
            f[b] = eps
            H[:, a, b] = (phi(x + e + f) - phi(x + e - f) - phi(x - e + f) + phi(x - e - f)) / (4 * eps * eps)
    return g, H


def test_gyroid_gradient_and_hessian_against_finite_differences():
    phi, gh = CV.gyroid_fields()
    x = np.random.default_rng(2).uniform(0, 1, (50, 3))
    g, H = gh(x)
    gf, Hf = _fd_grad_hess(phi, x)
    np.testing.assert_allclose(g, gf, rtol=1e-7, atol=1e-6)
    np.testing.assert_allclose(H, Hf, rtol=1e-4, atol=2e-3 * np.abs(H).max())


def test_implicit_curvature_of_a_sphere_is_two_over_r():
    c, R, h = np.array([0.4, 0.5, 0.6]), 0.23, 2.0 ** -9

    def phi(x):
        return ((x - c) ** 2).sum(axis=1) - R * R

    def gh(x):
        H = np.broadcast_to(2 * np.eye(3), (x.shape[0], 3, 3)).copy()
        return 2 * (x - c), H

    d = np.random.default_rng(4).standard_normal((40, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = c + d * (R + np.random.default_rng(5).uniform(-1.5, 1.5, (40, 1)) * h)
    np.testing.assert_allclose(CV.implicit_closest_hk(phi, gh, pts, h), 2 * h / R, rtol=1e-12)


def test_wave_curvature_agrees_with_the_implicit_formula_on_the_surface():
    rng = np.random.default_rng(8)
    h = 2.0 ** -7
    for _ in range(4):
        c = 0.5 + rng.uniform(-h, h, 3)
        Q = CV.random_frame(rng)
        assert np.allclose(Q.T @ Q, np.eye(3), atol=1e-12)
        A, om, a, b = rng.uniform(0.01, 0.2), rng.uniform(2 * math.pi, 12 * math.pi), rng.uniform(0, 6), rng.uniform(0, 6)
        u, v = rng.uniform(-0.4, 0.4, (2, 60))
        w = A * np.sin(om * u + a) * np.sin(om * v + b)
        pts = c + np.stack([u, v, w], axis=1) @ Q.T
        assert np.max(np.abs(CV.wave_phi0(pts, c, Q, A, om, a, b))) < 1e-14

        def phi(x):
            return CV.wave_phi0(x, c, Q, A, om, a, b)

        def gh(x):
            p = (x - c) @ Q
            su, cu = np.sin(om * p[:, 0] + a), np.cos(om * p[:, 0] + a)
            sv, cv = np.sin(om * p[:, 1] + b), np.cos(om * p[:, 1] + b)
            gl = np.stack([-A * om * cu * sv, -A * om * su * cv, np.ones(len(p))], axis=1)
            Hl = np.zeros((len(p), 3, 3))
            Hl[:, 0, 0] = A * om * om * su * sv
            Hl[:, 1, 1] = A * om * om * su * sv
            Hl[:, 0, 1] = Hl[:, 1, 0] = -A * om * om * cu * cv
            return gl @ Q.T, np.einsum("ia,nab,jb->nij", Q, Hl, Q)

        e = CV.wave_closest_hk(pts, c, Q, A, om, a, b, h)
        f = CV.implicit_closest_hk(phi, gh, pts, h)
        np.testing.assert_allclose(e, f, rtol=1e-9, atol=1e-12)


def test_shipped_3d_weights_file():
    import os
    path = os.path.join(os.path.dirname(CV.__file__), "weights", "curvature3d.npz")
    z = np.load(path, allow_pickle=False)
    dims = tuple(int(d) for d in z["dims"])
    assert dims == CV.DIMS_3D
    nw = sum(i * o for i, o in zip(dims[:-1], dims[1:]))
    assert z["w"].shape == (nw,) and z["b"].shape == (sum(dims[1:]),)
    assert z["mean"].shape == (28,) and np.all(z["std"] > 0)
    assert np.all(np.isfinite(z["w"])) and np.all(np.isfinite(z["b"]))
