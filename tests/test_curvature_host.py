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
