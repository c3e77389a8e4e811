"""Pins of the grid-sample, 3D normalisation and 3D hybrid parts of the
learned-curvature oracle (oracle/curvature_oracle.py):

  * the 3^dim block found by walking the construct_neighbors tables is the
    block found from node coordinates, on a uniform and on an adaptive tree;
  * the interpolated h*kappa reproduces a linear kappa field exactly at the
    (clamped) projection point, so a swapped weight or corner fails;
  * the 3D normalisation maps all 48 symmetries of a stencil to one row,
    whose canonical stencil (recovered from the output basis) holds the
    input's values and has non-negative, non-increasing centred differences;
    a linear field leaves no plane deviations;
  * the 3D hybrid at its end points and signs.
"""
import itertools

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import curvature_oracle as O


def _grid(dim, lmin, lmax, shape):
    f = C.Forest(dim).refine_levelset(shape, lmin, lmax, lip=1.2).build_nodes()
    tab = C.flat_stencil_from_forest(f)
    ixyz, xyz = f.nodes()
    return f, tab, ixyz, xyz


def _coord_index(ixyz):
    return {tuple(int(v) for v in row): n for n, row in enumerate(ixyz)}


@pytest.mark.parametrize("dim,lmin,lmax", [(2, 5, 5), (3, 3, 3), (2, 3, 6), (3, 2, 4)])
def test_grid_block_matches_coordinates(dim, lmin, lmax):
    shape = ls.Sphere([0.47] * dim, 0.31)
    f, tab, ixyz, xyz = _grid(dim, lmin, lmax, shape)
    h = float(f.dxyz_min()[0])
    step = 2 ** (C.forest.P4EST_MAXLEVEL[dim] - f.max_level)
    where = _coord_index(ixyz)
    found = 0
    for n in range(tab.n_nodes):
        blk = O.grid_block(tab, n, h)
        if blk is None:
            continue
        found += 1
        for o, m in blk.items():
            key = tuple(int(ixyz[n, a]) + o[a] * step for a in range(dim))
            assert where.get(key) == m, (n, o)
    assert found > 0
    if lmin == lmax:
        # every node at least one cell from every wall has a block
        M = 2 ** lmax
        inner = np.all((ixyz >= step) & (ixyz <= (M - 1) * step), axis=1)
        assert found == int(inner.sum())


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_samples_interpolate_linear_fields_exactly(dim):
    shape = ls.Sphere([0.5] * dim, 0.3)
    f, tab, ixyz, xyz = _grid(dim, 3, 5 if dim == 3 else 6, shape)
    h = float(f.dxyz_min()[0])
    phi = f.sample(shape)
    N = tab.n_nodes
    rng = np.random.default_rng(5)
    # normals: the sphere's, scaled so some projections leave the block (clamped)
    r = xyz - 0.5
    rn = np.linalg.norm(r, axis=1, keepdims=True)
    nrm = (r / np.where(rn > 0, rn, 1.0)).T.copy()
    phi_s = phi * 1.0
    a = rng.standard_normal(dim)
    b = rng.standard_normal()
    kappa = xyz @ a + b
    feat, off, node, skipped = O.grid_samples(tab, phi_s, nrm.ravel(), kappa, h, 1.5 * h)
    assert len(node) > 0
    for s in range(len(node)):
        n = node[s]
        u = np.clip(off[s] / h, -1.0, 1.0)
        expect = h * (float(xyz[n] @ a) + b + float(h * (u @ a)))
        assert feat[s, -1] == pytest.approx(expect, rel=2e-6, abs=1e-7)
        assert np.allclose(off[s], -phi[n] * nrm[:, n])
    # the phi features are phi/h over the block in x-slowest order
    where = _coord_index(ixyz)
    step = 2 ** (C.forest.P4EST_MAXLEVEL[dim] - f.max_level)
    for s in range(0, len(node), max(1, len(node) // 20)):
        n = node[s]
        for q, o in enumerate(O._block_offsets(dim)):
            m = where[tuple(int(ixyz[n, k]) + o[k] * step for k in range(dim))]
            assert feat[s, q] == np.float32(phi[m] / h)


def _apply_symmetry(stencil, perm, refl):
    """The stencil seen through axis permutation perm and reflections refl."""
    offs = O._block_offsets(3)
    pos = {o: q for q, o in enumerate(offs)}
    out = np.empty_like(stencil)
    for q, o in enumerate(offs):
        src = [0, 0, 0]
        for b in range(3):
            src[perm[b]] = o[b] * refl[perm[b]]
        out[q] = stencil[pos[tuple(src)]]
    return out


def test_normalize_3d_is_invariant_under_the_cube_symmetries():
    rng = np.random.default_rng(11)
    for trial in range(6):
        base = rng.standard_normal(28).astype(np.float32)
        base[27] = np.float32(rng.choice([-1, 1]) * rng.uniform(0.01, 0.5))
        rows = []
        for perm in itertools.permutations(range(3)):
            for refl in itertools.product((-1, 1), repeat=3):
                s = base.copy()
                s[:27] = _apply_symmetry(base[:27], perm, refl)
                rows.append(s)
        out = O.normalize_3d(np.asarray(rows))
        assert np.all(out == out[0])
        v = O.denormalize_3d(out[:1])[0]
        offs = O._block_offsets(3)
        pos = {o: q for q, o in enumerate(offs)}
        g = [v[pos[(1, 0, 0)]] - v[pos[(-1, 0, 0)]], v[pos[(0, 1, 0)]] - v[pos[(0, -1, 0)]],
             v[pos[(0, 0, 1)]] - v[pos[(0, 0, -1)]]]
        assert g[0] >= g[1] >= g[2] >= 0
        assert out[0, 27] == -abs(base[27])
        # the canonical stencil holds the input's values, negated when ihk > 0
        sg = -1.0 if base[27] > 0 else 1.0
        assert np.allclose(np.sort(v), np.sort(sg * base[:27].astype(np.float64)), rtol=0, atol=1e-5)


def test_normalize_3d_keeps_a_canonical_stencil():
    """A stencil already in canonical orientation keeps its values; the
    basis holds f0, the half differences and the plane deviations, which
    vanish for a linear field."""
    offs = np.array(O._block_offsets(3), dtype=np.float64)
    quad = 0.25 * (offs ** 2).sum(1)
    s = np.concatenate([0.5 + offs @ np.array([0.75, 0.5, 0.25]) + quad, [-0.2]]).astype(np.float32)
    out = O.normalize_3d(s[None])
    np.testing.assert_allclose(O.denormalize_3d(out)[0], s[:27], atol=1e-6)
    pos = {tuple(int(v) for v in o): q for q, o in enumerate(offs)}
    assert out[0, pos[(0, 0, 0)]] == np.float32(0.5)
    assert [out[0, pos[o]] for o in ((-1, 0, 0), (0, -1, 0), (0, 0, -1))] == [0.75, 0.5, 0.25]
    for q, o in enumerate(offs):
        if tuple(o) not in ((0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1)):
            assert out[0, q] == pytest.approx(quad[q], abs=1e-6)
    lin = np.concatenate([0.5 + offs @ np.array([0.75, 0.5, 0.25]), [-0.2]]).astype(np.float32)
    out = O.normalize_3d(lin[None])[0]
    rest = [q for q, o in enumerate(offs) if tuple(o) not in ((0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1))]
    assert np.max(np.abs(out[rest])) <= 1e-6


def test_hybrid_3d_end_points():
    feat = np.zeros((5, 28), np.float32)
    feat[:, 27] = [0.001, -0.003, 0.004, -0.007, 0.2]
    pred = np.array([9.0, 9.0, 0.05, 0.05, 0.05], np.float32)
    hk = O.hybrid_3d(feat, pred, residual=False)
    assert hk[0] == np.float32(0.001) and hk[1] == np.float32(-0.003)
    assert hk[2] == pytest.approx(0.004)  # lambda = 1 at lo: the numerical value
    assert hk[3] == pytest.approx(-0.05)  # lambda = 0 at hi: the network
    assert hk[4] == pytest.approx(0.05)
    hk = O.hybrid_3d(feat, pred, residual=True)
    assert hk[4] == pytest.approx(abs(0.05 - 0.2))
    # a correction inside the deadband leaves the numerical value
    hk = O.hybrid_3d(feat, pred, residual=True, deadband=0.06)
    assert hk[4] == pytest.approx(0.2)
    hk = O.hybrid_3d(feat, pred, residual=True, deadband=0.04)
    assert hk[4] == pytest.approx(abs(0.05 - 0.2))
