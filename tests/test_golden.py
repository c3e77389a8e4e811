"""The committed golden fixtures: the host flattener and the oracle still
reproduce them exactly (regression pins); the scalar known answers hold."""
import json
import os

import numpy as np
import pytest

import casl_p4est_amd as C
from casl_p4est_amd import levelset as ls
from oracle import pyoracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(G, name), allow_pickle=False)
    tab = C.FlatStencil(int(z["dim"]) if "dim" in z else 2, z["nbr"], z["d"], z["d_tr"], z["wall"], float(z["zdt"]))
    return z, tab


SHAPES = {
    "star2d.npz": (2, ls.Star()),
    "spheres3d.npz": (3, ls.Spheres([[0.3, 0.4, 0.5], [0.62, 0.55, 0.45], [0.5, 0.7, 0.65]], [0.15, 0.12, 0.1])),
}


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_flattener_reproduces_fixture(name):
    z, tab = load(name)
    dim, shape = SHAPES[name]
    f = C.Forest(dim).refine_levelset(shape, int(z["lmin"]), int(z["lmax"]), lip=2.4).build_nodes()
    t2 = C.flat_stencil_from_forest(f)
    assert np.array_equal(t2.nbr, tab.nbr) and np.array_equal(t2.d, tab.d) and np.array_equal(t2.d_tr, tab.d_tr)
    assert np.array_equal(f.sample(shape), z["phi0"])


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_oracle_reproduces_fixture(name):
    z, tab = load(name)
    phi0 = z["phi0"]
    assert np.array_equal(O.second_derivatives(tab, phi0), z["dd0"])
    for it in (1, 2, 20):
        assert np.array_equal(O.reinit(tab, phi0, 2, 2, it), z[f"rk2_{it}"])
    for os_, ot in ((1, 1), (2, 1), (1, 2)):
        assert np.array_equal(O.reinit(tab, phi0, os_, ot, 5), z[f"order{os_}{ot}_5"])


def test_uniform_fixture_second_differences():
    z, tab = load("uniform16.npz")
    dd = O.second_derivatives(tab, z["g"])
    ok = ~np.isnan(z["fd"])
    assert np.max(np.abs(dd[ok] - z["fd"][ok])) < 1e-10


def test_scalar_known_answers():
    kat = json.load(open(os.path.join(G, "scalars.json")))
    for a, b, r in kat["minmod"]:
        assert O.minmod(a, b) == r
    for a, b, fa, fb, r in kat["iface"]:
        assert O.interface_location(a, b, fa, fb) == r
    for row in kat["iface2"]:
        args, r, root = row[:6], row[6], row[7]
        assert O.interface_location_2nd(*args) == r
        if root is not None:
            assert abs(r - root) < 1e-14
