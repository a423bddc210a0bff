"""The oracle against the golden fixtures written from the reference's own code by
tests/golden/make_golden.py (reference B-splines at the reference Gauss points, MKL
factors through lin::factorize, the reference ads_solve). These run anywhere, including
where the reference tree is absent."""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("p,ne", [(1, 4), (2, 4), (3, 5), (2, 64)])
def test_basis_tables_match_reference(p, ne):
    g = np.load(os.path.join(GOLD, f"basis_p{p}_ne{ne}.npz"))
    assert np.array_equal(po.knots(0.0, 1.0, p, ne), g["knot"])
    X, W = po.gauss(p + 1)
    assert np.max(np.abs(X - g["gauss_x"])) <= 1.2e-16 and np.max(np.abs(W - g["gauss_w"])) <= 2.3e-16
    s = po.Sim(1, [p], [ne])
    t = s.tables(0)
    # same points (to the Gauss-node rounding), same values and derivatives
    assert np.max(np.abs(t["x"] - g["points"])) <= 1e-15
    assert np.max(np.abs(t["b"] - g["ders"])) <= 1e-12 * max(1.0, np.abs(g["ders"]).max())
    for e in range(ne):
        for q in range(p + 1):
            assert po.find_span(g["points"][e, q], g["knot"], p) == g["spans"][e, q]


@pytest.mark.parametrize("p,ne", [(2, 8), (3, 8)])
def test_band_lu_matches_reference(p, ne):
    g = np.load(os.path.join(GOLD, f"lu_p{p}_ne{ne}.npz"))
    s = po.Sim(1, [p], [ne])
    s.fix_dof(0, 0)
    s.fix_dof(0, s.shape[0] - 1)
    ab, _ = s.matrix(0)
    assert np.allclose(ab, g["band"], rtol=1e-14, atol=1e-18)
    fac = ab.copy()
    ip, info = po.dgbtrf(fac, s.shape[0], p, p)
    assert info == 0 and np.array_equal(ip, g["ipiv"])
    assert np.max(np.abs(fac - g["factors"])) <= 1e-15 * np.abs(g["factors"]).max()


@pytest.mark.parametrize("name", ["2d_p2", "3d_p2", "3d_p3"])
def test_ads_solve_matches_reference(name):
    g = np.load(os.path.join(GOLD, f"ads_{name}.npz"))
    p, nes = int(g["p"]), [int(v) for v in g["elements"]]
    s = po.Sim(len(nes), [p] * len(nes), nes)
    s.prepare_matrices()
    x = s.solve(g["rhs"].copy())
    assert po.rel_l2(x, g["solution"]) <= 1e-13
