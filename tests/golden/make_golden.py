"""Writes the golden fixtures in this folder from the reference's own code.

Run in a checkout that has the reference tree and its build (oracle/_ref/libadsref.so,
made by oracle/Makefile from the reference sources where they lie):

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickled objects):
  basis_p{p}_ne{ne}.npz   reference B-spline values/derivatives (bspline.cpp:102-160) at
                          the reference's Gauss points (quad/gauss.hpp) of every element,
                          as basis_data.cpp:63-114 tabulates them; span and point per (e, q)
  lu_p{p}_ne{ne}.npz      a Gram band matrix and its dgbtrf factors + pivots through the
                          reference's lin::factorize (band_solve.hpp:16-18, MKL LAPACK)
  ads_{shape}.npz         reference ads_solve (solver.hpp:222-226) on seeded right-hand
                          sides with the reference-factored Gram matrices of each direction
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

R = po.ref_lib()
if R is None:
    sys.exit("oracle/_ref/libadsref.so is not built (needs the reference tree)")


def ref_gauss(n):
    X, W = np.zeros(n), np.zeros(n)
    assert R.ref_gauss(n, po._d(X), po._d(W)) == 0
    return X, W


def basis_fixture(p, ne):
    ks = R.ref_create_basis(0.0, 1.0, p, ne, 0, None)
    knot = np.zeros(ks)
    R.ref_create_basis(0.0, 1.0, p, ne, 0, po._d(knot))
    X, W = ref_gauss(p + 1)
    pts = np.unique(knot)
    spans = np.zeros((ne, p + 1), dtype=np.int32)
    xs = np.zeros((ne, p + 1))
    ders = np.zeros((ne, p + 1, 2, p + 1))
    for e in range(ne):
        x1, x2 = pts[e], pts[e + 1]
        for q in range(p + 1):
            t = 0.5 * (X[q] + 1)
            x = (1 - t) * x1 + t * x2  # util.hpp:17-20 lerp
            s = R.ref_find_span(x, po._d(knot), ks, p)
            out = np.zeros((2, p + 1))
            R.ref_eval_basis_with_derivatives(s, x, po._d(knot), ks, p, 1, po._d(out))
            spans[e, q], xs[e, q], ders[e, q] = s, x, out
    np.savez(os.path.join(HERE, f"basis_p{p}_ne{ne}.npz"), knot=knot, gauss_x=X, gauss_w=W,
             spans=spans, points=xs, ders=ders)


def factored_gram(p, ne, fix):
    s = po.Sim(1, [p], [ne])
    if fix:
        s.fix_dof(0, 0)
        s.fix_dof(0, s.shape[0] - 1)
    ab, _ = s.matrix(0)
    n = s.shape[0]
    fac = ab.copy()
    ip = np.zeros(n, dtype=np.int32)
    info = R.ref_factorize(n, p, p, po._d(fac), po._i(ip))
    assert info == 0
    return ab, fac, ip


def lu_fixture(p, ne):
    ab, fac, ip = factored_gram(p, ne, fix=True)
    np.savez(os.path.join(HERE, f"lu_p{p}_ne{ne}.npz"), band=ab, factors=fac, ipiv=ip)


def ads_fixture(name, p, nes, seed):
    rng = np.random.default_rng(seed)
    mats = [factored_gram(p, ne, fix=False) for ne in nes]
    sizes = np.array([m[0].shape[0] for m in mats], dtype=np.int32)
    b = rng.standard_normal(int(np.prod(sizes)))
    x = b.copy()
    rank = len(nes)
    abp = (po._dp * rank)(*[po._d(m[1]) for m in mats])
    ipp = (po._ip * rank)(*[po._i(m[2]) for m in mats])
    kl = np.full(rank, p, dtype=np.int32)
    R.ref_ads_solve(rank, po._i(sizes), po._d(x), abp, ipp, po._i(kl), po._i(kl))
    np.savez(os.path.join(HERE, f"ads_{name}.npz"), p=p, elements=np.array(nes), rhs=b, solution=x)


if __name__ == "__main__":
    for p, ne in [(1, 4), (2, 4), (3, 5), (2, 64)]:
        basis_fixture(p, ne)
    for p, ne in [(2, 8), (3, 8)]:
        lu_fixture(p, ne)
    ads_fixture("2d_p2", 2, [5, 7], 1)
    ads_fixture("3d_p2", 2, [3, 4, 5], 2)
    ads_fixture("3d_p3", 3, [4, 3, 5], 3)
    print("golden fixtures written to", HERE)
