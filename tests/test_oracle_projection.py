"""Pins of the oracle's L2 projection (orc_sim_project, projection.hpp:11-153) and error
functional (orc_sim_error_sq, basic_simulation_2d.hpp:318-336) before they are trusted as
the checkers of adsg_project and adsg_error. The reference's projector needs basis_data and
Boost, so it is not compiled here (oracle/_ref); the pins are what mathematics fixes:

* a spline f = sum_j c_j B_j is integrated exactly by p+1 Gauss points per direction
  (degree 2p), so its load vector is (M_z (x) M_y (x) M_x) c with the 1D Gram matrices
  (themselves pinned by closed forms in test_oracle_pins), and the projection followed by
  the Gram solve returns c;
* f = 1 gives b = M 1, whose entries sum to the volume of the box (partition of unity);
* the load is linear in f and adds into the given vector;
* the error of u against its own spline values (and derivatives) at the Gauss points is 0,
  against the spline of v it is the norm of u - v (orc_sim_norm_sq, pinned by quadratic
  forms in test_oracle_rhs), and with u = 0 it is the sum of the load of s^2.
"""
import numpy as np
import pytest

import pyoracle as po

CASES = [(2, [1, 1], [6, 5], [0.0, -1.0], [2.0, 1.0]),
         (2, [2, 2], [7, 9], None, None),
         (2, [3, 3], [8, 6], [-0.5, 0.0], [0.5, 3.0]),
         (3, [1, 1, 1], [4, 5, 3], None, None),
         (3, [2, 2, 2], [5, 4, 6], [0.0, 1.0, -2.0], [1.0, 2.5, 2.0]),
         (3, [3, 3, 3], [5, 5, 4], None, None)]


def _dense(ab, n, p):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - p), min(n, i + p + 1)):
            A[i, j] = ab[j, 2 * p + i - j]
    return A


def _collocation(sim, d, der=0):
    """Phi[e*q + qx, first(e) + a] = B_a(x_q) (der=1: B_a'(x_q)) on element e: the spline's
    values (derivatives) at the Gauss points of direction d are Phi @ c along d."""
    t = sim.tables(d)
    ne, q = t["x"].shape
    Phi = np.zeros((ne * q, sim.shape[d]))
    for e in range(ne):
        for k in range(q):
            Phi[e * q + k, t["first"][e]:t["first"][e] + sim.p[d] + 1] = t["b"][e, k, der, :]
    return Phi


def _spline_planes(sim, c):
    """[value | d/dx | d/dy (| d/dz)] of the spline with coefficients c at the Gauss points."""
    V = [_collocation(sim, d) for d in range(sim.dim)]
    D = [_collocation(sim, d, 1) for d in range(sim.dim)]
    planes = [_apply(V, c, sim.shape)]
    for d in range(sim.dim):
        mats = [D[i] if i == d else V[i] for i in range(sim.dim)]
        planes.append(_apply(mats, c, sim.shape))
    return np.concatenate(planes)


def _apply(mats, c, shape):
    """(A_z (x) A_y (x) A_x) c for c with x fastest."""
    C = c.reshape(shape[::-1])
    if len(shape) == 2:
        return np.einsum("yj,xi,ji->yx", mats[1], mats[0], C).ravel()
    return np.einsum("zk,yj,xi,kji->zyx", mats[2], mats[1], mats[0], C).ravel()


def _sim(dim, p, ne, a, b):
    return po.Sim(dim, p, ne, a=(a + [0.0] * (3 - dim)) if a else None,
                  b=(b + [1.0] * (3 - dim)) if b else None)


@pytest.mark.parametrize("dim,p,ne,a,b", CASES)
def test_projection_of_a_spline_is_its_gram_product(dim, p, ne, a, b):
    s = _sim(dim, p, ne, a, b)
    c = np.random.default_rng(3).standard_normal(s.size)
    Phis = [_collocation(s, d) for d in range(dim)]
    fq = _apply(Phis, c, s.shape)  # the spline at the Gauss points, x fastest
    M = [_dense(s.matrix(d, 0)[0], s.shape[d], p[d]) for d in range(dim)]
    load = s.project(fq)
    expect = _apply(M, c, s.shape)
    assert po.rel_l2(load, expect) <= 1e-13
    s.prepare_matrices()
    back = s.solve(load.copy())
    assert po.rel_l2(back, c) <= 1e-11


@pytest.mark.parametrize("dim,p,ne,a,b", CASES)
def test_projection_of_one_sums_to_the_volume(dim, p, ne, a, b):
    s = _sim(dim, p, ne, a, b)
    nq = int(np.prod([ne[d] * (p[d] + 1) for d in range(dim)]))
    load = s.project(np.ones(nq))
    lo = a if a else [0.0] * dim
    hi = b if b else [1.0] * dim
    vol = float(np.prod([hi[d] - lo[d] for d in range(dim)]))
    assert abs(load.sum() - vol) <= 1e-13 * vol
    M = [_dense(s.matrix(d, 0)[0], s.shape[d], p[d]) for d in range(dim)]
    assert po.rel_l2(load, _apply(M, np.ones(s.size), s.shape)) <= 1e-14


def test_projection_is_linear_and_adds_into_the_vector():
    s = po.Sim(2, [2, 2], [9, 7])
    rng = np.random.default_rng(5)
    nq = 9 * 3 * 7 * 3
    f, g = rng.standard_normal(nq), rng.standard_normal(nq)
    v0 = rng.standard_normal(s.size)
    lf, lg = s.project(f), s.project(g)
    both = s.project(2.0 * f - 3.0 * g, v0.copy())
    assert po.rel_l2(both, v0 + 2.0 * lf - 3.0 * lg) <= 1e-15


def test_quad_points_are_the_mapped_gauss_points():
    s = po.Sim(2, [2, 3], [5, 4], a=[0.0, -1.0, 0.0], b=[2.0, 1.0, 1.0])
    for d, (lo, hi, ne, q) in enumerate([(0.0, 2.0, 5, 3), (-1.0, 1.0, 4, 4)]):
        X, _ = np.polynomial.legendre.leggauss(q)
        h = (hi - lo) / ne
        expect = np.concatenate([lo + h * e + h * (X + 1) / 2 for e in range(ne)])
        assert np.max(np.abs(s.quad_points(d) - expect)) <= 1e-15


@pytest.mark.parametrize("dim,p,ne,a,b", CASES)
def test_error_functional_pins(dim, p, ne, a, b):
    s = _sim(dim, p, ne, a, b)
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal(s.size), rng.standard_normal(s.size)
    su, sv = _spline_planes(s, u), _spline_planes(s, v)
    for kind in (0, 1, 2):
        ref = s.norm_sq(u, kind)
        assert s.error_sq(u, su, kind) <= 1e-26 * ref
        assert abs(s.error_sq(u, sv, kind) - s.norm_sq(u - v, kind)) <= 1e-13 * s.norm_sq(u - v, kind)
        assert abs(s.error_sq(None, su, kind) - ref) <= 1e-13 * ref
    # u = 0, kind 0: the integral of s^2, i.e. the sum of the load vector of s^2
    Q = su.size // (1 + dim)
    load = s.project(su[:Q] ** 2)
    assert abs(s.error_sq(None, su[:Q], 0) - load.sum()) <= 1e-13 * load.sum()
