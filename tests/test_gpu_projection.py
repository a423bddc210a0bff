"""GPU L2 projection (adsg_project, projection.hpp:11-153) against the oracle's projector:
the load vector <= 1e-13 relative L2 in FP64, the reference's before() flow (fix the
boundary, project, solve; validation.hpp:65-71) <= 1e-12, FP32 handles (FP32 basis
tables, FP64 accumulation) <= 1e-6."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _f(*xs):
    v = np.sin(3.0 * xs[0] + 0.5) * np.cos(2.0 * xs[1] - 0.25)
    if len(xs) == 3:
        v = v * (1.0 + xs[2] ** 2)
    return v


def _fq(o):
    pts = [o.quad_points(d) for d in range(o.dim)]
    return np.ascontiguousarray(_f(*np.meshgrid(*pts, indexing="ij")).transpose()).ravel()


CASES = [(2, 1, (33, 20), None, None), (2, 2, (64, 47), [0.0, -1.0], [2.0, 1.0]),
         (2, 3, (30, 41), None, None), (3, 1, (17, 12, 9), None, None),
         (3, 2, (20, 17, 19), [0.0, 1.0, -2.0], [1.0, 2.5, 2.0]), (3, 3, (13, 11, 12), None, None)]


def _pair(dim, p, ne, a, b, precision=ag.FP64):
    g = ag.Simulation(dim, [p] * dim, list(ne), precision=precision, a=a, b=b)
    o = po.Sim(dim, [p] * dim, list(ne), a=(a + [0.0] * (3 - dim)) if a else None,
               b=(b + [1.0] * (3 - dim)) if b else None)
    return g, o


@pytest.mark.parametrize("dim,p,ne,a,b", CASES)
def test_load_vector_matches_oracle(dim, p, ne, a, b):
    _log(f"case: projection load dim={dim} p={p} ne={ne}")
    g, o = _pair(dim, p, ne, a, b)
    for d in range(dim):
        assert np.max(np.abs(g.quad_points(d) - o.quad_points(d))) <= 1e-15
    fq = _fq(o)
    g.set(ag.U, np.zeros(g.size))
    g.project(fq)
    expect = o.project(fq)
    assert po.rel_l2(g.get(ag.U), expect) <= 1e-13
    # the callable path evaluates f at the same points in the same layout
    g.set(ag.U, np.zeros(g.size))
    g.project(_f)
    assert po.rel_l2(g.get(ag.U), expect) <= 1e-13
    # adds into what the vector holds, like the reference's projector
    v0 = np.random.default_rng(1).standard_normal(g.size)
    g.set(ag.U_PREV, v0)
    g.project(fq, which=ag.U_PREV)
    assert po.rel_l2(g.get(ag.U_PREV), v0 + expect) <= 1e-13
    g.close()


@pytest.mark.parametrize("dim,p,ne", [(2, 2, (64, 64)), (3, 2, (12, 14, 10)), (3, 3, (10, 9, 11))])
def test_before_flow_matches_oracle(dim, p, ne):
    """validation.hpp:65-71: fix the boundary rows, prepare, project, solve."""
    _log(f"case: projection + solve dim={dim} p={p} ne={ne}")
    g, o = _pair(dim, p, ne, None, None)
    g.fix_all_boundaries()
    o.fix_all_boundaries()
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    fq = _fq(o)
    g.set(ag.U, np.zeros(g.size))
    g.project(fq)
    g.ads_solve()
    expect = o.solve(o.project(fq))
    assert po.rel_l2(g.get(ag.U), expect) <= 1e-12
    g.close()


def test_fp32_handle_projects_in_fp64():
    _log("case: projection FP32 handle 3D p=2")
    g, o = _pair(3, 2, (20, 17, 19), None, None, precision=ag.FP32)
    fq = _fq(o)
    g.set(ag.U, np.zeros(g.size))
    g.project(fq)
    assert po.rel_l2(g.get(ag.U), o.project(fq)) <= 1e-6
    with pytest.raises(ag.AdsgError):  # the FP32 increment buffer is not an FP64 vector
        g.project(fq, which=ag.BUFFER)
    with pytest.raises(ValueError):
        g.project(fq[:-1])
    g.close()


def _exact(dim, grid):
    """s = sin(pi x) sin(pi y) (sin(pi z)) and its gradient at the Gauss points."""
    s = [np.sin(np.pi * g) for g in grid]
    c = [np.pi * np.cos(np.pi * g) for g in grid]
    val = np.prod(s, axis=0)
    planes = [val]
    for d in range(dim):
        planes.append(np.prod([c[i] if i == d else s[i] for i in range(dim)], axis=0))
    return np.concatenate([p.ravel() for p in planes])


@pytest.mark.parametrize("dim,p,ne,a,b", CASES)
def test_error_functionals_match_oracle(dim, p, ne, a, b):
    _log(f"case: error functionals dim={dim} p={p} ne={ne}")
    g, o = _pair(dim, p, ne, a, b)
    sq = _exact(dim, g.quad_grid())
    u = np.random.default_rng(2).standard_normal(g.size) * 1e-2
    g.set(ag.U, u)
    for kind in (0, 1, 2):
        for which, uu in ((ag.U, u), (None, None)):
            got = g.error(sq, kind, which)
            expect = np.sqrt(o.error_sq(uu, sq, kind))
            assert abs(got - expect) <= 1e-12 * expect, (kind, which)
    g.close()


def test_error_of_the_projection_converges():
    """The L2 projection of sin(pi x) sin(pi y) at p = 2 converges at order p + 1 = 3 in L2
    and p = 2 in H1 (halving h divides the errors by about 8 and 4)."""
    _log("case: projection error convergence 2D p=2")
    errs = []
    for n in (16, 32):
        g = ag.Simulation(2, [2, 2], [n, n])
        g.prepare_matrices()
        g.set(ag.U, np.zeros(g.size))
        g.project(lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y))
        g.ads_solve()
        sq = _exact(2, g.quad_grid())
        errs.append((g.error(sq, 0) / g.error(sq, 0, None), g.error(sq, 2) / g.error(sq, 2, None)))
        g.close()
    assert 7.0 < errs[0][0] / errs[1][0] < 9.0
    assert 3.5 < errs[0][1] / errs[1][1] < 4.5
