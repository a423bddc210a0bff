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
