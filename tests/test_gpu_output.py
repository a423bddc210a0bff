"""GPU grid sampling (adsg_sample, bspline::eval / eval_ders as the output manager uses them)
against the oracle's sampler, <= 1e-13 relative to the largest value."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,p,ne,a,b", [(2, 1, (33, 20), None, None), (2, 2, (64, 47), [0.0, -1.0], [2.0, 1.0]),
                                          (2, 3, (30, 41), None, None), (3, 1, (17, 12, 9), None, None),
                                          (3, 2, (20, 17, 19), [0.0, 1.0, -2.0], [1.0, 2.5, 2.0]),
                                          (3, 3, (13, 11, 12), None, None)])
def test_sampling_matches_oracle(dim, p, ne, a, b):
    print(f"case: sample dim={dim} p={p} ne={ne}", flush=True)
    g = ag.Simulation(dim, [p] * dim, list(ne), a=a, b=b)
    o = po.Sim(dim, [p] * dim, list(ne), a=(a + [0.0] * (3 - dim)) if a else None,
               b=(b + [1.0] * (3 - dim)) if b else None)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(g.size)
    g.set(ag.U, u)
    lo, hi = g.a, g.b
    pts = [np.sort(np.concatenate([rng.uniform(lo[d], hi[d], 40 - 10 * d), [lo[d], hi[d]]])) for d in range(dim)]
    for ders in (False, True):
        got = g.sample(pts, ders=ders).ravel()
        expect = o.sample(u, pts, ders=ders)
        assert np.max(np.abs(got - expect)) <= 1e-13 * np.max(np.abs(expect)), ders
    # the output manager's uniform grid
    n = [12, 7, 5][:dim]
    got = g.sample(intervals=n).ravel()
    expect = o.sample(u, [ag.linspace(lo[d], hi[d], n[d]) for d in range(dim)])
    assert np.max(np.abs(got - expect)) <= 1e-13 * np.max(np.abs(expect))
    with pytest.raises(ag.AdsgError):
        g.sample(pts, which=ag.SOURCE)  # not set
    g.close()
