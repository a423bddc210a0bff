"""The seeded workload generators of the package against the oracle's generator."""
import ctypes as C

import numpy as np

import pyoracle as po
from adsgpu import workloads as w


def test_splitmix_stream_matches_oracle_draws():
    st = C.c_ulonglong(2024)
    expect = [po.lib().orc_uniform(C.byref(st)) for _ in range(50)]
    s = w.Stream(2024)
    got = [s.uniform() for _ in range(20)] + list(s.uniform(30))
    assert got == expect


def test_knots_and_greville_match_oracle():
    for p, ne in ((1, 4), (2, 64), (3, 192), (2, 7)):
        kn = po.knots(0.0, 1.0, p, ne)
        assert np.array_equal(w.knots(p, ne), kn)
        g = np.zeros(len(kn) - p - 1)
        po.lib().orc_greville(po._d(kn), len(kn), p, po._d(g))
        assert np.array_equal(w.greville(p, ne), g)


def test_gaussians_match_oracle_generator():
    # same draws and points; exp() of numpy and of the C library may differ in the last bit
    for dim, p, ne, seed, count, sig, cen, noise in [
        (2, [2, 2], [64, 64], 7, 1, (0.1, 0.1), (0.3, 0.7), 1e-3),
        (3, [2, 2, 2], [20, 18, 16], 2024, 8, (0.05, 0.15), (0.2, 0.8), 1e-3),
        (3, [3, 3, 3], [12, 11, 10], 5, 3, (0.05, 0.15), (0.2, 0.8), 0.0),
    ]:
        o = po.Sim(dim, p, ne)
        a = o.init_gaussians(seed=seed, count=count, sigma=sig, centre=cen, noise=noise)
        b = w.gaussians(p, ne, seed, count, sig, cen, noise)
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(a).max())


def test_sine_source_is_separable_and_seeded():
    f1 = w.sine_source([3, 3, 3], [8, 9, 10], seed=77)
    f2 = w.sine_source([3, 3, 3], [8, 9, 10], seed=77)
    assert np.array_equal(f1, f2)
    assert not np.array_equal(f1, w.sine_source([3, 3, 3], [8, 9, 10], seed=78))
    # every term vanishes on the boundary (sin(k pi x) at x = 0, 1)
    f = f1.reshape(13, 12, 11)
    assert np.abs(f[0]).max() < 1e-14 and np.abs(f[:, :, -1]).max() < 1e-14
    assert np.abs(f).max() > 0.1
