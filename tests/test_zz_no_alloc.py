"""adsg_norm_noalloc (no device allocation; partial sums in a static device array) against
adsg_norm and adsg_error on the same handle, and against the oracle's norm functionals
(basic_simulation_2d.hpp:266-362, basic_simulation_3d.hpp:281-400).

Two handles are measured one after the other, the second smaller than the first, so the
second call reads a static array that still holds the first call's partial sums; a third
mesh has more elements than 1024 blocks of 256 lanes cover in one pass (the kernel strides).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "iga-ads_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import adsgpu as ag  # noqa: E402
import pyoracle as po  # noqa: E402

pytestmark = pytest.mark.gpu


class DeviceArray:
    """FP64 device memory owned by the test (the entry point under test allocates none): the
    state vector of a second handle, so the memory comes from the same HIP runtime as the
    handle under test."""

    def __init__(self, host):
        ne = int(np.ceil(np.sqrt(host.size))) + 1
        self.holder = ag.Simulation(2, [1, 1], [ne, ne])
        v = np.zeros(self.holder.size)
        v[: host.size] = host
        self.holder.set(ag.U, v)
        self.holder.synchronize()
        self.ptr = ag.lib().adsg_device_ptr(self.holder.h, ag.U)
        assert self.ptr

    def free(self):
        self.holder.close()


def quad_count(sim):
    return int(np.prod([sim.elements[d] * (sim.p[d] + 1) for d in range(sim.dim)]))


CASES = [
    # (dim, p, elements): first larger, then smaller; the last one needs the grid stride
    (3, 2, (24, 20, 18)),
    (2, 3, (10, 12)),
    (3, 1, (70, 70, 70)),
]


@pytest.mark.parametrize("order", [0, 1])
def test_norm_noalloc_matches_norm_and_error(order):
    rng = np.random.default_rng(7 + order)
    cases = CASES if order == 0 else CASES[::-1]
    for dim, p, ne in cases:
        print(f"norm_noalloc: dim {dim} p {p} elements {ne}", flush=True)
        g = ag.Simulation(dim, [p] * dim, list(ne))
        u = rng.standard_normal(g.size)
        g.set(ag.U, u)
        g.set(ag.U_PREV, 0.5 * u)
        for kind in (0, 1, 2):
            for which in (ag.U, ag.U_PREV):
                ref = g.norm(kind, which)
                got = g.norm_noalloc(kind, which)
                assert got == ref, (dim, p, ne, kind, which, got, ref)
        # the error functional against values given on the device
        Q = quad_count(g)
        for kind in (0, 2):
            planes = 1 if kind == 0 else 1 + dim
            sq = rng.standard_normal(planes * Q)
            dev = DeviceArray(sq)
            try:
                for which in (ag.U, None):
                    ref = g.error(sq, kind, which)
                    got = g.norm_noalloc(kind, which, dev.ptr)
                    assert got == ref, (dim, p, ne, kind, which, got, ref)
            finally:
                dev.free()
        # the oracle's norm functional (pinned by the 1D Gram / stiffness matrices)
        if int(np.prod(ne)) <= 20000:
            o = po.Sim(dim, [p] * dim, list(ne))
            for kind in (0, 2):
                ref = np.sqrt(o.norm_sq(u, kind))
                got = g.norm_noalloc(kind, ag.U)
                assert abs(got - ref) <= 1e-12 * ref, (kind, got, ref)
        g.close()


def test_norm_noalloc_rejects_bad_arguments():
    g = ag.Simulation(2, [2, 2], [8, 8])
    out = np.zeros(1)
    L = ag.lib()
    dp = C.POINTER(C.c_double)
    assert L.adsg_norm_noalloc(g.h, -1, 0, None, out.ctypes.data_as(dp)) == ag.EINVAL
    assert L.adsg_norm_noalloc(g.h, ag.U, 3, None, out.ctypes.data_as(dp)) == ag.EINVAL
    assert L.adsg_norm_noalloc(g.h, ag.BUFFER, 0, None, out.ctypes.data_as(dp)) == ag.EINVAL
    assert L.adsg_norm_noalloc(g.h, ag.U, 0, None, None) == ag.EINVAL
    g.close()
