"""GPU parity of batched 2D problems (grouped launches over many problems of uneven mesh
size and degree, one captured graph per step count) against the oracle per problem and
against the same problems run one by one.

Workload shape: the reaction-diffusion ensemble u_t = eps Lap u + u (1 - u) with natural
boundaries, explicit ADS steps (rhs = M u - dt eps K u + dt u (1 - u), then M^-1).
Bars: FP64 batch <= 1e-12 relative L2 per problem against the oracle; FP32 batch <= 1e-6
against the FP64 oracle; batch equals the single-problem path bitwise.
"""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po
from adsgpu import workloads as wl

pytestmark = pytest.mark.gpu

DT = 2e-5
PROBLEMS = [(1, 32), (2, 40), (3, 33), (2, 64), (1, 50), (3, 48), (2, 35), (1, 64), (3, 40),
            (2, 32), (1, 37), (3, 36)]
EPS = [1e-4, 3e-3, 1e-2, 5e-4, 2e-3, 1e-2, 7e-3, 1e-3, 4e-3, 2e-4, 6e-3, 8e-3]


def _log(msg):
    print(msg, flush=True)


def _forms():
    return [ag.rhs_form(DT, grad=(e, e, 0.0), reaction=1.0) for e in EPS]


def _u0(i):
    p, ne = PROBLEMS[i]
    return wl.smooth_unit(p, ne, seed=100 + i)


def _batch(precision, shape=0, sweep=0):
    b = ag.Batch(PROBLEMS, precision=precision)
    b.prepare()
    b.set_forms(_forms())
    b.set_tuning(shape)
    b.set_tuning(sweep, ag.TUNE_SWEEP_VARIANT)
    for i in range(len(PROBLEMS)):
        b.set(i, _u0(i))
    return b


def _oracle(i, steps):
    p, ne = PROBLEMS[i]
    o = po.Sim(2, [p, p], [ne, ne])
    assert o.prepare_matrices() == 0
    f = po.heat_form(DT, 2, grad=(EPS[i], EPS[i], 0.0), reaction=1.0)
    u = _u0(i)
    for _ in range(steps):
        u = o.solve(o.compute_rhs(u, f))
    return u


@pytest.mark.parametrize("shape,sweep", [(0, 0), (1, 0), (0, 1)])
def test_fp64_batch_matches_oracle_per_problem(shape, sweep):
    _log(f"case: FP64 batch of {len(PROBLEMS)} problems, tile shape {shape}, sweep {sweep}, 10 steps")
    b = _batch(ag.FP64, shape, sweep)
    b.step(10)
    worst = 0.0
    for i in range(len(PROBLEMS)):
        worst = max(worst, po.rel_l2(b.get(i), _oracle(i, 10)))
    _log(f"  worst rel L2 {worst:.3e}")
    assert worst <= 1e-12
    b.close()


def test_fp32_batch_against_fp64_oracle():
    _log("case: FP32 batch, 20 steps, against the FP64 oracle")
    b = _batch(ag.FP32)
    b.step(20)
    worst = 0.0
    for i in range(len(PROBLEMS)):
        worst = max(worst, po.rel_l2(b.get(i), _oracle(i, 20)))
    _log(f"  worst rel L2 {worst:.3e}")
    assert worst <= 1e-6
    b.close()


@pytest.mark.parametrize("precision", [ag.FP64, ag.FP32])
def test_batch_equals_single_problem_path_bitwise(precision):
    _log(f"case: batch vs one-by-one, precision {precision}")
    b = _batch(precision, 0, sweep=1)  # one lane per line, as the single path below
    b.step(6)
    for i in (0, 2, 5, 11):
        p, ne = PROBLEMS[i]
        g = ag.Simulation(2, [p, p], [ne, ne], precision=precision)
        g.prepare_matrices()
        g.set_tuning(ag.TUNE_SWEEP_VARIANT, 1)  # one lane per line, as the batch
        g.set_tuning(ag.TUNE_RHS_TILING, 0)  # colour launches, as the batch's element kernels
        g.set(ag.U, _u0(i))
        g.step(6, _forms()[i])
        assert np.array_equal(g.get(ag.U), b.get(i)), i
        g.close()
    b.close()


def test_batch_graph_replay_and_autotune_keep_results():
    _log("case: batch graph chunks and autotune")
    runs = []
    for chunks, tune in (([7], False), ([1, 2, 4], False), ([3, 4], True)):
        b = _batch(ag.FP32, 0)
        if tune:
            shape, ms = b.autotune()
            sweep = b.get_tuning(ag.TUNE_SWEEP_VARIANT)
            _log(f"  autotune: shape {shape}, sweep variant {sweep}, ms per step {ms}")
            assert shape in (0, 1) and sweep in (0, 1)
            b.set_tuning(0)
            b.set_tuning(0, ag.TUNE_SWEEP_VARIANT)
        for n in chunks:
            b.step(n)
        runs.append([b.get(i) for i in range(len(PROBLEMS))])
        b.close()
    for r in runs[1:]:
        for a, c in zip(runs[0], r):
            assert np.array_equal(a, c)


def test_device_setup_on_another_interval_equals_single_problem_path_bitwise():
    """The batch tabulates, assembles and factors its problems on the device; on [-1, 2]^2
    (knots, Jacobians and Gauss points other than the unit square's) its steps still equal
    the single-problem path, whose set-up runs on the host, bit for bit."""
    _log("case: batch device set-up on [-1, 2]^2 vs one-by-one")
    probs = [(1, 33), (2, 47), (3, 29), (3, 64)]
    forms = [ag.rhs_form(DT, grad=(3e-3, 3e-3, 0.0), reaction=1.0)] * len(probs)
    for precision in (ag.FP64, ag.FP32):
        b = ag.Batch(probs, precision=precision, interval=(-1.0, 2.0))
        b.prepare()
        b.set_forms(forms)
        b.set_tuning(1, ag.TUNE_SWEEP_VARIANT)
        u0 = [wl.smooth_unit(p, ne, seed=7 + i) for i, (p, ne) in enumerate(probs)]
        for i, u in enumerate(u0):
            b.set(i, u)
        b.step(4)
        for i, (p, ne) in enumerate(probs):
            g = ag.Simulation(2, [p, p], [ne, ne], precision=precision, a=[-1.0, -1.0], b=[2.0, 2.0])
            g.prepare_matrices()
            g.set_tuning(ag.TUNE_SWEEP_VARIANT, 1)
            g.set_tuning(ag.TUNE_RHS_TILING, 0)
            g.set(ag.U, u0[i])
            g.step(4, forms[i])
            assert np.array_equal(g.get(ag.U), b.get(i)), (precision, i)
            g.close()
        b.close()


def test_batch_create_refuses_configurations_outside_the_batched_kernels():
    """Batches take 2D problems with p + 1 Gauss points, simple knots, one degree and a < b;
    every other configuration is refused with ADSG_EINVAL and a message naming the problem."""
    import ctypes as C

    L = ag.lib()

    def create(edit):
        cfgs = (ag.Config * 2)()
        for c in cfgs:
            c.dim = 2
            for dc in (c.x, c.y):
                dc.p, dc.elements, dc.a, dc.b, dc.quad_order, dc.repeated_nodes = 2, 8, 0.0, 1.0, 0, 0
            c.precision = ag.FP64
        edit(cfgs[1])
        h = C.c_void_p()
        rc = L.adsg_batch_create(cfgs, 2, C.byref(h))
        if rc == ag.OK:
            L.adsg_batch_destroy(h)
        return rc, L.adsg_batch_last_error(None).decode()

    def setter(**kw):
        def edit(c):
            for k, v in kw.items():
                if k in ("dim", "precision"):
                    setattr(c, k, v)
                else:
                    setattr(c.y, k, v)
        return edit

    assert create(setter())[0] == ag.OK
    for kw in (dict(repeated_nodes=1), dict(quad_order=5), dict(p=3), dict(dim=3), dict(precision=ag.FP32),
               dict(elements=0), dict(a=1.0)):
        rc, msg = create(setter(**kw))
        assert rc == ag.EINVAL, kw
        assert "problem 1" in msg, (kw, msg)
