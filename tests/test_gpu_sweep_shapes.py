"""GPU parity of the segmented line sweeps (dgbtrs on every line, band_solve.hpp:20-31, inside
ads_solve, solver.hpp:17-41) at line lengths beyond what the x lines' LDS staging holds: up to
16 segments of 16 rows, then segments of 32 rows (lines of up to 512 DOFs), with the x lines
read lane by lane where the staging does not fit, and with sweep variant 2 (x lines always
read lane by lane). Against the oracle: the plain solve and the transposed solve <= 1e-13
relative L2 (FP64), explicit steps with fixed boundary rows <= 1e-12 (FP64) and <= 1e-6
(FP32 increment steps against the FP64 oracle)."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu

# (p, elements): 2D lines of n = elements + p DOFs in x
CASES = [(2, (254, 30)), (2, (300, 23)), (3, (253, 21)), (1, (480, 17)), (2, (128, 40))]


def _pair(p, ne, precision=ag.FP64, dirichlet=False):
    g = ag.Simulation(2, [p, p], list(ne), precision=precision)
    o = po.Sim(2, [p, p], list(ne))
    if dirichlet:
        g.fix_all_boundaries()
        o.fix_all_boundaries()
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    return g, o


@pytest.mark.parametrize("sweep", [0, 2])
@pytest.mark.parametrize("p,ne", CASES)
def test_solves_at_long_lines_match_oracle(p, ne, sweep):
    print(f"case: long-line solves p={p} ne={ne} sweep={sweep}", flush=True)
    g, o = _pair(p, ne)
    g.set_tuning(ag.TUNE_SWEEP_VARIANT, sweep)
    b = np.random.default_rng(11 + p).standard_normal(g.size)
    for transposed in (False, True):
        g.set(ag.U, b)
        if transposed:
            g.ads_solve_transposed()
        else:
            g.ads_solve()
        expect = o.solve(b.copy(), transposed=transposed)
        assert po.rel_l2(g.get(ag.U), expect) <= 1e-13, transposed
    g.close()


@pytest.mark.parametrize("sweep", [0, 2])
@pytest.mark.parametrize("p,ne", CASES[:3])
def test_steps_at_long_lines_match_oracle(p, ne, sweep):
    print(f"case: long-line steps p={p} ne={ne} sweep={sweep}", flush=True)
    dt = 1e-6
    for prec in (ag.FP64, ag.FP32):
        g, o = _pair(p, ne, prec, dirichlet=True)
        g.set_tuning(ag.TUNE_SWEEP_VARIANT, sweep)
        u = o.init_gaussians(seed=5, count=2, sigma=(0.05, 0.15), centre=(0.3, 0.7), noise=1e-3)
        o.zero_boundary(u)
        g.set(ag.U, u)
        for _ in range(3):
            rhs = o.compute_rhs(u, po.heat_form(dt, 2))
            o.zero_boundary(rhs)
            u = o.solve(rhs)
            g.step(1, ag.rhs_form(dt), ag.STEP_DIRICHLET)
            err = po.rel_l2(g.get(ag.U), u)
            assert err <= (1e-12 if prec == ag.FP64 else 1e-6), (prec, err)
        g.close()


@pytest.mark.parametrize("p,ne", [(2, (250, 18, 20)), (1, (400, 17, 17))])
def test_3d_solves_with_long_x_lines_match_oracle(p, ne):
    """3D: x lines of more than 224 DOFs (read lane by lane, the staging does not fit)."""
    print(f"case: 3D long x lines p={p} ne={ne}", flush=True)
    g = ag.Simulation(3, [p] * 3, list(ne))
    o = po.Sim(3, [p] * 3, list(ne))
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    b = np.random.default_rng(23 + p).standard_normal(g.size)
    for transposed in (False, True):
        g.set(ag.U, b)
        if transposed:
            g.ads_solve_transposed()
        else:
            g.ads_solve()
        assert po.rel_l2(g.get(ag.U), o.solve(b.copy(), transposed=transposed)) <= 1e-13, transposed
    g.close()


@pytest.mark.parametrize("prec,flags", [(ag.FP64, 0), (ag.FP32, 0), (ag.FP32, ag.STEP_REFINE)])
def test_delta_steps_at_long_lines_match_oracle(prec, flags):
    """Delta (implicit) steps u' = Z(u + A^-1 Z r(u)) with A = (M + dt/2 K) per direction on
    lines of 302 DOFs (segments of 32 rows), fixed boundary rows: the last sweep's increment
    update and, in FP32, the refined solves. FP64 <= 1e-12, FP32 <= 1e-6 against the oracle."""
    p, ne, dt = 2, (300, 21), 1e-4
    print(f"case: long-line delta steps prec={prec} flags={flags}", flush=True)
    g = ag.Simulation(2, [p, p], list(ne), precision=prec)
    o = po.Sim(2, [p, p], list(ne))
    g.fix_all_boundaries()
    o.fix_all_boundaries()
    for d in range(2):
        g.set_matrix(d, 1, 1.0, 0.5 * dt)
        o.set_matrix(d, 1, 1.0, 0.5 * dt)
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    u = o.init_gaussians(seed=9, count=2, sigma=(0.05, 0.15), centre=(0.3, 0.7), noise=1e-3)
    o.zero_boundary(u)
    g.set(ag.U, u)
    oform = po.heat_form(dt, 2)
    oform.mass = 0.0
    for _ in range(3):
        r = o.compute_rhs(u, oform)
        o.zero_boundary(r)
        o.solve(r, slots=[1, 1, 0])
        u = u + r
        o.zero_boundary(u)
        g.step_delta(1, ag.rhs_form(dt, mass=0.0), [1, 1], flags | ag.STEP_DIRICHLET)
        err = po.rel_l2(g.get(ag.U), u)
        assert err <= (1e-12 if prec == ag.FP64 else 1e-6), (prec, flags, err)
    g.close()
