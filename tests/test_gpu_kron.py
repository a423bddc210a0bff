"""Separable-integrand rhs (ADSG_TUNE_RHS_KRONECKER): forms without k field or reaction
assembled from the 1D Gram and stiffness matrices, against the element kernels and the
oracle's element loop (FP64 <= 1e-13 per rhs, <= 1e-12 per step; FP32 <= 1e-6)."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu

SHAPES = [(2, (20, 17, 19)), (1, (40, 9, 33)), (3, (13, 11, 12)), (2, (64, 20, 35))]


def _pair(p, ne, precision=ag.FP64):
    g = ag.Simulation(3, [p] * 3, list(ne), precision=precision)
    o = po.Sim(3, [p] * 3, list(ne))
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    return g, o


@pytest.mark.parametrize("p,ne", SHAPES)
def test_kronecker_rhs_equals_element_loop(p, ne):
    print(f"case: kron rhs p={p} ne={ne}", flush=True)
    g, o = _pair(p, ne)
    rng = np.random.default_rng(13)
    c = rng.standard_normal(g.size)
    f = rng.standard_normal(g.size)
    g.set(ag.U_PREV, c)
    g.set(ag.SOURCE, f)
    cases = [(dict(), dict()), (dict(grad=(0.0, 1.0, 0.0)), dict(grad=(0.0, 1.0, 0.0))),
             (dict(mass=0.5, grad=(2.0, 1.0, 0.5)), dict(mass=0.5, grad=(2.0, 1.0, 0.5))),
             (dict(mass=0.0, source=True), dict(mass=0.0, source=f))]
    for gk, ok in cases:
        form = ag.rhs_form(2e-3, **gk)
        out = []
        for kron in (0, 1):
            g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
            g.compute_rhs(form)
            out.append(g.get(ag.U))
        expect = o.compute_rhs(c, po.heat_form(2e-3, 3, **ok), threads=8)
        assert po.rel_l2(out[1], out[0]) <= 1e-13, gk
        assert po.rel_l2(out[1], expect) <= 1e-13, gk
    # forms with a k field or a reaction term keep the element kernels (bitwise the same)
    g.set(ag.KFIELD, np.exp(rng.uniform(-0.5, 0.5, g.size)))
    for gk in (dict(kfield=True), dict(reaction=1.0)):
        out = []
        for kron in (0, 1):
            g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
            g.compute_rhs(ag.rhs_form(2e-3, **gk))
            out.append(g.get(ag.U))
        assert np.array_equal(out[0], out[1]), gk
    g.close()


@pytest.mark.parametrize("precision", [ag.FP64, ag.FP32])
def test_kronecker_steps_match_oracle(precision):
    """dt K_d / M_d ~ 1e-2 (the explicit configurations' regime): in FP32 the increment
    -dt K u is assembled from the FP32 shadow of u either way, and both paths stay within
    1e-6 of the FP64 oracle."""
    p, ne = 2, (40, 36, 34)
    print(f"case: kron steps precision {precision}", flush=True)
    g = ag.Simulation(3, [p] * 3, list(ne), precision=precision)
    o = po.Sim(3, [p] * 3, list(ne))
    g.fix_all_boundaries()
    o.fix_all_boundaries()
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    u = o.init_gaussians(seed=7, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    o.zero_boundary(u)
    u0 = u.copy()
    dt = 1e-5
    for _ in range(4):
        rhs = o.compute_rhs(u, po.heat_form(dt, 3), threads=8)
        o.zero_boundary(rhs)
        u = o.solve(rhs)
    errs = []
    for kron in (0, 1):
        g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
        g.set(ag.U, u0)
        g.step(4, ag.rhs_form(dt), ag.STEP_DIRICHLET)
        errs.append(po.rel_l2(g.get(ag.U), u))
    print(f"  element {errs[0]:.3e}  kronecker {errs[1]:.3e}", flush=True)
    assert max(errs) <= (1e-12 if precision == ag.FP64 else 1e-6)
    g.close()


def test_kronecker_implicit_delta_step_with_source():
    """The c5 form: u += A^-1 dt (f - K u) with A = (x)_d (M_d + dt/2 K_d), p = 3."""
    p, ne, dt = 3, (24, 20, 18), 1e-3
    print("case: kron implicit delta step p=3", flush=True)
    g, o = _pair(p, ne)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(g.size)
    u0 = rng.standard_normal(g.size)
    out = []
    for kron in (0, 1):
        for d in range(3):
            g.set_matrix(d, 1, 1.0, 0.5 * dt)
        g.prepare_matrices()
        g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
        g.set(ag.U, u0)
        g.set(ag.SOURCE, f)
        g.step_delta(3, ag.rhs_form(dt, mass=0.0, source=True), [1, 1, 1])
        out.append(g.get(ag.U))
    assert po.rel_l2(out[1], out[0]) <= 1e-12
    g.close()


def test_kronecker_implicit_fp32_refined_within_1e6_of_fp64():
    """The c5 regime at 96^3 (dt = 1e-4, p = 3, smooth state and source, 10 steps): FP32 +
    refinement through the separable rhs and through the element kernels both stay within
    1e-6 of the FP64 element-loop steps. The stiffness rows sum to zero, so an rhs contracted
    in FP32 loses digits to cancellation (4e-6 after 22 steps at 192^3); the separable path
    contracts in FP64 from the FP32 shadow of u."""
    from adsgpu import workloads as wl

    p, ne, dt, steps = 3, [96] * 3, 1e-4, 10
    print("case: kron implicit fp32 refined 96^3", flush=True)
    u0 = wl.gaussians([p] * 3, ne, seed=2025, count=8, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    f = wl.sine_source([p] * 3, ne, seed=2026)
    form = ag.rhs_form(dt, mass=0.0, source=True)
    out = {}
    for name, prec, kron, flags in (("fp64", ag.FP64, 0, 0), ("kron32", ag.FP32, 1, ag.STEP_REFINE),
                                    ("elem32", ag.FP32, 0, ag.STEP_REFINE)):
        g = ag.Simulation(3, [p] * 3, ne, precision=prec)
        for d in range(3):
            g.set_matrix(d, 1, 1.0, 0.5 * dt)
        g.prepare_matrices()
        g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
        g.set(ag.U, u0)
        g.set(ag.SOURCE, f)
        g.step_delta(steps, form, [1, 1, 1], flags)
        out[name] = g.get(ag.U)
        g.close()
    errs = {k: po.rel_l2(out[k], out["fp64"]) for k in ("kron32", "elem32")}
    print(f"  {errs}", flush=True)
    assert max(errs.values()) <= 1e-6


def test_kronecker_adjoint_delta_step_equals_element_adjoint():
    """adsg_adjoint_step_delta applies R^T = R (the linear form without source) to the
    transposed solve's result: the 1D-matrix form gives the element kernels' result."""
    p, ne, dt = 3, (24, 20, 18), 1e-3
    print("case: kron adjoint delta step p=3", flush=True)
    g, _ = _pair(p, ne)
    for d in range(3):
        g.set_matrix(d, 1, 1.0, 0.5 * dt)
    g.prepare_matrices()
    lam0 = np.random.default_rng(6).standard_normal(g.size)
    out = []
    for kron in (0, 1):
        g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
        g.set(ag.U, lam0)
        g.adjoint_step_delta(3, ag.rhs_form(dt, mass=0.0, source=True), [1, 1, 1])
        out.append(g.get(ag.U))
    assert po.rel_l2(out[1], out[0]) <= 1e-12
    g.close()
