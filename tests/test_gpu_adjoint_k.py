"""GPU parity of the explicit-step adjoint with respect to a B-spline diffusion field k
(u_t = div(k grad u), explicit ADS step with Dirichlet boundary as in validation.hpp),
through the C ABI against the oracle's discrete adjoint, and the full gradient of
L = ||u(T) - u_target||^2 against central finite differences (h = 1e-6).

Bars: one adjoint step <= 1e-12 relative L2 (lam and dL/dk); the gradient through all
steps <= 1e-10 against the oracle's discrete adjoint; directional derivatives <= 1e-5
against central differences.
"""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po
from adsgpu import workloads as wl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _kfield(p, ne, seed):
    """k = exp(g), g a smooth seeded field scaled so that k lies in [0.5, 2]."""
    g = wl.sine_source([p] * 2, list(ne), seed=seed)
    g = g / max(np.abs(g).max(), 1e-300) * np.log(2.0)
    return np.exp(g)


def _pair(p, ne, dirichlet):
    g = ag.Simulation(2, [p, p], list(ne))
    o = po.Sim(2, [p, p], list(ne))
    if dirichlet:
        g.fix_all_boundaries()
        o.fix_all_boundaries()
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    return g, o


def _oracle_forward(o, u, k, dt, steps, dirichlet):
    traj = [u]
    form = po.heat_form(dt, 2, kfield=k)
    for _ in range(steps):
        r = o.compute_rhs(traj[-1], form)
        if dirichlet:
            o.zero_boundary(r)
        traj.append(o.solve(r))
    return traj


def _oracle_adjoint(o, traj, lam, k, dt, dirichlet):
    """dL/dk and dL/du0 by the discrete adjoint of _oracle_forward, given lam = dL/du(T)."""
    form = po.heat_form(dt, 2, kfield=k)
    dk = np.zeros_like(lam)
    for n in range(len(traj) - 2, -1, -1):
        w = lam.copy()
        o.solve(w, transposed=True)
        if dirichlet:
            o.zero_boundary(w)
        dk += o.kgrad(traj[n], w, form)
        lam = o.compute_rhs(w, form)
    return dk, lam


@pytest.mark.parametrize("dirichlet,sweep", [(True, 0), (False, 0), (True, 1), (True, 2), (False, 2)])
def test_adjoint_step_matches_oracle(dirichlet, sweep):
    """sweep: the line-sweep variant; the transposed sweeps are segmented in all of them (the
    first reads lambda and writes w), with x lines staged through LDS (0, 1) or read lane by
    lane (2)."""
    p, ne, dt = 2, (40, 36), 1e-5
    _log(f"case: explicit adjoint step 2D p=2 {ne} dirichlet={dirichlet} sweep variant {sweep}")
    g, o = _pair(p, ne, dirichlet)
    g.set_tuning(ag.TUNE_SWEEP_VARIANT, sweep)
    k = _kfield(p, ne, seed=31)
    rng = np.random.default_rng(5)
    un = rng.standard_normal(g.size)
    lam = rng.standard_normal(g.size)
    g.set(ag.KFIELD, k)
    g.set(ag.GRAD_K, np.zeros(g.size))
    g.set(ag.U, lam)
    g.set(ag.U_PREV, un)
    flags = ag.STEP_DIRICHLET if dirichlet else 0
    g.adjoint_step(ag.rhs_form(dt, kfield=True), flags)
    dk, lam0 = _oracle_adjoint(o, [un, None], lam, k, dt, dirichlet)
    e_lam = po.rel_l2(g.get(ag.U), lam0)
    e_dk = po.rel_l2(g.get(ag.GRAD_K), dk)
    _log(f"  lam {e_lam:.3e}  dk {e_dk:.3e}")
    assert e_lam <= 1e-12 and e_dk <= 1e-12


def _gpu_forward(g, k, u0, dt, steps, flags):
    g.set(ag.KFIELD, k)
    g.set(ag.U, u0)
    g.step(steps, ag.rhs_form(dt, kfield=True), flags)
    return g.get(ag.U)


def _gpu_gradient(g, k, u0, ut, dt, steps, every, flags):
    """dL/dk, dL/du0 with checkpoints every `every` steps (states recomputed per segment)."""
    form = ag.rhs_form(dt, kfield=True)
    g.set(ag.KFIELD, k)
    g.set(ag.U, u0)
    checkpoints = [u0.copy()]
    for s in range(steps):
        g.step(1, form, flags)
        if (s + 1) % every == 0 and s + 1 < steps:
            checkpoints.append(g.get(ag.U))
    uN = g.get(ag.U)
    lam = 2.0 * (uN - ut)
    g.set(ag.GRAD_K, np.zeros(g.size))
    for c in range(len(checkpoints) - 1, -1, -1):
        start = c * every
        stop = min(steps, start + every)
        seg = [checkpoints[c]]
        g.set(ag.U, checkpoints[c])
        for _ in range(start, stop - 1):
            g.step(1, form, flags)
            seg.append(g.get(ag.U))
        for n in range(stop - 1, start - 1, -1):
            g.set(ag.U, lam)
            g.set(ag.U_PREV, seg[n - start])
            g.adjoint_step(form, flags)
            lam = g.get(ag.U)
    return g.get(ag.GRAD_K), lam, uN


def test_gradient_through_steps_matches_oracle_and_finite_differences():
    p, ne, dt, steps, every = 2, (40, 36), 1e-5, 30, 7
    _log(f"case: dL/dk through {steps} steps, checkpoints every {every}")
    g, o = _pair(p, ne, True)
    flags = ag.STEP_DIRICHLET
    k = _kfield(p, ne, seed=41)
    kt = _kfield(p, ne, seed=42)
    u0 = wl.gaussians([p, p], list(ne), seed=43, count=2, sigma=(0.1, 0.15), noise=0.0)
    o.zero_boundary(u0)
    ut = _gpu_forward(g, kt, u0, dt, steps, flags)
    dk, du0, uN = _gpu_gradient(g, k, u0, ut, dt, steps, every, flags)
    traj = _oracle_forward(o, u0, k, dt, steps, True)
    assert po.rel_l2(uN, traj[-1]) <= 1e-12
    dk_o, du0_o = _oracle_adjoint(o, traj, 2.0 * (traj[-1] - ut), k, dt, True)
    _log(f"  dk vs oracle {po.rel_l2(dk, dk_o):.3e}  du0 vs oracle {po.rel_l2(du0, du0_o):.3e}")
    assert po.rel_l2(dk, dk_o) <= 1e-10 and po.rel_l2(du0, du0_o) <= 1e-10

    def loss(kk):
        v = _gpu_forward(g, kk, u0, dt, steps, flags)
        return float((v - ut) @ (v - ut))

    rng = np.random.default_rng(44)
    h = 1e-6
    for j in range(4):
        d = rng.standard_normal(g.size)
        d /= np.linalg.norm(d)
        fd = (loss(k + h * d) - loss(k - h * d)) / (2 * h)
        ad = float(dk @ d)
        rel = abs(fd - ad) / abs(ad)
        _log(f"  direction {j}: adjoint {ad:.9e} fd {fd:.9e} rel {rel:.2e}")
        assert rel <= 1e-5


def test_adjoint_step_rejects_bad_state():
    g, _ = _pair(2, (32, 32), False)
    g.set(ag.U_PREV, np.zeros(g.size))
    with pytest.raises(ag.AdsgError):  # GRAD_K not set
        g.adjoint_step(ag.rhs_form(1e-5, kfield=True))
    with pytest.raises(ag.AdsgError):  # nonlinear form
        g.adjoint_step(ag.rhs_form(1e-5, reaction=1.0))


@pytest.mark.parametrize("every", [7, 6])
def test_native_checkpointed_adjoint_equals_stepwise_adjoint(every):
    """adsg_forward_checkpointed + adsg_adjoint_checkpointed against the same adjoint driven
    step by step from the host (above): same kernels in the same order. The segments after
    the first are replayed from a graph captured once per starting state (every = 6: the
    recompute's 5 steps swap u and u_prev, so two graphs alternate); the adjoint is run three
    times, with eager work and synchronisations between, and every run must equal the first
    bitwise."""
    p, ne, dt, steps = 2, (40, 36), 1e-5, 30
    _log(f"case: native checkpointed adjoint, {steps} steps every {every}")
    g, o = _pair(p, ne, True)
    flags = ag.STEP_DIRICHLET
    k = _kfield(p, ne, seed=51)
    u0 = wl.gaussians([p, p], list(ne), seed=52, count=2, sigma=(0.1, 0.15), noise=0.0)
    o.zero_boundary(u0)
    ut = np.zeros(g.size)
    dk_ref, du0_ref, uN = _gpu_gradient(g, k, u0, ut, dt, steps, every, flags)
    form = ag.rhs_form(dt, kfield=True)
    g.set(ag.KFIELD, k)
    g.set(ag.U, u0)
    g.forward_checkpointed(steps, every, form, flags)
    assert np.array_equal(g.get(ag.U), uN)
    g.set(ag.GRAD_K, np.zeros(g.size))
    g.set(ag.U, 2.0 * (uN - ut))
    g.adjoint_checkpointed(form, flags)
    dk1, du01 = g.get(ag.GRAD_K), g.get(ag.U)
    assert po.rel_l2(dk1, dk_ref) <= 1e-13
    assert po.rel_l2(du01, du0_ref) <= 1e-13
    rng = np.random.default_rng(53)
    for _ in range(2):
        g.set(ag.U, rng.standard_normal(g.size))  # eager work between the replays
        g.set(ag.U_PREV, rng.standard_normal(g.size))
        g.adjoint_step(form, flags)
        g.synchronize()
        g.set(ag.GRAD_K, np.zeros(g.size))
        g.set(ag.U, 2.0 * (uN - ut))
        g.adjoint_checkpointed(form, flags)
        assert np.array_equal(g.get(ag.GRAD_K), dk1) and np.array_equal(g.get(ag.U), du01)
    with pytest.raises(ag.AdsgError):  # a different form than the forward run's
        g.adjoint_checkpointed(ag.rhs_form(2 * dt, kfield=True), flags)


def test_source_gradient_explicit_steps():
    """dL/df for a source field f (the rhs gains dt M f): accumulated into GRAD_F by the
    checkpointed adjoint, against the oracle's discrete adjoint (<= 1e-10) and central
    differences (L is quadratic in f: exact up to rounding for any h)."""
    p, ne, dt, steps, every = 2, (40, 36), 1e-5, 20, 6
    _log(f"case: source gradient, explicit, {steps} steps every {every}")
    g, o = _pair(p, ne, True)
    flags = ag.STEP_DIRICHLET
    k = _kfield(p, ne, seed=61)
    rng = np.random.default_rng(62)
    f = wl.sine_source([p, p], list(ne), seed=63) * 50.0
    u0 = wl.gaussians([p, p], list(ne), seed=64, count=2, sigma=(0.1, 0.15), noise=0.0)
    o.zero_boundary(u0)
    ut = rng.standard_normal(g.size) * 0.01
    form = ag.rhs_form(dt, kfield=True, source=True)

    def forward(ff):
        g.set(ag.KFIELD, k)
        g.set(ag.SOURCE, ff)
        g.set(ag.U, u0)
        g.step(steps, form, flags)
        return g.get(ag.U)

    g.set(ag.KFIELD, k)
    g.set(ag.SOURCE, f)
    g.set(ag.U, u0)
    g.forward_checkpointed(steps, every, form, flags)
    uT = g.get(ag.U)
    g.set(ag.GRAD_K, np.zeros(g.size))
    g.set(ag.GRAD_F, np.zeros(g.size))
    g.set(ag.U, 2.0 * (uT - ut))
    g.adjoint_checkpointed(form, flags)
    df = g.get(ag.GRAD_F)
    # oracle: forward with the source, adjoint with dL/df += dt M w
    fo = po.heat_form(dt, 2, kfield=k, source=f)
    traj = [u0]
    for _ in range(steps):
        r = o.compute_rhs(traj[-1], fo)
        o.zero_boundary(r)
        traj.append(o.solve(r))
    assert po.rel_l2(uT, traj[-1]) <= 1e-12
    lam = 2.0 * (traj[-1] - ut)
    df_o = np.zeros(g.size)
    lin = po.heat_form(dt, 2, kfield=k)
    mass = po.heat_form(0.0, 2)
    for n in range(steps - 1, -1, -1):
        w = lam.copy()
        o.solve(w, transposed=True)
        o.zero_boundary(w)
        df_o += dt * o.compute_rhs(w, mass)
        lam = o.compute_rhs(w, lin)
    _log(f"  df vs oracle {po.rel_l2(df, df_o):.3e}")
    assert po.rel_l2(df, df_o) <= 1e-10
    for j in range(3):
        d = rng.standard_normal(g.size)
        h = 1e-3
        lp = forward(f + h * d)
        lm = forward(f - h * d)
        fd = (float((lp - ut) @ (lp - ut)) - float((lm - ut) @ (lm - ut))) / (2 * h)
        ad = float(df @ d)
        _log(f"  direction {j}: adjoint {ad:.9e} fd {fd:.9e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_source_gradient_implicit_delta_steps():
    """The c5 step u += A^-1 dt (f - K u): dL/df of L = ||u(T)||^2 from adsg_adjoint_step_delta
    with GRAD_F set, against central differences; also through the 1D-matrix path."""
    p, ne, dt, steps = 2, (20, 18, 16), 1e-3, 5
    _log("case: source gradient, implicit delta steps 3D p=2")
    g = ag.Simulation(3, [p] * 3, list(ne))
    for d in range(3):
        g.set_matrix(d, 1, 1.0, 0.5 * dt)
    g.prepare_matrices()
    rng = np.random.default_rng(71)
    f = rng.standard_normal(g.size)
    u0 = rng.standard_normal(g.size) * 0.1
    form = ag.rhs_form(dt, mass=0.0, source=True)
    slots = [1, 1, 1]

    def forward(ff):
        g.set(ag.SOURCE, ff)
        g.set(ag.U, u0)
        g.step_delta(steps, form, slots)
        return g.get(ag.U)

    grads = []
    for kron in (0, 1):
        g.set_tuning(ag.TUNE_RHS_KRONECKER, kron)
        uT = forward(f)
        g.set(ag.GRAD_F, np.zeros(g.size))
        g.set(ag.U, 2.0 * uT)
        g.adjoint_step_delta(steps, form, slots)
        grads.append(g.get(ag.GRAD_F))
    assert po.rel_l2(grads[1], grads[0]) <= 1e-12
    g.set_tuning(ag.TUNE_RHS_KRONECKER, 0)
    for j in range(3):
        d = rng.standard_normal(g.size)
        h = 1e-2
        lp, lm = forward(f + h * d), forward(f - h * d)
        fd = (float(lp @ lp) - float(lm @ lm)) / (2 * h)
        ad = float(grads[0] @ d)
        _log(f"  direction {j}: adjoint {ad:.9e} fd {fd:.9e}")
        assert abs(fd - ad) <= 1e-8 * abs(ad)
    g.close()
