"""GPU parity of the linearly implicit (delta-form) step, the transposed solve and the
adjoint step, through the C ABI against the CPU oracle on the same inputs.

The implicit heat step with source is u' = u + A^-1 r(u), r(u) = dt (f - K u) and
A = (M_x + theta dt K_x) (x) (M_y + theta dt K_y) (x) (M_z + theta dt K_z), theta = 1/2:
the all-direction form of the directional splitting of implicit.hpp:102-111 (whose 3D
three-substep version is not stable at dt = 1e-4 on a 192^3 p = 3 mesh). Bars: FP64
<= 1e-12 relative L2 after every step; FP32 with one FP64-residual refinement per line
solve <= 1e-6 against FP64; adjoint against the oracle's discrete adjoint <= 1e-12 and
against central finite differences (h = 1e-6) <= 1e-5.
"""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po
from adsgpu import workloads as wl

pytestmark = pytest.mark.gpu

FP64_STEP_TOL = 1e-12
FP32_TOL = 1e-6
THETA = 0.5


def _log(msg):
    print(msg, flush=True)


def _implicit_pair(dim, p, ne, dt, precision=ag.FP64, fix=None):
    g = ag.Simulation(dim, [p] * dim, list(ne), precision=precision)
    o = po.Sim(dim, [p] * dim, list(ne))
    for d, k in fix or []:
        g.fix_dof(d, k)
        o.fix_dof(d, k)
    for d in range(dim):
        g.set_matrix(d, 1, 1.0, THETA * dt)
        o.set_matrix(d, 1, 1.0, THETA * dt)
    assert g.prepare_matrices() == 0
    assert o.prepare_matrices() == 0
    return g, o


def _oracle_form(dt, dim, source=None):
    f = po.heat_form(dt, dim, source=source)
    f.mass = 0.0
    return f


def _oracle_delta_step(o, u, form, slots, dirichlet=False):
    r = o.compute_rhs(u, form, threads=8)
    if dirichlet:
        o.zero_boundary(r)
    o.solve(r, slots=slots)
    un = u + r
    if dirichlet:
        o.zero_boundary(un)
    return un


def _oracle_adjoint_step(o, lam, form_lin, slots, dirichlet=False):
    mu = lam.copy()
    if dirichlet:
        o.zero_boundary(mu)
    v = mu.copy()
    o.solve(v, slots=slots, transposed=True)
    if dirichlet:
        o.zero_boundary(v)
    return mu + o.compute_rhs(v, form_lin, threads=8)


# ------------------------------------------------------------------ transposed solve
@pytest.mark.parametrize("dim,p,ne", [(2, 2, (40, 36)), (3, 3, (40, 36, 34)), (3, 2, (30, 33, 31))])
def test_transposed_solve_matches_oracle(dim, p, ne):
    _log(f"case: transposed solve dim={dim} p={p} ne={ne}")
    # a fixed dof per direction makes slot 0 non-symmetric, so A^-T differs from A^-1
    fix = [(d, 0 if d % 2 == 0 else ne[d] + p - 1) for d in range(dim)]
    g, o = _implicit_pair(dim, p, ne, 1e-3, fix=fix)
    rng = np.random.default_rng(17 + dim)
    for slots in ([0] * dim, [1] * dim, [1, 0, 1][:dim]):
        b = rng.standard_normal(g.size)
        g.set(ag.U, b)
        g.ads_solve_transposed(slots)
        x = g.get(ag.U)
        y = o.solve(b.copy(), slots=slots + [0] * (3 - dim), transposed=True)
        assert po.rel_l2(x, y) <= FP64_STEP_TOL, slots


# ------------------------------------------------------------------ implicit delta step
@pytest.mark.parametrize("dirichlet", [False, True])
def test_delta_step_fp64_every_step(dirichlet):
    dim, p, ne, dt = 3, 3, (40, 36, 34), 1e-4
    _log(f"case: implicit delta step FP64 p=3 {ne} dirichlet={dirichlet}")
    g, o = _implicit_pair(dim, p, ne, dt)
    f = wl.sine_source([p] * dim, list(ne), seed=505)
    u = wl.gaussians([p] * dim, list(ne), seed=77, count=4, sigma=(0.05, 0.15),
                     centre=(0.2, 0.8), noise=1e-3)
    g.set(ag.U, u)
    g.set(ag.SOURCE, f)
    form = ag.rhs_form(dt, mass=0.0, source=True)
    oform = _oracle_form(dt, dim, source=f)
    flags = ag.STEP_DIRICHLET if dirichlet else 0
    for step in range(4):
        g.step_delta(1, form, [1, 1, 1], flags)
        u = _oracle_delta_step(o, u, oform, [1, 1, 1], dirichlet)
        err = po.rel_l2(g.get(ag.U), u)
        _log(f"  step {step}: rel L2 {err:.3e}")
        assert err <= FP64_STEP_TOL


def test_delta_step_graph_replay_equals_single_steps_bitwise():
    dim, p, ne, dt = 3, 3, (20, 18, 17), 1e-4
    _log("case: delta step graph replay vs single steps")
    runs = []
    for chunks in ([5], [1, 1, 1, 1, 1], [2, 3]):
        g, _ = _implicit_pair(dim, p, ne, dt)
        g.set(ag.U, wl.gaussians([p] * dim, list(ne), seed=3, count=2, noise=0.0))
        g.set(ag.SOURCE, wl.sine_source([p] * dim, list(ne), seed=4))
        for n in chunks:
            g.step_delta(n, ag.rhs_form(dt, mass=0.0, source=True), [1, 1, 1])
        runs.append(g.get(ag.U))
        g.close()
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])


def test_delta_step_fp32_refined_against_fp64():
    dim, p, ne, dt, steps = 3, 3, (40, 36, 34), 1e-4, 6
    _log(f"case: implicit delta step FP32+refinement p=3 {ne}")
    f = wl.sine_source([p] * dim, list(ne), seed=505)
    u0 = wl.gaussians([p] * dim, list(ne), seed=77, count=4, sigma=(0.05, 0.15),
                      centre=(0.2, 0.8), noise=1e-3)
    form = ag.rhs_form(dt, mass=0.0, source=True)
    out = {}
    for name, prec, flags in (("fp64", ag.FP64, 0), ("fp32", ag.FP32, 0),
                              ("fp32r", ag.FP32, ag.STEP_REFINE)):
        g, _ = _implicit_pair(dim, p, ne, dt, precision=prec)
        g.set(ag.U, u0)
        g.set(ag.SOURCE, f)
        traj = []
        for _ in range(steps):
            g.step_delta(1, form, [1, 1, 1], flags)
            traj.append(g.get(ag.U))
        out[name] = traj
        g.close()
    for s in range(steps):
        e32 = po.rel_l2(out["fp32"][s], out["fp64"][s])
        e32r = po.rel_l2(out["fp32r"][s], out["fp64"][s])
        _log(f"  step {s}: fp32 {e32:.3e}  fp32+refine {e32r:.3e}")
        assert e32r <= FP32_TOL


def test_refinement_repairs_the_fp32_line_solves():
    """One FP64-residual refinement per line solve makes the FP32 solve of the FP32-assembled
    increment accurate to FP32 storage: compared with the exact (FP64) solve of that same
    FP32 right-hand side, the refined increment must be at least twice as close as the
    unrefined one. (Against the FP64 step both stay at the FP32 assembly error, ~1e-7.)"""
    dim, p, ne, dt = 3, 3, (40, 36, 34), 1e-2
    _log(f"case: refined FP32 line solves, p=3 {ne}, dt={dt}")
    f = wl.sine_source([p] * dim, list(ne), seed=606)
    form = ag.rhs_form(dt, mass=0.0, source=True)

    def handle(prec):
        g, _ = _implicit_pair(dim, p, ne, dt, precision=prec)
        g.set(ag.U, np.zeros(g.size))
        g.set(ag.SOURCE, f)
        return g

    g = handle(ag.FP32)
    g.compute_rhs(form)  # FP32 increment rhs from u = 0: dt (f, v)
    r32 = g.get(ag.BUFFER)
    g.close()
    g64 = handle(ag.FP64)
    g64.set(ag.U, r32)
    g64.ads_solve([1, 1, 1])
    exact = g64.get(ag.U)
    g64.close()
    err = {}
    for name, flags in (("fp32", 0), ("fp32r", ag.STEP_REFINE)):
        g = handle(ag.FP32)
        g.step_delta(1, form, [1, 1, 1], flags)
        err[name] = po.rel_l2(g.get(ag.U), exact)
        g.close()
    _log(f"  against the exact solve of the FP32 rhs: fp32 {err['fp32']:.3e}  "
         f"fp32+refine {err['fp32r']:.3e}")
    assert err["fp32r"] <= 1e-7
    assert err["fp32r"] < 0.5 * err["fp32"]


def test_refine_flag_rejected_for_fp64():
    g, _ = _implicit_pair(2, 2, (32, 32), 1e-4)
    with pytest.raises(ag.AdsgError):
        g.step_delta(1, ag.rhs_form(1e-4, mass=0.0), [1, 1], ag.STEP_REFINE)


# ------------------------------------------------------------------ adjoint
@pytest.mark.parametrize("dirichlet", [False, True])
def test_adjoint_delta_matches_oracle_discrete_adjoint(dirichlet):
    dim, p, ne, dt, steps = 3, 3, (40, 36, 34), 1e-4, 3
    _log(f"case: adjoint delta step p=3 {ne} dirichlet={dirichlet}")
    g, o = _implicit_pair(dim, p, ne, dt)
    lam = np.random.default_rng(8).standard_normal(g.size)
    g.set(ag.U, lam)
    flags = ag.STEP_DIRICHLET if dirichlet else 0
    form = ag.rhs_form(dt, mass=0.0, source=True)  # the source is ignored by the adjoint
    g.set(ag.SOURCE, np.ones(g.size))
    oform = _oracle_form(dt, dim)
    for step in range(steps):
        g.adjoint_step_delta(1, form, [1, 1, 1], flags)
        lam = _oracle_adjoint_step(o, lam, oform, [1, 1, 1], dirichlet)
        err = po.rel_l2(g.get(ag.U), lam)
        _log(f"  step {step}: rel L2 {err:.3e}")
        assert err <= FP64_STEP_TOL


def test_adjoint_gradient_of_final_norm_against_finite_differences():
    """L(u0) = sum_i u_N,i^2 after N delta steps; dL/du0 from the adjoint steps against
    central differences with h = 1e-6 along seeded directions."""
    dim, p, ne, dt, steps = 3, 3, (24, 22, 20), 1e-4, 5
    _log(f"case: adjoint vs FD p=3 {ne} {steps} steps")
    f = wl.sine_source([p] * dim, list(ne), seed=9)
    u0 = wl.gaussians([p] * dim, list(ne), seed=10, count=3, noise=1e-3)
    form = ag.rhs_form(dt, mass=0.0, source=True)
    g, _ = _implicit_pair(dim, p, ne, dt)
    g.set(ag.SOURCE, f)

    def loss(u):
        g.set(ag.U, u)
        g.step_delta(steps, form, [1, 1, 1])
        uN = g.get(ag.U)
        return float(uN @ uN), uN

    L0, uN = loss(u0)
    g.set(ag.U, 2.0 * uN)
    g.adjoint_step_delta(steps, form, [1, 1, 1])
    grad = g.get(ag.U)
    rng = np.random.default_rng(12)
    h = 1e-6
    for k in range(4):
        d = rng.standard_normal(g.size)
        d /= np.linalg.norm(d)
        fd = (loss(u0 + h * d)[0] - loss(u0 - h * d)[0]) / (2 * h)
        ad = float(grad @ d)
        rel = abs(fd - ad) / max(abs(ad), 1e-30)
        _log(f"  direction {k}: adjoint {ad:.9e} fd {fd:.9e} rel {rel:.2e}")
        assert rel <= 1e-5
