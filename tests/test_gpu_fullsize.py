"""The benchmark workloads at their full stated sizes, through the C ABI, against the oracle
(the project's restatement of the reference CPU step) and against the FP64 GPU step.

* 3D explicit heat, p = 2, 128^3 elements (2,197,000 DOFs), dt = 1e-6, natural boundary
  (heat_3d.hpp:31-67 with the validation.hpp:95-119 element loop): FP64 against the
  threaded oracle after each of 3 steps (<= 1e-12); FP32 increment steps against FP64
  after each of 50 steps (<= 1e-6).
* 3D implicit heat with a source, p = 3, 192^3 elements (7,414,875 DOFs), dt = 1e-4: the
  all-direction delta form of the directional splitting of implicit.hpp:102-111 and
  elasticity/implicit.hpp:464-511 (whose 3-substep form is unstable at this size and dt):
  one FP64 step against the threaded oracle (<= 1e-12), 20 FP32 + refinement steps against
  FP64 after every step (<= 1e-6), and the adjoint gradient of ||u(T)||^2 with respect to
  u(0) against central differences (h = 1e-6) along 4 directions (<= 1e-5).
* 256 batched 2D reaction-diffusion problems (elements per side 32..256, p = 1..3): the
  reference runs one simulation per process (validation.hpp:73-81 loop); here all advance
  in one graph. FP64 after 10 steps against the oracle for the largest problem of every
  degree (<= 1e-12); FP32 against FP64 for all 256 after 200 steps in one graph launch
  (<= 1e-6); batched equals single-problem runs bitwise (one lane per line on both sides).
"""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po
from adsgpu import workloads as wl

pytestmark = pytest.mark.gpu

THREADS = 16


def _log(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------------- 3D p=2 128^3


def test_heat3d_p2_128_fp64_every_step_and_fp32_fifty_steps():
    ne, dt = 128, 1e-6
    _log("case: 3D p=2 128^3 FP64 3 steps vs oracle, FP32 50 steps vs FP64")
    u0 = wl.gaussians([2, 2, 2], [ne] * 3, seed=2024, count=8, sigma=(0.05, 0.15),
                      centre=(0.2, 0.8), noise=1e-3)
    g = ag.Simulation(3, [2, 2, 2], [ne] * 3)
    o = po.Sim(3, [2, 2, 2], [ne] * 3)
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    form = ag.rhs_form(dt)
    oform = po.heat_form(dt, 3)
    g.set(ag.U, u0)
    y = u0
    for k in range(3):
        g.step(1, form)
        y = o.solve(o.compute_rhs(y, oform, threads=THREADS))
        err = po.rel_l2(g.get(ag.U), y)
        _log(f"  FP64 step {k + 1}: rel L2 vs oracle {err:.3e}")
        assert err <= 1e-12
    # FP32 increment steps against the FP64 GPU step, after every step
    g.set(ag.U, u0)
    s = ag.Simulation(3, [2, 2, 2], [ne] * 3, precision=ag.FP32)
    s.prepare_matrices()
    s.set(ag.U, u0)
    worst = 0.0
    for k in range(50):
        g.step(1, form)
        s.step(1, form)
        worst = max(worst, po.rel_l2(s.get(ag.U), g.get(ag.U)))
    _log(f"  FP32 worst rel L2 vs FP64 over 50 steps {worst:.3e}")
    assert worst <= 1e-6
    s.close()
    g.close()


# ---------------------------------------------------------------------- 3D p=3 192^3 implicit

NE5, DT5, THETA = 192, 1e-4, 0.5


def _implicit_sim(precision=ag.FP64):
    g = ag.Simulation(3, [3, 3, 3], [NE5] * 3, precision=precision)
    for d in range(3):
        g.set_matrix(d, 1, 1.0, THETA * DT5)
    assert g.prepare_matrices() == 0
    return g


def _c5_inputs():
    p3 = [3, 3, 3]
    u0 = wl.gaussians(p3, [NE5] * 3, seed=2025, count=8, sigma=(0.05, 0.15), centre=(0.2, 0.8),
                      noise=1e-3)
    f = wl.sine_source(p3, [NE5] * 3, seed=2026)
    return u0, f


def test_implicit3d_p3_192_fp64_step_vs_oracle_and_fp32_refined_every_step():
    _log("case: 3D p=3 192^3 implicit FP64 step vs oracle, FP32+refinement 20 steps vs FP64")
    u0, f = _c5_inputs()
    form = ag.rhs_form(DT5, mass=0.0, source=True)
    slots = [1, 1, 1]
    g = _implicit_sim()
    g.set(ag.U, u0)
    g.set(ag.SOURCE, f)
    g.step_delta(1, form, slots)
    x1 = g.get(ag.U)
    o = po.Sim(3, [3, 3, 3], [NE5] * 3)
    for d in range(3):
        o.set_matrix(d, 1, 1.0, THETA * DT5)
    assert o.prepare_matrices() == 0
    oform = po.heat_form(DT5, 3, source=f)
    oform.mass = 0.0
    r = o.compute_rhs(u0, oform, threads=THREADS)
    o.solve(r, slots=slots)
    err = po.rel_l2(x1, u0 + r)
    _log(f"  FP64 step: rel L2 vs oracle {err:.3e}")
    assert err <= 1e-12
    del o, r
    # FP32 + refinement against FP64, after every step
    s = _implicit_sim(ag.FP32)
    s.set(ag.U, u0)
    s.set(ag.SOURCE, f)
    g.set(ag.U, u0)
    worst = 0.0
    for k in range(20):
        g.step_delta(1, form, slots)
        s.step_delta(1, form, slots, ag.STEP_REFINE)
        worst = max(worst, po.rel_l2(s.get(ag.U), g.get(ag.U)))
    _log(f"  FP32+refinement worst rel L2 vs FP64 over 20 steps {worst:.3e}")
    assert worst <= 1e-6
    s.close()
    g.close()


def test_implicit3d_p3_192_adjoint_of_final_norm_against_central_differences():
    steps = 20
    _log(f"case: 3D p=3 192^3 adjoint of ||u(T)||^2 wrt u(0), {steps} steps, 4 FD directions")
    u0, f = _c5_inputs()
    form = ag.rhs_form(DT5, mass=0.0, source=True)
    slots = [1, 1, 1]
    g = _implicit_sim()
    g.set(ag.SOURCE, f)

    def loss(u):
        g.set(ag.U, u)
        g.step_delta(steps, form, slots)
        v = g.get(ag.U)
        return float(v @ v)

    g.set(ag.U, u0)
    g.step_delta(steps, form, slots)
    uT = g.get(ag.U)
    g.set(ag.U, 2.0 * uT)
    g.adjoint_step_delta(steps, form, slots)
    grad = g.get(ag.U)
    rng = np.random.default_rng(2034)
    h = 1e-6
    for k in range(4):
        # entries of unit variance: the perturbation h d has norm h sqrt(N), so the central
        # difference stands well above the rounding of L (about 1e6)
        d = rng.standard_normal(g.size)
        fd = (loss(u0 + h * d) - loss(u0 - h * d)) / (2 * h)
        ad = float(grad @ d)
        rel = abs(fd - ad) / abs(ad)
        _log(f"  direction {k}: adjoint {ad:.9e} central difference {fd:.9e} rel {rel:.2e}")
        assert rel <= 1e-5
    g.close()


# ---------------------------------------------------------------------- 256-problem ensemble

DT3 = 2e-5


def _ensemble():
    probs = wl.ensemble(seed=2028, count=256)
    shape = [(q["p"], q["ne"]) for q in probs]
    forms = [ag.rhs_form(DT3, grad=(q["eps"], q["eps"], 0.0), reaction=1.0) for q in probs]
    u0 = [wl.smooth_unit(q["p"], q["ne"], q["seed"]) for q in probs]
    return probs, shape, forms, u0


def _batch(shape, forms, u0, precision, sweep=None):
    b = ag.Batch(shape, precision=precision)
    b.prepare()
    b.set_forms(forms)
    b.set_tuning(0)  # 16 x 16 element tiles, as the single-problem kernel
    if sweep is not None:
        b.set_tuning(sweep, ag.TUNE_SWEEP_VARIANT)
    for i, u in enumerate(u0):
        b.set(i, u)
    return b


def test_ensemble256_fp64_largest_per_degree_vs_oracle():
    probs, shape, forms, u0 = _ensemble()
    steps = 10
    _log(f"case: 256-problem ensemble FP64, {steps} steps in one graph, largest per degree vs oracle")
    b = _batch(shape, forms, u0, ag.FP64)
    b.step(steps)
    for p in (1, 2, 3):
        i = max((j for j, q in enumerate(probs) if q["p"] == p), key=lambda j: probs[j]["ne"])
        q = probs[i]
        o = po.Sim(2, [p, p], [q["ne"], q["ne"]])
        assert o.prepare_matrices() == 0
        of = po.heat_form(DT3, 2, grad=(q["eps"], q["eps"], 0.0), reaction=1.0)
        y = u0[i]
        for _ in range(steps):
            y = o.solve(o.compute_rhs(y, of, threads=THREADS))
        err = po.rel_l2(b.get(i), y)
        _log(f"  problem {i} (p={p}, {q['ne']}^2 elements): rel L2 vs oracle {err:.3e}")
        assert err <= 1e-12
    b.close()


def test_ensemble256_fp32_vs_fp64_200_steps_one_graph_and_batch_equals_single():
    probs, shape, forms, u0 = _ensemble()
    steps = 200
    _log(f"case: 256-problem ensemble, {steps} steps in one graph, FP32 vs FP64")
    out = {}
    for name, prec in (("fp64", ag.FP64), ("fp32", ag.FP32)):
        b = _batch(shape, forms, u0, prec)
        b.autotune()
        for i, u in enumerate(u0):
            b.set(i, u)
        b.step(steps)
        out[name] = [b.get(i) for i in range(len(probs))]
        b.close()
    worst = max(po.rel_l2(a, c) for a, c in zip(out["fp32"], out["fp64"]))
    _log(f"  FP32 worst rel L2 vs FP64 over 256 problems {worst:.3e}")
    assert worst <= 1e-6
    # one lane per line and colour launches on both sides: the batched problem equals the
    # problem run alone
    b = _batch(shape, forms, u0, ag.FP64, sweep=1)
    b.step(steps)
    picks = sorted({max(range(len(probs)), key=lambda j: (probs[j]["p"] == p, probs[j]["ne"]))
                    for p in (1, 2, 3)} | {0, 17})
    for i in picks:
        q = probs[i]
        g = ag.Simulation(2, [q["p"]] * 2, [q["ne"]] * 2)
        g.prepare_matrices()
        g.set_tuning(ag.TUNE_SWEEP_VARIANT, 1)
        g.set_tuning(ag.TUNE_RHS_TILING, 0)
        g.set(ag.U, u0[i])
        g.step(steps, forms[i])
        assert np.array_equal(g.get(ag.U), b.get(i)), i
        g.close()
    b.close()
