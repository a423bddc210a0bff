"""One-launch tiling of the 3D element kernels (ADSG_TUNE_RHS_TILING = 1): all tiles in one
launch, each adding the DOF-block entries it owns and spilling the rest, which a second
kernel adds in a fixed order. Checked against the oracle for every variant that has the
form (0, 1, 2, 4), every form without a k field or source in the kernel, in both
precisions, on meshes whose element counts are not multiples of the tile (partial last
tiles) and on meshes smaller than one tile in some direction; against the colour launches
of the same variant; and for determinism (repeated runs bitwise equal). Steps, which leave
the fix-up to their sweep phase, against the hook sequence that runs it at once."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _pair(p, ne, precision=ag.FP64, dirichlet=False):
    g = ag.Simulation(3, [p] * 3, list(ne), precision=precision)
    o = po.Sim(3, [p] * 3, list(ne))
    if dirichlet:
        g.fix_all_boundaries()
        o.fix_all_boundaries()
    assert g.prepare_matrices() == 0
    assert o.prepare_matrices() == 0
    return g, o


FORMS = [
    (dict(), dict()),
    (dict(grad=(0.5, 1.0, 2.0)), dict(grad=(0.5, 1.0, 2.0))),
    (dict(mass=0.0, grad=(1.0, 0.5, 2.0)), dict(mass=0.0, grad=(1.0, 0.5, 2.0))),
    (dict(reaction=1.0), dict(reaction=1.0)),
]


@pytest.mark.parametrize("variant", [0, 1, 2, 4])
@pytest.mark.parametrize("p,ne", [(2, (20, 17, 19)), (1, (8, 9, 7)), (3, (13, 11, 12)), (2, (5, 3, 21)),
                                  (3, (2, 9, 4)), (2, (16, 8, 16))])
def test_one_launch_rhs_matches_oracle_and_colours(p, ne, variant):
    _log(f"case: one-launch rhs p={p} ne={ne} variant={variant}")
    g, o = _pair(p, ne)
    g.set_tuning(ag.TUNE_RHS_VARIANT, variant)
    c = np.random.default_rng(31 + p).uniform(0, 1, g.size)
    g.set(ag.U_PREV, c)
    dt = 2e-3
    for gk, ok in FORMS:
        out = []
        for tiling in (0, 1, 1):
            g.set_tuning(ag.TUNE_RHS_TILING, tiling)
            g.compute_rhs(ag.rhs_form(dt, **gk))
            out.append(g.get(ag.U))
        expect = o.compute_rhs(c, po.heat_form(dt, 3, **ok))
        assert po.rel_l2(out[1], expect) < 1e-13, gk
        assert po.rel_l2(out[1], out[0]) < 1e-14, gk
        assert np.array_equal(out[1], out[2]), gk  # deterministic
    g.close()


@pytest.mark.parametrize("variant", [1, 4])
def test_one_launch_fp32_increment_step(variant):
    """FP32 increment steps with the one-launch tiling and fixed boundary rows track the FP64
    oracle like the colour launches do: <= 1e-6 relative L2 after every one of 5 steps."""
    p, ne, dt = 2, (20, 17, 19), 1e-5
    _log(f"case: one-launch fp32 variant={variant}")
    g, o = _pair(p, ne, ag.FP32, dirichlet=True)
    g.set_tuning(ag.TUNE_RHS_VARIANT, variant)
    g.set_tuning(ag.TUNE_RHS_TILING, 1)
    u = o.init_gaussians(seed=31, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    o.zero_boundary(u)
    g.set(ag.U, u)
    for _ in range(5):
        rhs = o.compute_rhs(u, po.heat_form(dt, 3), threads=8)
        o.zero_boundary(rhs)
        u = o.solve(rhs)
        g.step(1, ag.rhs_form(dt), ag.STEP_DIRICHLET)
        assert po.rel_l2(g.get(ag.U), u) <= 1e-6
    g.close()


def test_one_launch_graph_steps_equal_eager_and_autotune_keeps_tiling_valid():
    """Captured steps with the one-launch tiling equal the eager steps bit for bit, and the
    autotuner (which times both tilings) leaves a valid choice."""
    _log("case: one-launch graph vs eager, 3D p=2 24x20x18")
    a, o = _pair(2, (24, 20, 18))
    b, _ = _pair(2, (24, 20, 18))
    for h in (a, b):
        h.set_tuning(ag.TUNE_RHS_VARIANT, 4)
        h.set_tuning(ag.TUNE_RHS_TILING, 1)
    u = o.init_gaussians(seed=3, count=4, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    a.set(ag.U, u)
    b.set(ag.U, u)
    form = ag.rhs_form(1e-4)
    a.step(5, form)
    for _ in range(5):
        b.swap()
        b.compute_rhs(form)
        b.ads_solve()
    assert np.array_equal(a.get(ag.U), b.get(ag.U))
    a.autotune(form)
    assert a.get_tuning(ag.TUNE_RHS_TILING) in (0, 1)
    assert a.get_tuning(ag.TUNE_RHS_VARIANT) in range(6)
    a.close()
    b.close()


@pytest.mark.parametrize("dim,prec,dirichlet,tiling,sweep", [
    (3, ag.FP64, True, 1, 0), (3, ag.FP64, True, 0, 0), (3, ag.FP64, False, 1, 0), (3, ag.FP64, True, 1, 1),
    (2, ag.FP64, True, 1, 0), (2, ag.FP64, True, 0, 0), (2, ag.FP64, True, 1, 1),
    (3, ag.FP32, False, 1, 0), (3, ag.FP32, True, 1, 0), (2, ag.FP32, False, 1, 0),
])
def test_step_with_deferred_fixup_equals_the_hook_sequence(dim, prec, dirichlet, tiling, sweep):
    """adsg_step leaves the one-launch fix-up, and in FP64 the boundary zeroing, to the sweep
    phase (one pass before the x sweep; the zeroing alone inside the x sweep's line load with
    the colour launches; a separate pass before the one-lane-per-line sweep). The steps equal
    the reference's hook sequence (before_step swap, compute_rhs, boundary zeroing, ads_solve;
    validation.hpp:73-81) through the C ABI, where each part runs on its own, bit for bit."""
    ne = (21, 18, 19) if dim == 3 else (37, 40)
    _log(f"case: deferred fix-up {dim}D prec={prec} dirichlet={dirichlet} tiling={tiling} sweep={sweep}")
    sims = []
    for _ in range(2):
        h = ag.Simulation(dim, [2] * dim, list(ne), precision=prec)
        if dirichlet:
            h.fix_all_boundaries()
        assert h.prepare_matrices() == 0
        h.set_tuning(ag.TUNE_RHS_TILING, tiling)
        h.set_tuning(ag.TUNE_SWEEP_VARIANT, sweep)
        sims.append(h)
    a, b = sims
    u = np.random.default_rng(7 + dim).uniform(0, 1, a.size)
    a.set(ag.U, u)
    b.set(ag.U, u)
    form = ag.rhs_form(2e-4, grad=(1.0, 0.8, 1.2))
    flags = ag.STEP_DIRICHLET if dirichlet else 0
    a.step(3, form, flags)
    for _ in range(3):
        b.swap()
        b.compute_rhs(form)
        if dirichlet:
            b.zero_boundary()
        b.ads_solve()
    assert np.array_equal(a.get(ag.U), b.get(ag.U))
    a.close()
    b.close()


@pytest.mark.parametrize("p,ne", [(1, (33, 20)), (2, (20, 17)), (3, (40, 9)), (2, (5, 70)), (3, (16, 16))])
def test_one_launch_rhs_2d_every_form_matches_oracle_and_colours(p, ne):
    """2D one-launch tiling (ADSG_TUNE_RHS_TILING = 1; 16 x 16 tiles, up to 3 earlier tiles per
    DOF in the fix-up), every form including a k field and a source in the kernel: against the
    oracle (<= 1e-13) and the colour launches (<= 1e-14), and deterministic."""
    _log(f"case: one-launch rhs 2D p={p} ne={ne}")
    g = ag.Simulation(2, [p, p], list(ne))
    o = po.Sim(2, [p, p], list(ne))
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    rng = np.random.default_rng(17 + p)
    c = rng.uniform(0, 1, g.size)
    kf = np.exp(rng.uniform(-0.5, 0.5, g.size))
    src = rng.standard_normal(g.size)
    g.set(ag.U_PREV, c)
    g.set(ag.KFIELD, kf)
    g.set(ag.SOURCE, src)
    dt = 2e-3
    cases = [
        (dict(), dict()),
        (dict(mass=0.0, grad=(1.0, 0.5, 0.0)), dict(mass=0.0, grad=(1.0, 0.5, 0.0))),
        (dict(kfield=True), dict(kfield=kf)),
        (dict(reaction=1.0), dict(reaction=1.0)),
        (dict(source=True), dict(source=src)),
        (dict(kfield=True, source=True, reaction=0.5), dict(kfield=kf, source=src, reaction=0.5)),
    ]
    for gk, ok in cases:
        out = []
        for tiling in (0, 1, 1):
            g.set_tuning(ag.TUNE_RHS_TILING, tiling)
            g.compute_rhs(ag.rhs_form(dt, **gk))
            out.append(g.get(ag.U))
        expect = o.compute_rhs(c, po.heat_form(dt, 2, **ok))
        assert po.rel_l2(out[1], expect) < 1e-13, gk
        assert po.rel_l2(out[1], out[0]) < 1e-14, gk
        assert np.array_equal(out[1], out[2]), gk
    g.close()
