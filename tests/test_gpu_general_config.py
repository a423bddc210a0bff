"""Configurations beyond the defaults of dim_config (simulation/config.hpp:11-29): Gauss
points per element other than p+1 (quad_order), repeated interior knots (repeated_nodes,
bspline.cpp:26-43) and a degree per direction (config_2d/3d hold one dim_config per
direction, config.hpp:46-56). The general element loop assembles the rhs; the step, solve
and boundary handling are the same as for the defaults. Bars: FP64 rhs <= 1e-13 and steps
<= 1e-12 against the oracle; FP32 steps <= 1e-6 against the FP64 oracle."""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu

CONFIGS = [
    # dim, p, elements, quad_order, repeated_nodes
    (2, (2, 2), (21, 17), (5, 4), (0, 0)),
    (2, (3, 1), (15, 27), (0, 0), (0, 0)),
    (2, (3, 3), (12, 10), (2, 3), (0, 0)),  # fewer Gauss points than p+1
    (2, (2, 3), (14, 11), (0, 0), (1, 2)),
    (3, (1, 2, 3), (9, 7, 8), (0, 0, 0), (0, 0, 0)),
    (3, (2, 2, 2), (8, 6, 7), (4, 3, 5), (1, 0, 2)),
    (3, (3, 2, 1), (7, 9, 6), (6, 4, 2), (2, 1, 0)),  # fewer Gauss points than p+1 in z (p = 1)
]


def _log(msg):
    print(msg, flush=True)


def _pair(dim, p, ne, q, r, precision=ag.FP64, dirichlet=False):
    g = ag.Simulation(dim, list(p), list(ne), precision=precision, quad_order=list(q), repeated_nodes=list(r))
    o = po.Sim(dim, list(p), list(ne), quad_order=list(q), repeated_nodes=list(r))
    if dirichlet:
        g.fix_all_boundaries()
        o.fix_all_boundaries()
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    assert g.shape == o.shape
    return g, o


@pytest.mark.parametrize("dim,p,ne,q,r", CONFIGS)
def test_general_rhs_forms_match_oracle(dim, p, ne, q, r):
    _log(f"case: general rhs dim={dim} p={p} ne={ne} quad={q} repeated={r}")
    g, o = _pair(dim, p, ne, q, r)
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 1, g.size)
    kf = np.exp(rng.uniform(-0.5, 0.5, g.size))
    src = rng.standard_normal(g.size)
    g.set(ag.U_PREV, c)
    g.set(ag.KFIELD, kf)
    g.set(ag.SOURCE, src)
    dt = 2e-3
    for gk, ok in ((dict(), dict()), (dict(mass=0.0, grad=(1.0, 0.5, 2.0)), dict(mass=0.0, grad=(1.0, 0.5, 2.0))),
                   (dict(kfield=True, reaction=0.5), dict(kfield=kf, reaction=0.5)),
                   (dict(source=True), dict(source=src))):
        g.compute_rhs(ag.rhs_form(dt, **gk))
        expect = o.compute_rhs(c, po.heat_form(dt, dim, **ok))
        assert po.rel_l2(g.get(ag.U), expect) < 1e-13, gk
    g.close()


@pytest.mark.parametrize("dim,p,ne,q,r", CONFIGS)
@pytest.mark.parametrize("dirichlet", [False, True])
def test_general_steps_match_oracle(dim, p, ne, q, r, dirichlet):
    _log(f"case: general steps dim={dim} p={p} ne={ne} quad={q} repeated={r} dirichlet={dirichlet}")
    g, o = _pair(dim, p, ne, q, r, dirichlet=dirichlet)
    u = o.init_gaussians(seed=41, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    if dirichlet:
        o.zero_boundary(u)
    g.set(ag.U, u)
    flags = ag.STEP_DIRICHLET if dirichlet else 0
    dt = 1e-5
    for _ in range(3):
        rhs = o.compute_rhs(u, po.heat_form(dt, dim))
        if dirichlet:
            o.zero_boundary(rhs)
        u = o.solve(rhs)
        g.step(1, ag.rhs_form(dt), flags)
        assert po.rel_l2(g.get(ag.U), u) <= 1e-12
    g.close()


@pytest.mark.parametrize("dim,p,ne,q,r", [c for c in CONFIGS if len(set(c[1])) == 1] + [CONFIGS[1], CONFIGS[4]])
def test_general_fp32_steps_against_fp64_oracle(dim, p, ne, q, r):
    """FP32 increment steps (the Dirichlet boundary residual needs one degree in every
    direction, so handles with a degree per direction run natural boundaries here)."""
    mixed = len(set(p)) > 1
    _log(f"case: general FP32 dim={dim} p={p} ne={ne} quad={q} repeated={r}")
    g, o = _pair(dim, p, ne, q, r, precision=ag.FP32, dirichlet=not mixed)
    u = o.init_gaussians(seed=31, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    if not mixed:
        o.zero_boundary(u)
    g.set(ag.U, u)
    dt = 1e-5
    for _ in range(4):
        rhs = o.compute_rhs(u, po.heat_form(dt, dim))
        if not mixed:
            o.zero_boundary(rhs)
        u = o.solve(rhs)
        g.step(1, ag.rhs_form(dt), 0 if mixed else ag.STEP_DIRICHLET)
        assert po.rel_l2(g.get(ag.U), u) <= 1e-6
    g.close()


def test_general_delta_step_matches_oracle():
    """The implicit delta step with M + dt/2 K per direction on a general configuration."""
    dim, p, ne, q, r = CONFIGS[5]
    g, o = _pair(dim, p, ne, q, r)
    dt = 1e-3
    for d in range(dim):
        g.set_matrix(d, 1, 1.0, 0.5 * dt)
        o.set_matrix(d, 1, 1.0, 0.5 * dt)
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    u = o.init_gaussians(seed=5, count=3, sigma=(0.05, 0.15), centre=(0.2, 0.8), noise=1e-3)
    g.set(ag.U, u)
    f = po.heat_form(dt, dim)
    f.mass = 0.0
    for _ in range(3):
        rr = o.compute_rhs(u, f)
        o.solve(rr, slots=[1] * dim)
        u = u + rr
        g.step_delta(1, ag.rhs_form(dt, mass=0.0), [1] * dim)
        assert po.rel_l2(g.get(ag.U), u) <= 1e-12
    g.close()


def test_general_config_refusals():
    """Entry points whose kernels assume one degree and p+1 Gauss points refuse general
    handles with ADSG_EINVAL instead of computing with the wrong tables."""
    g = ag.Simulation(2, [3, 1], [8, 9])
    g.prepare_matrices()
    g.set(ag.U, np.ones(g.size))
    with pytest.raises(ag.AdsgError, match="norm"):
        g.norm(0)
    with pytest.raises(ag.AdsgError, match="sample"):
        g.sample(pts=[np.linspace(0, 1, 3), np.linspace(0, 1, 4)])
    g.close()
    # out-of-range configurations are refused at creation
    for kw in (dict(quad_order=[1, 3]), dict(quad_order=[65, 3]), dict(repeated_nodes=[3, 0])):
        with pytest.raises(ag.AdsgError):
            ag.Simulation(2, [2, 2], [6, 6], **kw)
