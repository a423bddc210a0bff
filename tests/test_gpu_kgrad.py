"""GPU parity of the k-field gradient of the explicit adjoint step, d/dk_b of w^T rhs(u; k):
the diffusion term of the element loop with a coefficient field (validation.hpp:110-115,
flow.hpp:73-100) is linear in k, so
    d/dk_b = -dt sum_e sum_q w_q J_e B_b(x_q) sum_d grad_d (d_d u)(d_d w),
which the oracle computes element by element and point by point (orc_sim_kgrad, pinned by
linearity against its element loop in test_oracle_rhs.py).

One adjoint step through the C ABI per case, on meshes that leave partial tiles (the
transposed solve of the step needs lines of more than 16 DOFs), on the tiles of the element
loop (2D: 16 x 16 elements; 3D: 8 x 8 x 4, at p = 3 8 x 8 x 3), in one launch with the
fix-up of its spill arrays (tiling 1) and in the parity-colour launches (tiling 0), and for a few
cases with the x lines of the transposed sweeps read lane by lane (sweep variant 2). Distinct gradient
weights per direction, so a swapped derivative fails. GRAD_K starts at a random vector, so
the kernels must add to it. Bar: the increment of GRAD_K and the new lambda <= 1e-12
relative L2 against the oracle.
"""
import numpy as np
import pytest

import adsgpu as ag
import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    (2, 1, (40, 36), 1), (2, 2, (40, 36), 1), (2, 3, (40, 36), 1),
    (2, 1, (40, 36), 0), (2, 2, (40, 36), 0), (2, 3, (40, 36), 0),
    (2, 2, (17, 50), 1), (2, 3, (30, 17), 1), (2, 2, (16, 16), 1), (2, 1, (33, 18), 0),
    (3, 1, (17, 18, 16), 1), (3, 2, (16, 17, 15), 1), (3, 3, (14, 15, 16), 1),
    (3, 1, (17, 18, 16), 0), (3, 2, (16, 17, 15), 0), (3, 3, (14, 15, 16), 0), (3, 2, (24, 16, 18), 1),
]
# (the same with x lines read lane by lane in the transposed sweeps: sweep variant 2)
SWEEP2 = [(2, 2, (40, 36), 1), (3, 2, (16, 17, 15), 1), (3, 3, (14, 15, 16), 0)]


@pytest.mark.parametrize("dim,p,ne,tiling,sweep", [c + (0,) for c in CASES] + [c + (2,) for c in SWEEP2])
def test_kgrad_adjoint_step_matches_oracle(dim, p, ne, tiling, sweep):
    print(f"case: k-field gradient {dim}D p={p} {ne} tiling={tiling} sweep variant {sweep}", flush=True)
    dt = 1e-5
    grad = (1.0, 0.7, 1.3)[:dim] + (1.0,) * (3 - dim)
    g = ag.Simulation(dim, [p] * dim, list(ne))
    o = po.Sim(dim, [p] * dim, list(ne))
    g.set_tuning(ag.TUNE_RHS_TILING, tiling)
    g.set_tuning(ag.TUNE_SWEEP_VARIANT, sweep)
    assert g.prepare_matrices() == 0 and o.prepare_matrices() == 0
    rng = np.random.default_rng(1000 * dim + 10 * p + sum(ne))
    k = 0.5 + rng.random(g.size)
    un = rng.standard_normal(g.size)
    lam = rng.standard_normal(g.size)
    dk0 = 1e-3 * rng.standard_normal(g.size)
    g.set(ag.KFIELD, k)
    g.set(ag.GRAD_K, dk0)
    g.set(ag.U, lam)
    g.set(ag.U_PREV, un)
    g.adjoint_step(ag.rhs_form(dt, grad=grad, kfield=True))

    oform = po.heat_form(dt, dim, grad=grad, kfield=k)
    w = lam.copy()
    o.solve(w, transposed=True)
    dk = o.kgrad(un, w, oform)
    lam1 = o.compute_rhs(w, oform)
    e_dk = po.rel_l2(g.get(ag.GRAD_K) - dk0, dk)
    e_lam = po.rel_l2(g.get(ag.U), lam1)
    print(f"  dk {e_dk:.3e}  lam {e_lam:.3e}", flush=True)
    assert e_dk <= 1e-12 and e_lam <= 1e-12
    g.close()
