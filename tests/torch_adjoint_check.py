"""Run as a script (in its own process, so that torch brings up the GPU runtime before
libadsgpu.so is loaded): the libtorch autograd Function over the ADS heat step with a k field: the 2D inverse
problem u_t = div(k grad u), p=2, 128^2 elements, 200 steps of dt = 1e-6, Dirichlet
boundary, k = exp(g) in [0.5, 2], L = ||u(T) - u_target||^2 with u_target from another
seeded k. Gradients from loss.backward() (checkpoints every 20 steps) against the oracle's
FP64 discrete adjoint (<= 1e-10) and central finite differences, h = 1e-6, on 32 seeded
directions with standard-normal entries (<= 1e-5). Then the Function with a source field f
as well (2D p=2 48x40, 40 steps of dt = 1e-5, checkpoints every 8): dL/dk, dL/df and dL/du0
against the oracle's discrete adjoint (<= 1e-10; dL/df = sum over steps of dt M w) and dL/df
against central differences on 8 directions (L is quadratic in f: <= 1e-6).
"""
import os
import sys

import numpy as np

import torch  # first: torch initialises the HIP runtime that libadsgpu.so then shares

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", "oracle"), os.path.join(HERE, "..", "iga-ads_amd")]

import adsgpu as ag
import pyoracle as po
from adsgpu import workloads as wl

P, NE, DT, STEPS, EVERY = 2, (128, 128), 1e-6, 200, 20


def _log(msg):
    print(msg, flush=True)


def _kfield(seed):
    g = wl.sine_source([P, P], list(NE), seed=seed)
    g = g / np.abs(g).max() * np.log(2.0)
    return np.exp(g)


def main():
    _log("case: torch heat_k 2D p=2 128^2 200 steps, checkpoints every 20")
    dev = torch.device("cuda:0")
    sim = ag.Simulation(2, [P, P], list(NE))
    sim.fix_all_boundaries()
    assert sim.prepare_matrices() == 0
    o = po.Sim(2, [P, P], list(NE))
    o.fix_all_boundaries()
    assert o.prepare_matrices() == 0
    u0 = wl.gaussians([P, P], list(NE), seed=2024, count=1, sigma=(0.1, 0.1), noise=1e-3)
    o.zero_boundary(u0)
    k = _kfield(61)
    kt = _kfield(62)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    with torch.no_grad():
        ut = ag.heat_k(sim, t(kt), t(u0), STEPS, EVERY, DT)
    kk = t(k).requires_grad_(True)
    uu = t(u0).requires_grad_(True)
    uT = ag.heat_k(sim, kk, uu, STEPS, EVERY, DT)
    loss = ((uT - ut) ** 2).sum()
    loss.backward()
    dk = kk.grad.cpu().numpy()
    du0 = uu.grad.cpu().numpy()

    # oracle: forward trajectory and discrete adjoint in FP64 on the CPU
    form = po.heat_form(DT, 2, kfield=k)
    traj = [u0]
    for _ in range(STEPS):
        r = o.compute_rhs(traj[-1], form, threads=8)
        o.zero_boundary(r)
        traj.append(o.solve(r))
    assert po.rel_l2(uT.detach().cpu().numpy(), traj[-1]) <= 1e-12
    lam = 2.0 * (traj[-1] - ut.cpu().numpy())
    dk_o = np.zeros_like(k)
    for n in range(STEPS - 1, -1, -1):
        w = o.solve(lam.copy(), transposed=True)
        o.zero_boundary(w)
        dk_o += o.kgrad(traj[n], w, form)
        lam = o.compute_rhs(w, form, threads=8)
    e_k, e_u = po.rel_l2(dk, dk_o), po.rel_l2(du0, lam)
    _log(f"  dL/dk vs oracle {e_k:.3e}  dL/du0 vs oracle {e_u:.3e}")
    assert e_k <= 1e-10 and e_u <= 1e-10

    def L(kv):
        with torch.no_grad():
            v = ag.heat_k(sim, t(kv), t(u0), STEPS, EVERY, DT)
            return float(((v - ut) ** 2).sum())

    rng = np.random.default_rng(63)
    h = 1e-6
    worst = 0.0
    for j in range(32):
        # standard-normal entries, the scale of k itself: with unit-norm directions the
        # signal 2h (dL/dk . d) drops to the rounding noise of L over 200 steps
        d = rng.standard_normal(k.size)
        fd = (L(k + h * d) - L(k - h * d)) / (2 * h)
        ad = float(dk @ d)
        worst = max(worst, abs(fd - ad) / abs(ad))
    _log(f"  32 directions: worst relative FD disagreement {worst:.2e}")
    assert worst <= 1e-5
    sim.close()


def check_source():
    p, ne, dt, steps, every = 2, (48, 40), 1e-5, 40, 8
    _log("case: torch heat_kf 2D p=2 48x40 40 steps, checkpoints every 8")
    dev = torch.device("cuda:0")
    sim = ag.Simulation(2, [p, p], list(ne))
    sim.fix_all_boundaries()
    assert sim.prepare_matrices() == 0
    o = po.Sim(2, [p, p], list(ne))
    o.fix_all_boundaries()
    assert o.prepare_matrices() == 0
    u0 = wl.gaussians([p, p], list(ne), seed=7, count=2, sigma=(0.1, 0.2), noise=1e-3)
    o.zero_boundary(u0)
    k = np.exp(0.5 * wl.sine_source([p, p], list(ne), seed=8))
    f = wl.sine_source([p, p], list(ne), seed=9)
    target = wl.gaussians([p, p], list(ne), seed=10, count=1, sigma=(0.2, 0.2))
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    kk, ff, uu = (t(a).requires_grad_(True) for a in (k, f, u0))
    uT = ag.heat_kf(sim, kk, ff, uu, steps, every, dt)
    loss = ((uT - t(target)) ** 2).sum()
    loss.backward()
    dk, df, du0 = (a.grad.cpu().numpy() for a in (kk, ff, uu))

    form = po.heat_form(dt, 2, kfield=k, source=f)
    gram = po.heat_form(0.0, 2)  # M v
    traj = [u0]
    for _ in range(steps):
        r = o.compute_rhs(traj[-1], form, threads=8)
        o.zero_boundary(r)
        traj.append(o.solve(r))
    assert po.rel_l2(uT.detach().cpu().numpy(), traj[-1]) <= 1e-12
    lam = 2.0 * (traj[-1] - target)
    dk_o, df_o = np.zeros_like(k), np.zeros_like(f)
    lin = po.heat_form(dt, 2, kfield=k)  # the part of the form that depends on u
    for n in range(steps - 1, -1, -1):
        w = o.solve(lam.copy(), transposed=True)
        o.zero_boundary(w)
        dk_o += o.kgrad(traj[n], w, lin)
        df_o += dt * o.compute_rhs(w, gram, threads=8)
        lam = o.compute_rhs(w, lin, threads=8)
    errs = (po.rel_l2(dk, dk_o), po.rel_l2(df, df_o), po.rel_l2(du0, lam))
    _log("  dL/dk, dL/df, dL/du0 vs oracle " + " ".join(f"{e:.3e}" for e in errs))
    assert max(errs) <= 1e-10

    def L(fv):
        with torch.no_grad():
            v = ag.heat_kf(sim, t(k), t(fv), t(u0), steps, every, dt)
            return float(((v - t(target)) ** 2).sum())

    rng = np.random.default_rng(11)
    h = 1e-3
    worst = 0.0
    for _ in range(8):
        d = rng.standard_normal(f.size)
        fd = (L(f + h * d) - L(f - h * d)) / (2 * h)
        ad = float(df @ d)
        worst = max(worst, abs(fd - ad) / abs(ad))
    _log(f"  8 directions: worst relative FD disagreement in f {worst:.2e}")
    assert worst <= 1e-6
    sim.close()


if __name__ == "__main__":
    main()
    check_source()
    print("torch adjoint check: ok", flush=True)
