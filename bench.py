#!/usr/bin/env python3
"""DOF-updates/s of the ADS time step on one MI355X.

Workload: 3D explicit heat u_t = Lap u on [0,1]^3, quadratic B-splines (p=2, C^1),
128^3 elements (130^3 = 2,197,000 DOFs), 27 Gauss points per element, dt = 1e-6,
natural boundary (heat_3d.hpp), initial coefficients = 8 seeded Gaussians
(sigma in [0.05, 0.15], centres in [0.2, 0.8]^3) + uniform noise in [-1e-3, 1e-3].
One step = swap, element-loop RHS assembly, ADS solve (x, y, z line sweeps), replayed
from a captured HIP graph. Inputs are resident in HBM before the timed region.

Prints one JSON line. `value` is the plain FP64 step; the FP32 increment step (FP32
assembly and substitution on FP64-computed factors, FP64 state update) is reported in
`forms` with its error against the FP64 step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "iga-ads_amd"))

METRIC = ("DOF-updates/s on 1 GPU, FP32/batched/adjoint vs plain FP64, rel L2 ≤1e-6 vs FP64")
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (and FP64 matrix) peak
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def f_elem(d, p):
    """Sum-factorised flops per element of the constant-coefficient heat RHS:
    2*C_d*2*n^(d+1) + (d+2)*n^d with C_2 = 5, C_3 = 9 one-dimensional contractions."""
    n = p + 1
    c = 9 if d == 3 else 5
    return 2 * c * 2 * n ** (d + 1) + (d + 2) * n ** d


def run_form(ag, precision, ne, steps, warmup, u0, dt, variant=None, kron=False):
    """Plain explicit heat step (c2 form) in the given precision: autotune, warm up, time
    `steps` graph-replayed steps, then time the phases of 3 eager steps. kron: the rhs of
    this separable integrand through the 1D matrices instead of the element kernels."""
    sim = ag.Simulation(3, [2, 2, 2], [ne, ne, ne], precision=precision)
    sim.prepare_matrices()
    sim.set(ag.U, u0)
    form = ag.rhs_form(dt)
    if variant is None:
        variant = sim.autotune(form)  # picks the element-kernel variant; state is restored
    else:
        sim.set_tuning(ag.TUNE_RHS_VARIANT, variant)
    sim.set_tuning(ag.TUNE_RHS_KRONECKER, 1 if kron else 0)
    sim.step(warmup, form)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps, form)
    sim.synchronize()
    t1 = time.perf_counter()
    out = sim.get(ag.U)
    phases = sim.time_phases(3, form)  # eager, events per phase
    tiling = sim.get_tuning(ag.TUNE_RHS_TILING)
    sim.close()
    return (t1 - t0), phases, out, variant, tiling


def run_increment_form(ag, precision, ne, steps, warmup, u0, dt):
    """The explicit heat step in increment form, u = u_prev + M^-1 (-dt K u_prev) (the delta
    step with the mass-free form and the Gram matrices in every direction): the FP64
    counterpart of the FP32 increment step, same element-kernel instantiation."""
    sim = ag.Simulation(3, [2, 2, 2], [ne, ne, ne], precision=precision)
    sim.prepare_matrices()
    sim.set(ag.U, u0)
    form = ag.rhs_form(dt, mass=0.0)
    slots = [0, 0, 0]
    variant = sim.autotune(form, slots)
    sim.step_delta(warmup, form, slots)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step_delta(steps, form, slots)
    sim.synchronize()
    t1 = time.perf_counter()
    out = sim.get(ag.U)
    phases = sim.time_phases(3, form, slots)
    sim.close()
    return (t1 - t0), phases, out, variant


def func_rel_l2(ag, p, ne, got, ref):
    """Relative error of the 3D spline with coefficients `got` against `ref` in the function
    L2 norm (the M-weighted norm sqrt(e^T M e), by Gauss quadrature on the device)."""
    s = ag.Simulation(3, [p] * 3, [ne] * 3, precision=ag.FP64)
    s.set(ag.U, got - ref)
    e = s.norm(0)
    s.set(ag.U, ref)
    r = s.norm(0)
    s.close()
    return e / r


def run_implicit(ag, precision, ne, steps, warmup, u0, f, dt, flags, kron=False):
    """Implicit delta-form heat step with source (p=3): u += A^-1 dt (f - K u),
    A = (x)_d (M_d + dt/2 K_d)."""
    sim = ag.Simulation(3, [3, 3, 3], [ne, ne, ne], precision=precision)
    for d in range(3):
        sim.set_matrix(d, 1, 1.0, 0.5 * dt)
    sim.prepare_matrices()
    sim.set(ag.U, u0)
    sim.set(ag.SOURCE, f)
    form = ag.rhs_form(dt, mass=0.0, source=True)
    slots = [1, 1, 1]
    variant = sim.autotune(form, slots, flags)
    sim.set_tuning(ag.TUNE_RHS_KRONECKER, 1 if kron else 0)
    sim.step_delta(warmup, form, slots, flags)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step_delta(steps, form, slots, flags)
    sim.synchronize()
    t1 = time.perf_counter()
    out = sim.get(ag.U)
    phases = sim.time_phases(2, form, slots, flags)
    return sim, (t1 - t0), phases, out, variant


def implicit_block(ag, np, ne, steps, warmup):
    """3D implicit heat with source, p=3, ne^3, dt=1e-4: FP64, FP32 + refinement, and the
    adjoint of ||u(T)||^2 with respect to u(0)."""
    from adsgpu import workloads as wl

    dt = 1e-4
    p3 = [3, 3, 3]
    u0 = wl.gaussians(p3, [ne] * 3, seed=2025, count=8, sigma=(0.05, 0.15), centre=(0.2, 0.8),
                      noise=1e-3)
    f = wl.sine_source(p3, [ne] * 3, seed=2026)
    N = u0.size
    total = warmup + steps
    sim64, t64, ph64, x64, v64 = run_implicit(ag, ag.FP64, ne, steps, warmup, u0, f, dt, 0)
    res = {
        "workload": (f"3D implicit heat with source, p=3, {ne}^3 elements ({N} DOFs), dt=1e-4, "
                     f"{steps} timed steps after {warmup}; u += A^-1 dt (f - K u), "
                     f"A = (x)_d (M_d + dt/2 K_d); natural BC"),
        "fp64": {"value": N * steps / t64, "ms_per_step": 1e3 * t64 / steps, "phases_ms": ph64,
                 "rhs_variant": v64,
                 "rhs_tflops": f_elem(3, 3) * ne ** 3 / (ph64["rhs"] * 1e-3) / 1e12},
    }
    for name, flags in (("fp32", 0), ("fp32_refined", ag.STEP_REFINE)):
        s32, t32, ph32, x32, v32 = run_implicit(ag, ag.FP32, ne, steps, warmup, u0, f, dt, flags)
        s32.close()
        res[name] = {"value": N * steps / t32, "ms_per_step": 1e3 * t32 / steps,
                     "speedup_vs_fp64": t64 / t32, "phases_ms": ph32, "rhs_variant": v32,
                     f"rel_l2_vs_fp64_after_{total}_steps":
                         float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
                     f"function_rel_l2_vs_fp64_after_{total}_steps": func_rel_l2(ag, 3, ne, x32, x64)}
    # the same steps with the rhs of this separable integrand through the 1D matrices
    kr = {}
    form = ag.rhs_form(dt, mass=0.0, source=True)
    slots = [1, 1, 1]
    for name, prec, flags in (("fp64", ag.FP64, 0), ("fp32_refined", ag.FP32, ag.STEP_REFINE)):
        sk, tk, phk, xk, _ = run_implicit(ag, prec, ne, steps, warmup, u0, f, dt, flags, kron=True)
        kr[name] = {"value": N * steps / tk, "ms_per_step": 1e3 * tk / steps, "phases_ms": phk,
                    "speedup_vs_element_fp64": t64 / tk,
                    f"rel_l2_vs_element_fp64_after_{total}_steps":
                        float(np.linalg.norm(xk - x64) / np.linalg.norm(x64)),
                    f"function_rel_l2_vs_element_fp64_after_{total}_steps":
                        func_rel_l2(ag, 3, ne, xk, x64)}
        if prec == ag.FP64:  # the adjoint of ||u(T)||^2 with the same rhs form
            sk.set(ag.U, 2.0 * xk)
            sk.adjoint_step_delta(1, form, slots)
            sk.set(ag.U, 2.0 * xk)
            sk.synchronize()
            ta = time.perf_counter()
            sk.adjoint_step_delta(total, form, slots)
            sk.synchronize()
            ta = time.perf_counter() - ta
            kr["adjoint_fp64"] = {"value": N * total / ta, "ms_per_step": 1e3 * ta / total, "steps": total}
        sk.close()
    res["kronecker_rhs"] = kr
    # adjoint: lam(T) = 2 u(T) for L = ||u(T)||^2, run back over the same number of steps
    form = ag.rhs_form(dt, mass=0.0, source=True)
    slots = [1, 1, 1]
    sim64.set(ag.U, 2.0 * x64)
    sim64.adjoint_step_delta(1, form, slots)  # graph capture and code load
    sim64.set(ag.U, 2.0 * x64)
    sim64.synchronize()
    ta = time.perf_counter()
    sim64.adjoint_step_delta(total, form, slots)
    sim64.synchronize()
    ta = time.perf_counter() - ta
    grad = sim64.get(ag.U)

    def loss(u):
        sim64.set(ag.U, u)
        sim64.step_delta(total, form, slots)
        v = sim64.get(ag.U)
        return float(v @ v)

    # L is exactly quadratic in u(0), so the central difference is exact up to rounding for
    # any h; h = 1e-3 keeps the rounding of L (about 1e6) below 1e-8 relative
    d = np.random.default_rng(2027).standard_normal(N)
    d /= np.linalg.norm(d)
    h = 1e-3
    fd = (loss(u0 + h * d) - loss(u0 - h * d)) / (2 * h)
    ad = float(grad @ d)
    sim64.close()
    res["adjoint_fp64"] = {"value": N * total / ta, "ms_per_step": 1e3 * ta / total,
                           "steps": total, "fd_direction_rel_err": abs(fd - ad) / abs(ad),
                           "fd_h": h}
    return res


def ensemble_block(ag, np, count, steps, dt):
    """Batched 2D reaction-diffusion ensemble: `count` problems of uneven size and degree,
    `steps` explicit steps from one captured graph; FP32 (increment scheme) and FP64."""
    from adsgpu import workloads as wl

    probs = wl.ensemble(seed=2028, count=count)
    shape = [(q["p"], q["ne"]) for q in probs]
    forms = [ag.rhs_form(dt, grad=(q["eps"], q["eps"], 0.0), reaction=1.0) for q in probs]
    u0 = [wl.smooth_unit(q["p"], q["ne"], q["seed"]) for q in probs]
    dofs = sum(u.size for u in u0)
    elems = sum(q["ne"] ** 2 for q in probs)
    flops = sum(f_elem(2, q["p"]) * q["ne"] ** 2 for q in probs)
    out, res = {}, {}
    for name, prec in (("fp32", ag.FP32), ("fp64", ag.FP64)):
        tc0 = time.perf_counter()
        b = ag.Batch(shape, precision=prec)
        tc1 = time.perf_counter()
        b.prepare()
        b.synchronize()
        tc2 = time.perf_counter()
        b.set_forms(forms)
        for i, u in enumerate(u0):
            b.set(i, u)
        tile, tile_ms = b.autotune()
        lines = ["segmented", "one lane per line"][b.get_tuning(ag.TUNE_SWEEP_VARIANT)]
        b.step(steps)  # capture, instantiate and warm the graph of `steps` steps
        b.synchronize()
        for i, u in enumerate(u0):
            b.set(i, u)
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        t = time.perf_counter() - t0
        out[name] = [b.get(i) for i in range(len(probs))]
        b.close()
        res[name] = {"value": dofs * steps / t, "ms_per_step": 1e3 * t / steps,
                     "setup_ms": {"create": 1e3 * (tc1 - tc0), "prepare": 1e3 * (tc2 - tc1),
                                  "in_steps": (tc2 - tc0) / (t / steps),
                                  "note": ("every problem's basis tables (adsg_batch_create), Gram "
                                           "bands, FP64 band LU and packed sweep rows "
                                           "(adsg_batch_prepare), computed on the device; the "
                                           "host lays out the arenas and uploads their descriptors")},
                     "tile_shape": ["16x16", "8x8"][tile], "line_sweeps": lines,
                     "autotune_ms_per_step": list(map(float, tile_ms)),
                     "element_tflops": flops * steps / t / 1e12}
    err = max(float(np.linalg.norm(a - c) / np.linalg.norm(c)) for a, c in zip(out["fp32"], out["fp64"]))
    res["fp32"]["speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res["fp32"]["ms_per_step"]
    res["fp32"][f"max_rel_l2_vs_fp64_after_{steps}_steps"] = err
    res["workload"] = (f"{count} 2D reaction-diffusion problems u_t = eps Lap u + u(1-u), "
                       f"elements per side in [32, 256], p in {{1,2,3}}, eps log-uniform in "
                       f"[1e-4, 1e-2], dt={dt}, {steps} steps in one graph launch, natural BC; "
                       f"{dofs} DOFs, {elems} elements in total")
    return res


def adjoint_k_block(ag, np, ne, steps, every, dt):
    """2D inverse-problem shape: u_t = div(k grad u), p=2, ne^2, Dirichlet, k in [0.5, 2];
    forward steps and the checkpointed adjoint of ||u(T) - u_target||^2 w.r.t. k and u(0)."""
    from adsgpu import workloads as wl

    p = 2
    def kfield(seed):
        g = wl.sine_source([p, p], [ne, ne], seed=seed)
        return np.exp(g / np.abs(g).max() * np.log(2.0))

    k, kt = kfield(2030), kfield(2031)
    sim = ag.Simulation(2, [p, p], [ne, ne])
    sim.fix_all_boundaries()
    sim.prepare_matrices()
    N = sim.size
    u0 = wl.gaussians([p, p], [ne, ne], seed=2032, count=1, sigma=(0.1, 0.1), noise=1e-3)
    nx = ne + p
    u0.reshape(nx, nx)[[0, -1], :] = 0.0
    u0.reshape(nx, nx)[:, [0, -1]] = 0.0
    form = ag.rhs_form(dt, kfield=True)
    flags = ag.STEP_DIRICHLET

    def forward(kv, u):
        sim.set(ag.KFIELD, kv)
        sim.set(ag.U, u)
        sim.step(steps, form, flags)
        return sim.get(ag.U)

    ut = forward(kt, u0)
    forward(k, u0)  # graph capture and code load
    sim.set(ag.U, u0)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps, form, flags)
    sim.synchronize()
    tf = time.perf_counter() - t0
    uT = sim.get(ag.U)
    sim.set(ag.U, u0)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.forward_checkpointed(steps, every, form, flags)
    sim.synchronize()
    tc = time.perf_counter() - t0
    for timed in (False, True):  # the first run loads the adjoint kernels and captures its graphs
        sim.set(ag.GRAD_K, np.zeros(N))
        sim.set(ag.U, 2.0 * (uT - ut))
        sim.synchronize()
        t0 = time.perf_counter()
        sim.adjoint_checkpointed(form, flags)
        sim.synchronize()
        ta = time.perf_counter() - t0
    dk = sim.get(ag.GRAD_K)
    d = np.random.default_rng(2033).standard_normal(N)
    h = 1e-6
    loss = lambda kv: float(((forward(kv, u0) - ut) ** 2).sum())
    fd = (loss(k + h * d) - loss(k - h * d)) / (2 * h)
    ad = float(dk @ d)
    sim.close()
    return {"workload": (f"2D heat u_t = div(k grad u), p=2, {ne}^2 elements ({N} DOFs), "
                         f"{steps} steps of dt={dt}, Dirichlet, k seeded in [0.5, 2]; "
                         f"loss ||u(T) - u_target||^2, checkpoints every {every} steps"),
            "forward": {"value": N * steps / tf, "ms_per_step": 1e3 * tf / steps},
            "forward_checkpointed_ms_per_step": 1e3 * tc / steps,
            "adjoint": {"value": N * steps / ta, "ms_per_step": 1e3 * ta / steps,
                        "note": "recompute of every segment from its checkpoint included"},
            "fd_direction_rel_err": abs(fd - ad) / abs(ad), "fd_h": h}


def copy_calibration(mib, reps):
    """A plain device-to-device copy of known size through the HIP runtime (its copy
    kernel reads and writes mib MiB per launch): the reference point for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the step's kernels."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_mib": mib, "launches": reps}), flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_rhs3_heat3d_p2_128.md")


def launch_traffic(kernel):
    """HBM bytes per launch of the element kernel `kernel` (k_rhs3, k_rhs3_pipe, k_rhs3_zpair) at 128^3
    from the committed counter summary (the line `traffic_bytes_per_launch <kernel>: <bytes>`),
    or None when there is none."""
    key = f"traffic_bytes_per_launch {kernel}:"
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.strip().startswith(key):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def cpu_baseline(np, ne, u0, dt, N, E):
    """The oracle (the project's restatement of the reference CPU step) on the host cores:
    one core with the sequential element loop, and all cores of this process's share (at
    most 16 threads) with the std::thread executor (lock-serialised scatter, the reference's
    executor contract). Medians of 3 timings each; the single-core element loop is timed on
    a bounded sample of E/8 consecutive elements and scaled to the full mesh."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po  # the CPU baseline leg only

    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    threads = max(1, min(16, avail))
    o = po.Sim(3, [2, 2, 2], [ne, ne, ne])
    o.prepare_matrices()
    form = po.heat_form(dt, 3)
    # one core: the element loop over E/8 elements (scaled by 8), then the full solve
    count = E // 8
    t_rhs1, t_solve1 = [], []
    for _ in range(3):
        rhs = o.zeros()
        t0 = time.perf_counter()
        o.compute_rhs_sample(u0, form, 0, count, rhs)
        t_rhs1.append((time.perf_counter() - t0) * E / count)
    rhs = o.compute_rhs(u0, form, threads=threads)
    for _ in range(3):
        r = rhs.copy()
        t0 = time.perf_counter()
        o.solve(r)
        t_solve1.append(time.perf_counter() - t0)
    one = float(np.median(t_rhs1) + np.median(t_solve1))
    # all cores: full steps (threaded element loop, then the single-threaded solve)
    t_all = []
    for _ in range(3):
        t0 = time.perf_counter()
        r = o.compute_rhs(u0, form, threads=threads)
        o.solve(r)
        t_all.append(time.perf_counter() - t0)
    allc = float(np.median(t_all))
    return {
        "value": N / allc,
        "unit": "DOF-updates/s",
        "cores": threads,
        "kind": "port",
        "nproc": nproc,
        "cores_available": avail,
        "sample": (f"all-core leg: median of 3 full steps of the same {ne}^3 workload (element "
                   f"loop on {threads} std::thread workers with a lock-serialised scatter, then "
                   f"the 3-direction dgbtrs+transpose ADS solve on one thread): {allc:.2f} s "
                   f"per step"),
        "legs": {
            "1_core": {"value": N / one, "s_per_step": one,
                       "rhs_s": float(np.median(t_rhs1)), "solve_s": float(np.median(t_solve1)),
                       "sample": (f"sequential element loop over {count} of {E} elements "
                                  f"(median of 3, scaled by {E // count}) + the full solve "
                                  f"(median of 3)")},
            "all_cores": {"value": N / allc, "s_per_step": allc, "threads": threads,
                          "runs_s": t_all},
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--elements", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--rhs-variant", type=int, default=None,
                    help="element-kernel variant (default: autotuned)")
    ap.add_argument("--rhs-tiling", type=int, default=None, choices=(0, 1),
                    help="--eager-only: element-kernel tiling, 0 colour launches, 1 one launch "
                         "(default: the handle's)")
    ap.add_argument("--no-implicit", action="store_true",
                    help="skip the 3D p=3 implicit / adjoint block")
    ap.add_argument("--implicit-elements", type=int, default=192)
    ap.add_argument("--implicit-steps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched ensemble block")
    ap.add_argument("--no-kron", action="store_true",
                    help="skip the separable-integrand (1D-matrix) rhs block")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 2D k-field adjoint block")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the 2D k-field adjoint block (for profiling) and print its JSON")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain D2D copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--eager-only", action="store_true",
                    help="only run `steps` eager steps of the 3D p=2 workload (no graph "
                         "replay; for counter collection) and exit")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64",
                    help="arithmetic of the --eager-only steps")
    ap.add_argument("--degree", type=int, default=2, help="B-spline degree of the --eager-only steps")
    ap.add_argument("--mass", type=float, default=1.0,
                    help="mass coefficient of the --eager-only form (0: the mass-free increment form)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 4)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the ADS step here runs on a single MI355X "
              "(--gpus 1); multi-GPU slab decomposition is described but not built.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import adsgpu as ag
    from adsgpu import workloads as wl

    ne, dt = args.elements, 1e-6
    u0 = wl.gaussians([2, 2, 2], [ne] * 3, seed=2024, count=8, sigma=(0.05, 0.15),
                      centre=(0.2, 0.8), noise=1e-3)
    N = u0.size
    E = ne ** 3

    if args.adjoint_only:
        print(json.dumps({"adjoint_k_2d": adjoint_k_block(ag, np, 128, 200, 20, 1e-6)}), flush=True)
        return

    if args.eager_only:
        prec = ag.FP64 if args.precision == "fp64" else ag.FP32
        pd = [args.degree] * 3
        sim = ag.Simulation(3, pd, [ne, ne, ne], precision=prec)
        sim.prepare_matrices()
        sim.set(ag.U, u0 if args.degree == 2 else
                wl.gaussians(pd, [ne] * 3, seed=2024, count=8, sigma=(0.05, 0.15), centre=(0.2, 0.8),
                             noise=1e-3))
        form = ag.rhs_form(dt, mass=args.mass)
        variant = args.rhs_variant
        if variant is None:
            variant = sim.autotune(form)
        else:
            sim.set_tuning(ag.TUNE_RHS_VARIANT, variant)
        if args.rhs_tiling is not None:
            sim.set_tuning(ag.TUNE_RHS_TILING, args.rhs_tiling)
        tiling = sim.get_tuning(ag.TUNE_RHS_TILING)
        phases = sim.time_phases(args.steps, form)
        sim.close()
        print(json.dumps({"eager_steps": args.steps, "precision": args.precision, "degree": args.degree,
                          "elements": ne, "rhs_variant": variant, "rhs_tiling": tiling,
                          "phases_ms": phases}), flush=True)
        return

    rates = ag.device_rates()  # measured FMA ceilings of this device (vector and matrix)
    t64, ph64, x64, var64, til64 = run_form(ag, ag.FP64, ne, args.steps, args.warmup, u0, dt,
                                            args.rhs_variant)
    value = N * args.steps / t64
    ms_step = 1e3 * t64 / args.steps

    # dominant kernel: the element kernel (k_rhs3; k_rhs3_pipe, k_rhs3_zpair for variants 3, 4),
    # launched once over all tiles (tiling 1; its spill fix-up k_rhs_fixup runs at the start
    # of the sweep phase), or once per parity colour (8 launches per step, each over E/8
    # elements); HIP events on the handle's stream bracket the launches of the phase
    rhs_kernel = {3: "k_rhs3_pipe", 4: "k_rhs3_zpair"}.get(var64, "k_rhs3")
    one = til64 == 1 and var64 != 3
    launches = 1 if one else 8
    launch_flops = f_elem(3, 2) * E / launches
    launch_ms = ph64["rhs"] / launches
    traffic_key = rhs_kernel + (" (one launch)" if one else "")
    launch_desc = ("one launch over all tiles; its spill fix-up k_rhs_fixup (memory-bound, "
                   "timed in the x-sweep phase) not included" if one else "one launch per parity colour")
    rhs_tflops = launch_flops / (launch_ms * 1e-3) / 1e12
    sweep_bytes = 2 * N * 8  # one directional sweep: read and write every coefficient once
    sweep_ms = [ph64["sweep_x"], ph64["sweep_y"], ph64["sweep_z"]]
    sweep_gbs = [sweep_bytes / (m * 1e-3) / 1e9 for m in sweep_ms]

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D explicit heat, p=2, {ne}^3 elements ({N} DOFs), dt=1e-6, "
                        f"{args.steps} timed steps, natural BC, seeded Gaussians + noise",
            "dim": 3, "p": 2, "elements": ne, "dofs": N, "dt": dt,
        },
        "roofline": {
            "kernel": f"{rhs_kernel} (element-loop RHS assembly, variant {var64}; {launch_desc})",
            "bound": "valu_fp64",
            "achieved": rhs_tflops,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": rhs_tflops / PEAK_FP64_TFLOPS,
            "traffic": launch_traffic(traffic_key),
            "launch_flops": launch_flops,
            "launch_ms": launch_ms,
            "note": ("FP64 vector-ALU bound (MI355X FP64 vector peak = FP64 matrix peak = "
                     "78.6 TFLOP/s); achieved = 3051 flop/element x the elements of one launch "
                     "(E, or E/8 per colour launch) / average launch time (HIP events around the "
                     "phase's launches on the handle's stream); traffic = HBM bytes per launch "
                     "from rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a copy-kernel "
                     "calibration"),
        },
        "phases_ms": ph64,
        "rhs_variant": var64,
        "rhs_tiling": til64,
        "device_rates_tflops": dict(rates, note=(
            "FMA ceilings measured on this device by adsg_device_rates (independent FMA chains "
            "filling the device): vector FP64 / FP32 (plain FMAs, no packed instructions) and "
            "16x16x4 MFMA FP64 / FP32")),
        "solve_roofline": {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
            "achieved_xyz": sweep_gbs,
            "frac_xyz": [g / PEAK_HBM_GBS for g in sweep_gbs],
            "bytes_per_sweep": sweep_bytes,
        },
    }

    if not args.no_fp32:
        t32, ph32, x32, var32, til32 = run_form(ag, ag.FP32, ne, args.steps, args.warmup, u0, dt,
                                                args.rhs_variant)
        ti, phi, xi, vari = run_increment_form(ag, ag.FP64, ne, args.steps, args.warmup, u0, dt)
        err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
        kern32 = {3: "k_rhs3_pipe", 4: "k_rhs3_zpair"}.get(var32, "k_rhs3")
        tf32 = f_elem(3, 2) * E / (ph32["rhs"] * 1e-3) / 1e12
        result["forms"] = {
            "fp64": {"value": value, "ms_per_step": ms_step},
            "fp64_increment": {
                "value": N * args.steps / ti,
                "ms_per_step": 1e3 * ti / args.steps,
                "speedup_vs_fp64": t64 / ti,
                "rel_l2_vs_fp64_after_all_steps": float(np.linalg.norm(xi - x64) / np.linalg.norm(x64)),
                "rhs_variant": vari,
                "phases_ms": phi,
                "note": ("u = u_prev + M^-1 (-dt K u_prev) in FP64: the same scheme and the same "
                         "mass-free element-kernel instantiation as the FP32 increment step"),
            },
            "fp32_increment": {
                "value": N * args.steps / t32,
                "ms_per_step": 1e3 * t32 / args.steps,
                "speedup_vs_fp64": t64 / t32,
                "speedup_vs_fp64_increment": ti / t32,
                "rel_l2_vs_fp64_after_all_steps": err,
                "function_rel_l2_vs_fp64_after_all_steps": func_rel_l2(ag, 2, ne, x32, x64),
                "rhs_variant": var32,
                "rhs_tiling": til32,
                "phases_ms": ph32,
                "rhs_roofline": {
                    "kernel": f"{kern32} (FP32, variant {var32}, tiling {til32})",
                    "bound": "valu_fp32",
                    "achieved": tf32, "unit": "TFLOP/s",
                    "peak": 78.6, "frac": tf32 / 78.6,
                    "peak_mfma_fp32": 157.3, "frac_of_mfma_fp32": tf32 / 157.3,
                    "note": ("3051 flop/element x E elements / the rhs phase's time; "
                             "FP32 VALU without packed instructions issues at the FP64 rate "
                             "(78.6 TFLOP/s); 157.3 TFLOP/s is the FP32 MFMA / packed peak"),
                },
            },
        }

    if not args.no_kron:
        kron = {}
        for name, prec in (("fp64", ag.FP64), ("fp32", ag.FP32)):
            tk, phk, xk, _, _ = run_form(ag, prec, ne, args.steps, args.warmup, u0, dt, args.rhs_variant,
                                      kron=True)
            kron[name] = {"value": N * args.steps / tk, "ms_per_step": 1e3 * tk / args.steps,
                          "speedup_vs_element_fp64": t64 / tk, "phases_ms": phk,
                          "rel_l2_vs_element_fp64_after_all_steps":
                              float(np.linalg.norm(xk - x64) / np.linalg.norm(x64)),
                          "function_rel_l2_vs_element_fp64_after_all_steps":
                              func_rel_l2(ag, 2, ne, xk, x64)}
            # the rhs kernel reads u and writes rhs once: 2 N s bytes per step
            nbytes = 2 * N * (8 if prec == ag.FP64 else 4)
            kron[name]["rhs_roofline"] = {
                "kernel": "k_rhs_kron3", "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                "achieved": nbytes / (phk["rhs"] * 1e-3) / 1e9,
                "frac": nbytes / (phk["rhs"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                "bytes_per_launch": nbytes}
        kron["note"] = ("the same heat step with the rhs of this separable integrand assembled "
                        "from the 1D Gram and stiffness matrices (ADSG_TUNE_RHS_KRONECKER); the "
                        "p+1-point Gauss rule is exact for it, so it equals the element loop to "
                        "rounding")
        result.setdefault("forms", {})["kronecker_rhs"] = kron

    if not args.no_implicit:
        result.setdefault("forms", {})["implicit_p3"] = implicit_block(
            ag, np, args.implicit_elements, args.implicit_steps, 2)

    if not args.no_adjoint:
        result.setdefault("forms", {})["adjoint_k_2d"] = adjoint_k_block(ag, np, 128, 200, 20, 1e-6)

    if not args.no_batch:
        result.setdefault("forms", {})["batched_ensemble"] = ensemble_block(
            ag, np, args.batch_problems, args.batch_steps, 2e-5)

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(np, ne, u0, dt, N, E)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
