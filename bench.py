"""Benchmark of the matrix-free heat operator path on one MI355X.

A step is one FP64 operator apply (vmult) of the Q3 heat operator
J = (rho c_p/dt) M + K_k on 128^3 cells (57,066,625 DoFs, 134,217,728
quadrature points) with the conductivity k = k0 (1 + 0.5 tanh((T - 1)/0.1))
given per quadrature point from T ~ U[0,2] (seed 2) and src ~ U[-1,1] (seed 3),
inputs resident in HBM. Prints one JSON line (see the repository's
benchmark contract): value = GDoF/s of the apply; roofline of the apply
kernel from algorithmic bytes (16 B per DoF for src/dst + 8 B per
quadrature point for k) over the kernel time measured with HIP events on the
launch stream; cpu_baseline = the CPU oracle's apply (same algorithm, OpenMP)
on a bounded sample. With --solve (default) it also times one heat-step
solve to ||b - J x|| <= 1e-10 ||b|| with FP64 CG + FP64 GMG and with
flexible CG + FP32 GMG. With --scan (default) it times the implicit steps of a
laser scan on the same mesh (Gaussian flux on the top face moving at speed 1,
sigma 0.05, peak 1e3, k = 1, dt = 1e-3, T0 = 0): steps 1..N_train run with
linear extrapolation while the network's training pairs are recorded, the
network is trained on the GPU, then steps N_train+1..N_train+N_eval are replayed
from the same state with each solver/predictor pair, each step timed from the
predictor to ||b - J T|| <= 1e-10 ||b||.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the apply kernel at Q3 128^3 (per-qp k), from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) scaled by a device copy of known size; the summary with the commands
# is profiles/heat_apply_q3_128_fp64_hbm_traffic.md. Counters need their own profiler runs, so
# the bench reports the profiled figure for that workload and null for others.
PROFILED_TRAFFIC = {(3, 128): 2.499e+09}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--solve", type=int, default=1)
    ap.add_argument("--cpu", type=int, default=1)
    ap.add_argument("--cpu-cells", type=int, default=64)
    ap.add_argument("--scan", type=int, default=1)
    ap.add_argument("--scan-train", type=int, default=30)
    ap.add_argument("--scan-eval", type=int, default=10)
    ap.add_argument("--scan-epochs", type=int, default=4)
    ap.add_argument("--scan-coarsening", type=int, default=4)
    ap.add_argument("--scan-law", type=int, default=0,
                    help="scan with k = k0 (1 + 0.5 tanh((T - 1)/0.1)) from the T DoFs (frozen at T_n each "
                         "step) and T0 = 0.95 instead of k = 1 and T0 = 0")
    ap.add_argument("--scan-coarse-cells", type=int, default=4,
                    help="cells per direction of the multigrid's coarse Q2 level in the scan")
    ap.add_argument("--scan-variants", default="all",
                    help="comma-separated solver/predictor rows of time_to_solve (default: all)")
    ap.add_argument("--copy", type=int, default=0,  # kept for old command lines; the copy rate is always measured
                    help="also run 3 device copies of the src vector (traffic calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def make_operator(torch, M, k, N, dt=1e-3):
    n1 = k + 1
    op = M.HeatDiffuseMultiPhaseOperator(k, (N, N, N), (1.0, 1.0, 1.0), rho_cp=1.0)
    op.reset_time_increment(dt)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    Tq = torch.rand(op.n_q_points, dtype=torch.float64, device="cuda", generator=g) * 2.0
    kq = 1.0 * (1.0 + 0.5 * torch.tanh((Tq - 1.0) / 0.1))
    del Tq
    op.set_conductivity_qp(kq)
    return op, kq


def cpu_baseline(k, N, threads):
    """CPU oracle apply (sum factorisation, OpenMP over coloured cells) on a bounded sample."""
    import numpy as np
    from oracle import oracle as O
    n1 = k + 1
    rng = np.random.default_rng(2)
    Tq = rng.uniform(0.0, 2.0, N ** 3 * n1 ** 3)
    op = O.Operator(k, (N, N, N), kappa_qp=1.0 + 0.5 * np.tanh((Tq - 1.0) / 0.1), c_mass=1e3)
    x = np.random.default_rng(3).uniform(-1.0, 1.0, op.n_dofs)
    op.vmult(x, threads=threads)
    reps, t0 = 0, time.perf_counter()
    while True:
        op.vmult(x, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if (reps >= 3 and el > 10.0) or reps >= 400:
            break
    return op.n_dofs * reps / el / 1e9, reps, el, op.n_dofs


def solve_timings(torch, M, op, reps=2):
    """Time-to-solve of one heat step (x0 = 0) to ||b - J x|| <= 1e-10 ||b||."""
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    b = torch.rand(op.n_dofs, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
    out = {}
    for name, prec, stype in (("fp64_cg_gmg64", M.FP64, M.SOLVER_CG), ("fp64_fcg_gmg32", M.FP32, M.SOLVER_FCG)):
        pc = M.GMGPreconditioner(op, precision=prec, n_cells_coarse=4, coarse_degree=2, smoother_degree=3)
        t0 = time.perf_counter()
        pc.update()
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        data = M.LinearSolverData(solver_type=stype, rel_tolerance=1e-10, criterion=M.CRIT_RHS, max_iterations=500)
        best, its = None, None
        for _ in range(reps):
            x = torch.zeros_like(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            its = M.LinearSolver.solve(op, x, b, data, pc)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        r = torch.empty_like(b)
        op.vmult(r, x)
        rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        out[name] = {"ms": round(best * 1e3, 3), "iterations": its, "final_rel_residual": rel,
                     "precond_setup_s": round(setup, 3)}
        del pc
    out["speedup_fp32_gmg_vs_fp64_gmg"] = round(out["fp64_cg_gmg64"]["ms"] / out["fp64_fcg_gmg32"]["ms"], 3)
    return out


def scan_timings(torch, M, k, N, n_train, n_eval, epochs, rho, only="all", law=False, coarse=4):
    """Time-to-solve per step of a laser scan (see the module docstring)."""
    import numpy as np
    from meltpooldg_amd.application import LaserScan, generate_trajectory, train
    dt = 1e-3
    kw = dict(n_cells_coarse=coarse)
    if law:
        kw["law"] = (M.LAW_TANH, (1.0, 0.5, 1.0, 0.1))
    t0 = time.perf_counter()
    _, Ts, qs, zs, _ = generate_trajectory(k, N, n_train, T_init=0.95 if law else None, dt=dt, precond="gmg32",
                                           predictor=M.PREDICTOR_LINEAR, **kw)
    torch.cuda.synchronize()
    t_traj = time.perf_counter() - t0
    wanted = None if only == "all" else set(only.split(","))
    net, losses, t_train = None, [None], 0.0
    if wanted is None or "fcg_gmg32_learned" in wanted:
        net = M.LearnedPredictor((N, N, N), degree=k, coarsening=rho, qscale=dt * N, seed=11)
        t0 = time.perf_counter()
        losses = train(net, Ts, qs, zs, epochs=epochs, lr=2e-3)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
    del qs, zs
    out = {"workload": "laser scan, Q%d %d^3 cells, %s, dt=1e-3, steps %d-%d, ||b-JT|| <= 1e-10 ||b||, coarse Q2 "
                       "%d^3; FP64 GMG: Chebyshev 3 pre / 3 post (CG), FP32 GMG: Chebyshev 1 pre / 3 post "
                       "(flexible CG)" % (k, N, "k = 1 + 0.5 tanh((T - 1)/0.1) from T (frozen at T_n), T0 = 0.95"
                                          if law else "k=1", n_train + 1, n_train + n_eval, coarse),
           "training": {"pairs": n_train - 2, "epochs": epochs, "coarsening": rho, "seconds": round(t_train, 2),
                        "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1],
                        "trajectory_seconds": round(t_traj, 2)}}
    variants = (("fp64_cg_gmg64_prev_step", "gmg64", M.PREDICTOR_NONE, 1),
                ("fp64_cg_gmg64_linear", "gmg64", M.PREDICTOR_LINEAR, 1),
                ("fp64_cg_gmg64_extrapolation6", "gmg64", M.PREDICTOR_EXTRAPOLATION, 6),
                ("fcg_gmg32_linear", "gmg32", M.PREDICTOR_LINEAR, 1),
                ("fcg_gmg32_extrapolation3", "gmg32", M.PREDICTOR_EXTRAPOLATION, 3),
                ("fcg_gmg32_extrapolation6", "gmg32", M.PREDICTOR_EXTRAPOLATION, 6),
                ("fcg_gmg32_learned", "gmg32", M.PREDICTOR_LEARNED, 1))
    variants = tuple(v for v in variants if wanted is None or v[0] in wanted or v[0] == "fp64_cg_gmg64_prev_step")
    for name, pc, pred, order in variants:
        case = LaserScan(k, N, dt=dt, precond=pc, predictor=pred,
                         learned=net if pred == M.PREDICTOR_LEARNED else None,
                         pre_smoother_degree=1 if pc == "gmg32" else 0, extrapolation_order=order, **kw)
        # restart from the recorded state; the replayed steps rebuild the history the predictor uses
        case.t = (n_train - order) * dt
        case.heat.set_initial_condition(Ts[n_train - order])
        for _ in range(order):
            case.step()
        ms, its, r0s, rs, fbs = [], [], [], [], []
        for _ in range(n_eval):
            it, el, (fb, r0, r) = case.step()
            ms.append(el * 1e3)
            its.append(it)
            r0s.append(r0)
            rs.append(r)
            fbs.append(fb)
        out[name] = {"ms_per_step": round(float(np.mean(ms)), 3), "iterations": float(np.mean(its)),
                     "log10_initial_rel_residual": round(float(np.mean(np.log10(r0s))), 3),
                     "fallbacks": int(sum(fbs)), "final_rel_residual_max": float(max(rs))}
        del case
        torch.cuda.empty_cache()
    base = out["fp64_cg_gmg64_prev_step"]["ms_per_step"]
    out["speedup_vs_fp64_prev_step"] = {n: round(base / out[n]["ms_per_step"], 3) for n, *_ in variants[1:]}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (the multi-GPU path is a design, not built)"
              % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    import meltpooldg_amd as M
    k, N = a.degree, a.cells
    op, kq = make_operator(torch, M, k, N)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    src = torch.rand(op.n_dofs, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
    dst = torch.empty_like(src)
    for _ in range(a.warmup):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    op.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern_ms, nlaunch = op.get_timing()
    op.set_timing(False)
    ms_step = el * 1e3 / a.steps
    kern_avg = kern_ms / max(nlaunch, 1)
    gdofs = op.n_dofs / (ms_step * 1e-3) / 1e9
    bytes_alg = 16.0 * op.n_dofs + 8.0 * op.n_q_points
    achieved = bytes_alg / (kern_avg * 1e-3) / 1e9
    line = {
        "metric": "GDoF/s operator apply; time-to-solve per heat step at 1e-10 (learned+FP32 vs FP64)",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (torch.Generator seeds 2 and 3 on the GPU)",
        "config": {"workload": "FP64 heat operator apply, Q%d hex, %d^3 cells, %d DoFs, QGauss(%d), k per quadrature "
                               "point = k0(1+0.5 tanh((T-1)/0.1)), T~U[0,2]" % (k, N, op.n_dofs, k + 1),
                   "degree": k, "cells": N, "n_dofs": op.n_dofs, "n_q_points": op.n_q_points},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": PROFILED_TRAFFIC.get((k, N)),
                     "kernel": "heat_apply_kernel<4,double,KAPPA_QP>", "kernel_avg_ms": round(kern_avg, 5),
                     "algorithmic_bytes": bytes_alg},
    }
    # measured copy bandwidth on the same vectors (device-to-device copy, read + write), the
    # attainable HBM rate next to the 8 TB/s spec
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 20
    line["hbm_copy"] = {"GBps": round(2 * 8 * op.n_dofs / (copy_ms * 1e-3) / 1e9, 1), "bytes_per_copy": 16 * op.n_dofs,
                        "note": "torch copy of one FP64 DoF vector, read + write"}
    del kq
    if a.solve and k == 3:
        try:
            line["solve"] = solve_timings(torch, M, op)
        except Exception as e:  # report, do not hide the apply line
            line["solve"] = {"error": str(e)}
    if a.scan:
        try:
            line["time_to_solve"] = scan_timings(torch, M, k, N, a.scan_train, a.scan_eval, a.scan_epochs,
                                                 a.scan_coarsening, a.scan_variants, law=bool(a.scan_law),
                                                 coarse=a.scan_coarse_cells)
        except Exception as e:  # report, do not hide the apply line
            line["time_to_solve"] = {"error": str(e)}
    if a.cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1))
        v, reps, el, nd = cpu_baseline(k, a.cpu_cells, threads)
        line["cpu_baseline"] = {"value": round(v, 5), "unit": "GDoF/s", "cores": threads, "kind": "port",
                                "sample": "CPU oracle apply (same sum-factorised algorithm, OpenMP, FP64), Q%d "
                                          "%d^3 cells (%d DoFs), k per quadrature point, %d applies in %.1f s"
                                          % (k, a.cpu_cells, nd, reps, el)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
