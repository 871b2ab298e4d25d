"""Benchmark of the matrix-free heat operator path on one MI355X.

A step is one FP64 operator apply (vmult) of the Q3 heat operator
J = (rho c_p/dt) M + K_k on 128^3 cells (57,066,625 DoFs, 134,217,728
quadrature points) with the conductivity k = k0 (1 + 0.5 tanh((T - 1)/0.1))
given per quadrature point from T ~ U[0,2] (seed 2) and src ~ U[-1,1] (seed 3),
inputs resident in HBM. Prints one JSON line (see the repository's
benchmark contract): value = GDoF/s of the apply; roofline of the apply
kernel from algorithmic bytes (16 B per DoF for src/dst + 8 B per quadrature
point for k) over the kernel time measured with HIP events on the launch
stream; hbm = the library's own copy / read / write streams next to the
8 TB/s spec; cpu_baseline = the CPU oracle (same algorithm, OpenMP) on bounded
samples: all-core and one-core applies and one CG + multigrid solve.

With --solve (default) it times one solve of a system on the same operator
(b ~ U[-1,1], x0 = 0) to ||b - J x|| <= 1e-10 ||b|| with FP64 CG + FP64
multigrid V(3,3), FP64 flexible CG + FP64 multigrid V(1,3) and flexible CG +
FP32 multigrid V(3,3) / V(1,3); each FP32-multigrid solution is compared with
the first FP64 one. With --scan (default) it times the implicit steps of a
laser scan on the same mesh (Gaussian flux on the top face moving at speed 1,
sigma 0.05, peak 1e3, dt = 1e-3) with the conductivity law above evaluated at
T_n (T0 = 0.95; --scan-law 0: k = 1, T0 = 0): steps 1..N_train run while the
networks' training pairs are recorded (linear and order-3 extrapolation as
base guesses), the networks are trained on the GPU, then steps
N_train+1..N_train+N_eval are replayed from the same state with each
solver/predictor pair, each step timed from the predictor to
||b - J T|| <= 1e-10 ||b||; every step's solution is compared with the FP64
baseline's (FP64 CG + FP64 multigrid from the previous step).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the apply kernel (degree, cells, conductivity, mapped), from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the library's streams of known size;
# the summaries with the commands are under profiles/ (heat_apply_*). Counters need their own
# profiler runs, so the bench reports the profiled figure for these workloads and null for others.
PROFILED_TRAFFIC = {(3, 128, "qp", False): 2.506e+09, (4, 96, "qp", False): 2.354e+09,
                    (4, 96, "law", False): 2.308e+09, (3, 96, "const", True): 3.759e+09}
# learned predictor: after PAUSE[0] consecutive steps whose learned guess loses the guess check, the
# network is not evaluated for PAUSE[1] steps, then tried again
PAUSE = (2, 8)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--kappa", default="qp", choices=("qp", "law", "const"),
                    help="conductivity of the benchmark apply: per quadrature point, law from T, constant")
    ap.add_argument("--mapped", type=int, default=0, help="apply on the curved box x + 0.05 sin(2 pi y) sin(2 pi z)")
    ap.add_argument("--solve", type=int, default=1)
    ap.add_argument("--cpu", type=int, default=1)
    ap.add_argument("--cpu-cells", type=int, default=64)
    ap.add_argument("--scan", type=int, default=1)
    ap.add_argument("--scan-train", type=int, default=30)
    ap.add_argument("--scan-eval", type=int, default=10)
    ap.add_argument("--scan-epochs", type=int, default=4)
    ap.add_argument("--scan-coarsening", type=int, default=4)
    ap.add_argument("--scan-law", type=int, default=1,
                    help="scan with k = k0 (1 + 0.5 tanh((T - 1)/0.1)) from the T DoFs (frozen at T_n each "
                         "step) and T0 = 0.95 (0: k = 1 and T0 = 0)")
    ap.add_argument("--scan-guess-smoothing", type=int, default=2,
                    help="Chebyshev steps of the finest multigrid level applied to an accepted learned guess")
    ap.add_argument("--learned-c4", type=int, default=1,
                    help="the learned guess on its training configuration (Q2 64^3, 200 steps)")
    ap.add_argument("--c4-coarsening", type=int, default=2, help="working-grid coarsening of that network")
    ap.add_argument("--c4-smoothing", type=int, default=2, help="guess smoothing steps in that block")
    ap.add_argument("--c4-orders", default="1,2", help="base extrapolation orders in that block")
    ap.add_argument("--cpu-solve-cells", type=int, default=32,
                    help="cells per direction of the CPU oracle's CG + multigrid solve sample")
    ap.add_argument("--scan-coarse-cells", type=int, default=4,
                    help="cells per direction of the multigrid's coarse Q2 level in the scan")
    ap.add_argument("--scan-variants", default="all",
                    help="comma-separated solver/predictor rows of time_to_solve (default: all)")
    ap.add_argument("--copy", type=int, default=0,  # kept for old command lines; the copy rate is always measured
                    help="also run 3 device copies of the src vector (traffic calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def make_operator(torch, M, k, N, dt=1e-3, kappa="qp", mapped=False):
    """The benchmark operator: k per quadrature point from T ~ U[0,2] (qp), the law evaluated from
    a T DoF vector ~ U[0,2] inside the apply (law), or k = 1 (const); mapped: the curved box
    x + 0.05 sin(2 pi y) sin(2 pi z) on all nodes. Returns the operator and the arrays it reads."""
    op = M.HeatDiffuseMultiPhaseOperator(k, (N, N, N), (1.0, 1.0, 1.0), rho_cp=1.0)
    op.reset_time_increment(dt)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    keep = []
    if kappa == "qp":
        Tq = torch.rand(op.n_q_points, dtype=torch.float64, device="cuda", generator=g) * 2.0
        kq = 1.0 * (1.0 + 0.5 * torch.tanh((Tq - 1.0) / 0.1))
        del Tq
        op.set_conductivity_qp(kq)
        keep.append(kq)
    elif kappa == "law":
        T = torch.rand(op.n_dofs, dtype=torch.float64, device="cuda", generator=g) * 2.0
        op.set_conductivity_law(M.LAW_TANH, (1.0, 0.5, 1.0, 0.1), T)
        keep.append(T)
    else:
        op.set_conductivity_constant(1.0)
    if mapped:
        X = M.node_coordinates(k, (N, N, N), mapping_amplitude=0.05)
        op.set_mapping(X)
        keep.append(X)
    return op, keep


def cpu_apply(k, N, threads, seconds=5.0):
    """CPU oracle apply (sum factorisation, OpenMP over coloured cells) on a bounded sample:
    GDoF/s, applies, seconds, DoFs."""
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
        if (reps >= 3 and el > seconds) or reps >= 400:
            break
    return op.n_dofs * reps / el / 1e9, reps, el, op.n_dofs


def cpu_solve(k, N, threads):
    """CPU oracle CG + geometric multigrid (Chebyshev 3, coarse Q2 on 4^3 cells) to
    ||b - J x|| <= 1e-10 ||b|| from x0 = 0, the same conductivity law per quadrature point."""
    import numpy as np
    from oracle import oracle as O
    n1 = k + 1
    rng = np.random.default_rng(2)
    Tq = rng.uniform(0.0, 2.0, N ** 3 * n1 ** 3)
    op = O.Operator(k, (N, N, N), kappa_qp=1.0 + 0.5 * np.tanh((Tq - 1.0) / 0.1), c_mass=1e3)
    b = np.random.default_rng(4).uniform(-1.0, 1.0, op.n_dofs)
    t0 = time.perf_counter()
    gmg = O.GMG(op, nc_coarse=4, coarse_degree=2, cheb_degree=3, threads=threads)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, it, res = op.cg(b, rel_tol=1e-10, precond=2, gmg=gmg, threads=threads)
    el = time.perf_counter() - t0
    return {"workload": "Q%d %d^3 cells (%d DoFs), dt = 1e-3, k per quadrature point, b ~ U[-1,1]" % (k, N, op.n_dofs),
            "ms": round(el * 1e3, 2), "iterations": int(it), "setup_s": round(setup, 3),
            "final_rel_residual": float(np.linalg.norm(b - op.vmult(x, threads=threads)) / np.linalg.norm(b)),
            "us_per_kdof": round(el * 1e6 / (op.n_dofs / 1e3), 3)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def stream_rates(torch, M, n):
    """The library's own copy / read / write streams over FP64 vectors of n entries (GB/s)."""
    n -= n % 2
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    part = torch.empty(4096, dtype=torch.float64, device="cuda")
    out = {}
    for name, kind, nbytes in (("copy", 0, 16 * n), ("read", 1, 8 * n), ("write", 2, 8 * n)):
        dst = part if kind == 1 else b
        for _ in range(3):
            M.stream_probe(kind, dst, a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            M.stream_probe(kind, dst, a)
        e1.record()
        torch.cuda.synchronize()
        out[name + "_GBps"] = round(nbytes / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9, 1)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    out["torch_copy_GBps"] = round(16 * n / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9, 1)
    out["bytes_per_vector"] = 8 * n
    del a, b, part
    return out


def solve_timings(torch, M, op, reps=2):
    """Time-to-solve of one system on the benchmark operator (x0 = 0) to ||b - J x|| <= 1e-10 ||b||,
    with the error of each solution against the FP64 CG + FP64 V(3,3) one."""
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    b = torch.rand(op.n_dofs, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
    out = {"workload": "b ~ U[-1,1] (seed 4), x0 = 0, ||b - J x|| <= 1e-10 ||b||, multigrid: Chebyshev degree 3, "
                       "coarse Q2 on 4^3 cells solved by a dense FP64 inverse"}
    x_ref = None
    rows = (("fp64_cg_gmg64_v33", M.FP64, M.SOLVER_CG, 0), ("fp64_fcg_gmg64_v13", M.FP64, M.SOLVER_FCG, 1),
            ("fcg_gmg32_v33", M.FP32, M.SOLVER_FCG, 0), ("fcg_gmg32_v13", M.FP32, M.SOLVER_FCG, 1))
    for name, prec, stype, pre in rows:
        pc = M.GMGPreconditioner(op, precision=prec, n_cells_coarse=4, coarse_degree=2, smoother_degree=3,
                                 pre_smoother_degree=pre)
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
        row = {"ms": round(best * 1e3, 3), "iterations": its, "final_rel_residual": rel,
               "precond_setup_s": round(setup, 3)}
        if x_ref is None:
            x_ref = x
        else:
            row["rel_error_vs_fp64"] = float(torch.linalg.norm(x - x_ref) / torch.linalg.norm(x_ref))
        out[name] = row
        del pc, r
    base = out["fp64_cg_gmg64_v33"]["ms"]
    out["speedup_vs_fp64_cg_gmg64_v33"] = {n: round(base / out[n]["ms"], 3) for n, *_ in rows[1:]}
    return out


def scan_timings(torch, M, k, N, n_train, n_eval, epochs, rho, only="all", law=False, coarse=4, smoothing=0):
    """Time-to-solve per step of a laser scan (see the module docstring)."""
    import numpy as np
    from meltpooldg_amd.application import LaserScan, generate_trajectory, train
    dt = 1e-3
    kw = dict(n_cells_coarse=coarse)
    if law:
        kw["law"] = (M.LAW_TANH, (1.0, 0.5, 1.0, 0.1))
    T_init = 0.95 if law else None
    wanted = None if only == "all" else set(only.split(","))
    # trajectories with the base guesses the networks correct: linear and order-3 extrapolation
    nets, training, Ts = {}, {}, None
    t_traj = 0.0
    for order in (1, 3):
        name = "fcg_gmg32_learned_on_" + ("linear" if order == 1 else "extrapolation3")
        if wanted is not None and name not in wanted and Ts is not None:
            continue
        t0 = time.perf_counter()
        _, Ts_o, qs, zs, _, bs = generate_trajectory(
            k, N, n_train, T_init=T_init, dt=dt, precond="gmg32", pre_smoother_degree=1,
            predictor=M.PREDICTOR_LINEAR if order == 1 else M.PREDICTOR_EXTRAPOLATION, extrapolation_order=order,
            with_base=True, **kw)
        torch.cuda.synchronize()
        t_traj += time.perf_counter() - t0
        if Ts is None:
            Ts = Ts_o  # the replayed steps start from the first trajectory's states
        if wanted is None or name in wanted:
            net = M.LearnedPredictor((N, N, N), degree=k, coarsening=rho, qscale=dt * N, seed=11)
            t0 = time.perf_counter()
            losses = train(net, Ts_o, qs, zs, epochs=epochs, lr=2e-3, bs=bs, lr_decay=0.7)
            torch.cuda.synchronize()
            nets[order] = net
            training[name] = {"pairs": n_train - 2, "epochs": epochs, "lr": 2e-3, "lr_decay": 0.7, "coarsening": rho,
                              "guess_pause": list(PAUSE), "seconds": round(time.perf_counter() - t0, 2), "loss_first_epoch": losses[0],
                              "loss_last_epoch": losses[-1]}
        if Ts_o is not Ts:
            del Ts_o
        del qs, zs, bs
    out = {"workload": "laser scan, Q%d %d^3 cells, %s, dt=1e-3, steps %d-%d, ||b-JT|| <= 1e-10 ||b||, coarse Q2 "
                       "%d^3, multigrid rebuilt every step (_update4: every 4th step); FP64 multigrid V(3,3) in CG, "
                       "or V(1,3) in flexible CG; FP32 multigrid V(1,3) in flexible CG" % (
                           k, N, "k = 1 + 0.5 tanh((T - 1)/0.1) from T (frozen at T_n), T0 = 0.95" if law else "k=1",
                           n_train + 1, n_train + n_eval, coarse),
           "trajectory_seconds": round(t_traj, 2), "training": training}
    variants = (("fp64_cg_gmg64_prev_step", "gmg64", M.PREDICTOR_NONE, 1, 0),
                ("fp64_cg_gmg64_extrapolation6", "gmg64", M.PREDICTOR_EXTRAPOLATION, 6, 0),
                ("fp64_fcg_gmg64_extrapolation6", "gmg64_fcg", M.PREDICTOR_EXTRAPOLATION, 6, 1),
                ("fcg_gmg32_linear", "gmg32", M.PREDICTOR_LINEAR, 1, 1),
                ("fcg_gmg32_extrapolation3", "gmg32", M.PREDICTOR_EXTRAPOLATION, 3, 1),
                ("fcg_gmg32_extrapolation6", "gmg32", M.PREDICTOR_EXTRAPOLATION, 6, 1),
                ("fcg_gmg32_learned_on_linear", "gmg32", M.PREDICTOR_LEARNED, 1, 1),
                ("fcg_gmg32_learned_on_extrapolation3", "gmg32", M.PREDICTOR_LEARNED, 3, 1),
                ("fcg_gmg32_extrapolation6_update4", "gmg32", M.PREDICTOR_EXTRAPOLATION, 6, 1))
    variants = tuple(v for v in variants if (wanted is None or v[0] in wanted or v[0] == "fp64_cg_gmg64_prev_step")
                     and (v[2] != M.PREDICTOR_LEARNED or v[3] in nets))
    T_ref = []
    for name, pc, pred, order, pre in variants:
        # _update4: the multigrid is rebuilt every 4th step only (any fixed SPD preconditioner
        # keeps CG exact; the FP64 residual test decides)
        # the learned predictor pauses for PAUSE[1] steps after PAUSE[0] consecutive rejections
        case = LaserScan(k, N, dt=dt, precond=pc, predictor=pred,
                         learned=nets.get(order) if pred == M.PREDICTOR_LEARNED else None,
                         pre_smoother_degree=pre, extrapolation_order=order, guess_smoothing=smoothing,
                         preconditioner_update_interval=4 if name.endswith("_update4") else 1,
                         guess_pause_after=PAUSE[0], guess_pause_steps=PAUSE[1], **kw)
        # restart from the recorded state; the replayed steps rebuild the history the predictor uses
        case.t = (n_train - order) * dt
        case.heat.set_initial_condition(Ts[n_train - order])
        for _ in range(order):
            case.step()
        ms, its, r0s, rs, fbs, errs = [], [], [], [], [], []
        for j in range(n_eval):
            it, el, (fb, r0, r) = case.step()
            ms.append(el * 1e3)
            its.append(it)
            r0s.append(r0)
            rs.append(r)
            fbs.append(fb)
            T = case.solution()
            if name == "fp64_cg_gmg64_prev_step":
                T_ref.append(T.clone())
            else:
                errs.append(float(torch.linalg.norm(T - T_ref[j]) / torch.linalg.norm(T_ref[j])))
        row = {"ms_per_step": round(float(np.mean(ms)), 3), "iterations": float(np.mean(its)),
               "log10_initial_rel_residual": round(float(np.mean(np.log10(r0s))), 3),
               "fallbacks": fbs.count(M.GUESS_REJECTED), "paused": fbs.count(M.GUESS_PAUSED),
               "final_rel_residual_max": float(max(rs))}
        if errs:
            row["rel_error_vs_fp64_max"] = float(max(errs))
        out[name] = row
        del case
        torch.cuda.empty_cache()
    del T_ref
    base = out["fp64_cg_gmg64_prev_step"]["ms_per_step"]
    out["speedup_vs_fp64_prev_step"] = {n: round(base / out[n]["ms_per_step"], 3) for n, *_ in variants[1:]}
    gains = {}
    for order, bname in ((1, "fcg_gmg32_linear"), (3, "fcg_gmg32_extrapolation3")):
        lname = "fcg_gmg32_learned_on_" + ("linear" if order == 1 else "extrapolation3")
        if lname in out and bname in out:
            gains[lname] = {"vs": bname, "time_ratio": round(out[bname]["ms_per_step"] / out[lname]["ms_per_step"], 3),
                            "iterations": [out[bname]["iterations"], out[lname]["iterations"]],
                            "log10_initial_rel_residual": [out[bname]["log10_initial_rel_residual"],
                                                           out[lname]["log10_initial_rel_residual"]]}
    out["learned_vs_its_base"] = gains
    learned = [n for n in out["speedup_vs_fp64_prev_step"] if "learned" in n]
    out["target_3x"] = {"best_mixed_precision": max((v, n) for n, v in out["speedup_vs_fp64_prev_step"].items()
                                                    if n.startswith("fcg_gmg32"))[::-1] if len(variants) > 1 else None,
                        "best_learned": max((out["speedup_vs_fp64_prev_step"][n], n) for n in learned)[::-1]
                        if learned else None}
    return out


def learned_c4(torch, M, n_steps=200, n_train=150, epochs=8, cells=64, rho=2, smoothing=2, orders=(1, 2)):
    """The learned guess on the training configuration: Q2 on 64^3 cells, 200 implicit steps of
    the laser scan (k = 1, T0 = 0, FP32 multigrid in flexible CG), for base guesses of order 1 and
    2. A network per order is trained on the pairs of steps 2..n_train of the solver's own
    trajectory (with that base), then steps n_train+1..n_steps are replayed from the same state
    with the base alone and with the network on top."""
    import numpy as np
    from meltpooldg_amd.application import LaserScan, generate_trajectory, train
    k, dt = 2, 1e-3
    out = {"workload": "laser scan, Q2 %d^3 cells (%d DoFs), k = 1, dt = 1e-3, T0 = 0, %d steps; networks trained on "
                       "steps 2-%d, evaluated on steps %d-%d; FP32 multigrid V(1,3), coarse Q2 4^3, flexible CG "
                       "to ||b-JT|| <= 1e-10 ||b||" % (cells, (2 * cells + 1) ** 3, n_steps, n_train, n_train + 1,
                                                       n_steps)}
    out["network"] = {"working_grid_coarsening": rho, "guess_smoothing": smoothing, "guess_pause": list(PAUSE)}
    for order in orders:
        pred = M.PREDICTOR_LINEAR if order == 1 else M.PREDICTOR_EXTRAPOLATION
        t0 = time.perf_counter()
        _, Ts, qs, zs, _, bs = generate_trajectory(k, cells, n_steps, dt=dt, precond="gmg32", pre_smoother_degree=1,
                                                   predictor=pred, extrapolation_order=order, with_base=True)
        torch.cuda.synchronize()
        t_traj = time.perf_counter() - t0
        net = M.LearnedPredictor((cells,) * 3, degree=k, coarsening=rho, qscale=dt * cells, seed=11)
        t0 = time.perf_counter()
        losses = train(net, Ts[:n_train + 1], qs[:n_train], zs[:n_train], epochs=epochs, lr=2e-3, bs=bs[:n_train],
                       lr_decay=0.7)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        del qs, zs, bs
        rows = {}
        for name, p in (("base", pred), ("learned", M.PREDICTOR_LEARNED)):
            case = LaserScan(k, cells, dt=dt, predictor=p, learned=net if p == M.PREDICTOR_LEARNED else None,
                             extrapolation_order=order, pre_smoother_degree=1, guess_smoothing=smoothing,
                             guess_pause_after=PAUSE[0], guess_pause_steps=PAUSE[1])
            case.t = (n_train - order) * dt
            case.heat.set_initial_condition(Ts[n_train - order])
            for _ in range(order):
                case.step()
            ms, its, r0s, fbs = [], [], [], []
            for _ in range(n_steps - n_train):
                it, el, (fb, r0, r) = case.step()
                ms.append(el * 1e3)
                its.append(it)
                r0s.append(r0)
                fbs.append(fb)
            rows[name] = {"ms_per_step": round(float(np.mean(ms)), 3), "iterations": float(np.mean(its)),
                          "log10_initial_rel_residual": round(float(np.mean(np.log10(r0s))), 3),
                          "fallbacks": fbs.count(M.GUESS_REJECTED), "paused": fbs.count(M.GUESS_PAUSED)}
            del case
        name = "base_linear" if order == 1 else "base_extrapolation2"
        out[name] = {"trajectory_seconds": round(t_traj, 2), "training_seconds": round(t_train, 2),
                     "training_pairs": n_train - 2, "epochs": epochs, "lr": 2e-3, "lr_decay": 0.7,
                     "loss_first_epoch": losses[0],
                     "loss_last_epoch": losses[-1], "without_network": rows["base"], "with_network": rows["learned"],
                     "time_ratio": round(rows["base"]["ms_per_step"] / rows["learned"]["ms_per_step"], 3)}
        del Ts, net
        torch.cuda.empty_cache()
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
    op, kq = make_operator(torch, M, k, N, kappa=a.kappa, mapped=bool(a.mapped))
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
    # algorithmic bytes per apply: src + dst, the conductivity data, the mapping nodes
    bytes_alg = 16.0 * op.n_dofs + {"qp": 8.0 * op.n_q_points, "law": 8.0 * op.n_dofs, "const": 0.0}[a.kappa]
    if a.mapped:
        bytes_alg += 24.0 * op.n_dofs
    kname = "heat_apply_kernel<%d,double,%s%s>" % (k + 1, {"qp": "KAPPA_QP", "law": "KAPPA_LAW",
                                                          "const": "KAPPA_CONST"}[a.kappa],
                                                   ",mapped" if a.mapped else "")
    if a.mapped and a.kappa != "const":
        kname = "heat_apply_kernel<%d,double,KAPPA_CONST,mapped> (conductivity folded into the metric)" % (k + 1)
    kdesc = {"qp": "k per quadrature point = k0(1+0.5 tanh((T-1)/0.1)), T~U[0,2]",
             "law": "k = k0(1+0.5 tanh((T-1)/0.1)) evaluated in the apply from a T DoF vector, T~U[0,2]",
             "const": "k = 1"}[a.kappa] + (", mesh mapped by x + 0.05 sin(2 pi y) sin(2 pi z)" if a.mapped else "")
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
        "config": {"workload": "FP64 heat operator apply, Q%d hex, %d^3 cells, %d DoFs, QGauss(%d), %s"
                               % (k, N, op.n_dofs, k + 1, kdesc),
                   "degree": k, "cells": N, "n_dofs": op.n_dofs, "n_q_points": op.n_q_points},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": PROFILED_TRAFFIC.get((k, N, a.kappa, bool(a.mapped))),
                     "kernel": kname, "kernel_avg_ms": round(kern_avg, 5),
                     "algorithmic_bytes": bytes_alg},
    }
    # the library's own streams over a DoF vector's size: the attainable HBM rates next to the
    # 8 TB/s spec
    del dst
    line["hbm"] = dict(spec_GBps=HBM_PEAK_GBS, **stream_rates(torch, M, op.n_dofs))
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
                                                 coarse=a.scan_coarse_cells, smoothing=a.scan_guess_smoothing)
        except Exception as e:  # report, do not hide the apply line
            line["time_to_solve"] = {"error": str(e)}
    if a.learned_c4:
        try:
            line["learned_q2_64"] = learned_c4(torch, M, rho=a.c4_coarsening, smoothing=a.c4_smoothing,
                                               orders=tuple(int(o) for o in a.c4_orders.split(",")))
        except Exception as e:  # report, do not hide the apply line
            line["learned_q2_64"] = {"error": str(e)}
    if a.cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1))
        v, reps, el, nd = cpu_apply(k, a.cpu_cells, threads)
        v1, reps1, el1, nd1 = cpu_apply(k, a.cpu_solve_cells, 1)
        line["cpu_baseline"] = {
            "value": round(v, 5), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle apply (same sum-factorised algorithm, OpenMP, FP64), Q%d %d^3 cells (%d DoFs), k per "
                      "quadrature point, %d applies in %.1f s" % (k, a.cpu_cells, nd, reps, el),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "one_core": {"value": round(v1, 5), "unit": "GDoF/s", "sample": "Q%d %d^3 cells (%d DoFs), %d applies in "
                                                                              "%.1f s" % (k, a.cpu_solve_cells, nd1,
                                                                                          reps1, el1)},
            "cg_gmg_solve": cpu_solve(k, a.cpu_solve_cells, threads)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
