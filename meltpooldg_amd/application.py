"""Laser-scan heat transfer time loop on the MI355X operator path.

Mirrors the time loop of HeatTransferApplication::run
(applications/mp-heat-transfer/heat_transfer_application.cpp:46-140): advance
the time, move the laser (source/heat/laser_operation.cpp:245-268), evaluate the
heat source, solve the implicit step. The laser is a Gaussian surface flux on
the top face z = L2, q = peak exp(-((x - x0 - v t)^2 + (y - y0)^2)/(2 sigma^2))
(the reference's Gauss profile exp(-2 d^2/r^2) with r = 2 sigma,
include/meltpooldg/heat/laser_intensity_profiles.hpp:134-139). Every step is one
linear implicit solve with the conductivity frozen at T_n
(MPDG_STEP_LINEAR), stopped at ||b - J T|| <= rel_tol ||b||.
"""
import time

import meltpooldg_amd as M


class LaserScan:
    def __init__(self, degree, cells, dt=1e-3, rho_cp=1.0, k0=1.0, law=None, peak=1e3, sigma=0.05, speed=1.0,
                 x0=0.3, y0=0.5, precond="gmg32", predictor=M.PREDICTOR_NONE, learned=None, rel_tol=1e-10,
                 n_cells_coarse=4, coarse_degree=2, guess_check=True, smoother_degree=3, pre_smoother_degree=0,
                 guess_smoothing=0, extrapolation_order=2, n_old_solution_vectors=2,
                 preconditioner_update_interval=1, guess_pause_after=0, guess_pause_steps=0, vcycle_graph=True):
        import torch
        self.torch = torch
        self.dt, self.peak, self.sigma, self.speed, self.x0, self.y0 = dt, peak, sigma, speed, x0, y0
        nc = (cells, cells, cells) if isinstance(cells, int) else tuple(cells)
        self.op = M.HeatDiffuseMultiPhaseOperator(degree, nc, (1.0, 1.0, 1.0), rho_cp=rho_cp)
        self.op.reset_time_increment(dt)
        self.T0 = torch.zeros(self.op.n_dofs, dtype=torch.float64, device="cuda")
        if law is not None:
            # k(T) from the temperature; the operation re-binds it to T_n every step
            self.op.set_conductivity_law(law[0], law[1], self.T0)
        else:
            self.op.set_conductivity_constant(k0)
        if precond == "gmg32":
            self.pc = M.GMGPreconditioner(self.op, precision=M.FP32, n_cells_coarse=n_cells_coarse,
                                          coarse_degree=coarse_degree, smoother_degree=smoother_degree,
                                          pre_smoother_degree=pre_smoother_degree, vcycle_graph=vcycle_graph)
            solver = M.SOLVER_FCG
        elif precond in ("gmg64", "gmg64_fcg"):
            # FP64 multigrid inside CG, or inside flexible CG (for non-symmetric cycles, e.g. fewer
            # pre- than post-smoothing steps)
            self.pc = M.GMGPreconditioner(self.op, precision=M.FP64, n_cells_coarse=n_cells_coarse,
                                          coarse_degree=coarse_degree, smoother_degree=smoother_degree,
                                          pre_smoother_degree=pre_smoother_degree, vcycle_graph=vcycle_graph)
            solver = M.SOLVER_CG if precond == "gmg64" else M.SOLVER_FCG
        else:
            self.pc = M.JacobiPreconditioner(self.op)
            solver = M.SOLVER_CG
        self.q = torch.zeros_like(self.T0)
        self.heat = M.HeatDiffuseOperation(
            self.op, self.pc, predictor=predictor,
            linear=M.LinearSolverData(solver_type=solver, rel_tolerance=rel_tol, criterion=M.CRIT_RHS,
                                      max_iterations=2000),
            step_mode=M.STEP_LINEAR, learned=learned, q_field=self.q, guess_check=guess_check,
            guess_smoothing=guess_smoothing, extrapolation_order=extrapolation_order,
            n_old_solution_vectors=n_old_solution_vectors,
            preconditioner_update_interval=preconditioner_update_interval,
            guess_pause_after=guess_pause_after, guess_pause_steps=guess_pause_steps)
        self.heat.set_initial_condition(self.T0)
        self.t = 0.0
        self.dims = tuple(k * degree + 1 for k in nc)

    def set_source(self, t):
        """Laser position at time t: surface flux for create_rhs and the nodal field for the network."""
        self.op.set_surface_flux(self.peak, self.sigma, self.x0, self.y0, self.speed, t)
        self.op.surface_flux_field(self.q)

    def step(self):
        """One time step; returns (linear iterations, seconds from predictor to convergence, info)."""
        torch = self.torch
        self.t += self.dt
        self.set_source(self.t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its, _ = self.heat.solve(self.dt)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        return its, el, self.heat.last_step_info()

    def solution(self):
        return self.heat.solution()


def generate_trajectory(degree, cells, n_steps, T_init=None, with_base=False, **kw):
    """Run n_steps and keep every T_n (T_0 = 0, or the uniform T_init, included), the source field
    of each step and the network's residual input of each step: qs[n], zs[n] belong to the step
    T_n -> T_{n+1}. With with_base, also the base guess of each step (the predictor's
    extrapolation, bs[n]) as a sixth return value."""
    case = LaserScan(degree, cells, **kw)
    if T_init is not None:
        case.heat.set_initial_condition(case.torch.full_like(case.T0, T_init))
    z = case.torch.zeros_like(case.T0)
    base = case.torch.zeros_like(case.T0) if with_base else None
    case.heat.record_guess_feature(z, base)
    Ts = [case.solution()]
    qs, zs, bs = [], [], []
    stats = []
    for _ in range(n_steps):
        its, el, info = case.step()
        Ts.append(case.solution())
        qs.append(case.q.clone())
        zs.append(z.clone())
        if with_base:
            bs.append(base.clone())
        stats.append((its, el, info))
    case.heat.record_guess_feature(None)
    if with_base:
        return case, Ts, qs, zs, stats, bs
    return case, Ts, qs, zs, stats


def train(net, Ts, qs, zs, epochs=4, lr=1e-3, first=2, seed=0, log=None, bs=None, lr_decay=1.0):
    """Adam over the solver's own pairs ((T_{n-1}, T_n, q_{n+1}, z_n[, T_b,n]) -> T_{n+1}),
    n >= first, in a shuffled order per epoch; bs (the base guesses) None: the linear
    extrapolation. The step size of epoch e is lr * lr_decay**e (a decay of 0.7 keeps the loss
    falling from epoch to epoch where a constant rate lets it jump back up)."""
    import random
    rng = random.Random(seed)
    idx = list(range(first, len(Ts) - 1))
    losses = []
    for e in range(epochs):
        rng.shuffle(idx)
        tot = 0.0
        lr_e = lr * lr_decay ** e
        for n in idx:
            tot += net.train_step(Ts[n - 1], Ts[n], qs[n], zs[n], Ts[n + 1], lr_e,
                                  base=None if bs is None else bs[n])
        losses.append(tot / len(idx))
        if log:
            log("epoch %d loss %.4e" % (e, losses[-1]))
    return losses
