"""GPU parity of the solver stack: CG + Jacobi, the multigrid V-cycle (FP64 and
FP32 levels), CG + GMG, and the Newton time step against the reference's
golden norm sequences."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu


def test_cg_jacobi_q2_16(cuda):
    """Q2 16^3, k = 1, dt = 1e-3, b ~ U[-1,1] (seed 1), CG + Jacobi to 1e-10 relative to ||b||."""
    torch = cuda
    import meltpooldg_amd as M
    orc, gop, _ = make_pair(torch, 2, (16, 16, 16), mode="const")
    b = np.random.default_rng(1).uniform(-1.0, 1.0, orc.n_dofs)
    x_ref, it_ref, _ = orc.cg(b, rel_tol=1e-10, precond=1, criterion=1)
    pc = M.JacobiPreconditioner(gop)
    pc.update()
    x = torch.zeros(orc.n_dofs, dtype=torch.float64, device="cuda")
    it = M.LinearSolver.solve(gop, x, torch.tensor(b, device="cuda"),
                              M.LinearSolverData(rel_tolerance=1e-10, criterion=M.CRIT_RHS), pc)
    xs = x.cpu().numpy()
    assert abs(it - it_ref) <= 1
    assert np.linalg.norm(b - orc.vmult(xs)) <= 1e-10 * np.linalg.norm(b) * 1.0001
    assert rel_err(xs, x_ref) < 1e-8


@pytest.mark.parametrize("degree,nc,mode,dmask,coarse_degree", [
    (2, (8, 8, 8), "const", 0, 0),
    (3, (8, 8, 8), "qp", 0, 2),
    (3, (12, 12, 12), "law", 0, 2),
    (2, (8, 8, 8), "cell", 0b000011, 0),
    (3, (8, 8, 8), "qp", 0, -2),  # mapped mesh, coarse Q2
    (4, (8, 8, 8), "qp", 0, 2),   # the Q4 hierarchy of the Q4 96^3 configuration, coarse Q2
    (4, (8, 8, 8), "law", 0, 2),
])
def test_gmg_vcycle_fp64_matches_oracle(cuda, degree, nc, mode, dmask, coarse_degree):
    torch = cuda
    import meltpooldg_amd as M
    mapped = 0.05 if coarse_degree < 0 else None
    coarse_degree = abs(coarse_degree)
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, dirichlet=dmask, seed=4, mapped=mapped)
    og = O.GMG(orc, nc_coarse=2, coarse_degree=coarse_degree, cheb_degree=3, smoothing_range=20.0, eig_iters=20)
    pc = M.GMGPreconditioner(gop, precision=M.FP64, n_cells_coarse=2, coarse_degree=coarse_degree,
                             smoother_degree=3, smoothing_range=20.0, eig_cg_n_iterations=20)
    pc.update()
    lv_g, lv_o = pc.levels(), og.levels()
    assert [l[:2] for l in lv_g] == [l[:2] for l in lv_o]
    for (_, _, lg), (_, _, lo) in zip(lv_g[1:], lv_o[1:]):
        assert abs(lg - lo) <= 1e-9 * lo
    b = np.random.default_rng(9).uniform(-1, 1, orc.n_dofs)
    ref = og.vcycle(b)
    out = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    pc.vmult(out, torch.tensor(b, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(out.cpu().numpy(), ref) < 1e-9


@pytest.mark.parametrize("degree,nc,mode,mapped", [(3, (8, 8, 8), "qp", None), (4, (8, 8, 8), "qp", None),
                                                   (4, (8, 8, 8), "law", None), (3, (8, 8, 8), "const", 0.05),
                                                   (3, (12, 12, 12), "qp", 0.05)])
def test_gmg_vcycle_fp32(cuda, degree, nc, mode, mapped):
    """FP32 level operators, smoothers and transfers (FP64 coarse solve) against the FP64 oracle
    V-cycle: Cartesian Q3 and Q4 hierarchies and the mapped Q3 one, coarse Q2."""
    torch = cuda
    import meltpooldg_amd as M
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, seed=4, mapped=mapped)
    og = O.GMG(orc, nc_coarse=2, coarse_degree=2)
    pc = M.GMGPreconditioner(gop, precision=M.FP32, n_cells_coarse=2, coarse_degree=2)
    pc.update()
    b = np.random.default_rng(9).uniform(-1, 1, orc.n_dofs)
    out = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    pc.vmult(out, torch.tensor(b, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(out.cpu().numpy(), og.vcycle(b)) < 1e-4


@pytest.mark.parametrize("precision,degree,nc,mapped", [("FP32", 3, (16, 16, 12), None), ("FP64", 3, (16, 16, 12), None),
                                                     ("FP32", 2, (16, 12, 8), None), ("FP32", 3, (12, 12, 12), 0.05)])
def test_gmg_vcycle_graph_replays_the_eager_cycle(cuda, precision, degree, nc, mapped):
    """The V-cycle replayed from its captured HIP graph gives the bits of the cycle launched
    kernel by kernel: on several replays with other work of the operator and of the eager
    preconditioner between them, and again after a new conductivity and update() (the graph is
    captured anew)."""
    torch = cuda
    import meltpooldg_amd as M
    orc, gop, data = make_pair(torch, degree, nc, mode="qp", seed=31, dirichlet=0b000011 if mapped is None else 0,
                               mapped=mapped)
    rng = np.random.default_rng(32)
    kw = dict(precision=getattr(M, precision), n_cells_coarse=4, coarse_degree=2, pre_smoother_degree=1)
    pcg = M.GMGPreconditioner(gop, vcycle_graph=True, **kw)
    pce = M.GMGPreconditioner(gop, vcycle_graph=False, **kw)
    y = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    for rebuild in range(2):
        if rebuild:
            # a new conductivity: both preconditioners are rebuilt, the graph is captured again
            data["k"].copy_(torch.tensor(rng.uniform(0.5, 1.5, data["k"].numel()), device="cuda"))
        pcg.update()
        pce.update()
        for rep in range(4):
            b = torch.tensor(rng.uniform(-1, 1, orc.n_dofs), device="cuda")
            og, oe = torch.empty_like(b), torch.empty_like(b)
            pcg.vmult(og, b)
            gop.vmult(y, og)  # other work on the stream between replays
            pce.vmult(oe, b)
            torch.cuda.synchronize()
            assert torch.isfinite(og).all()
            assert torch.equal(og, oe), (rebuild, rep)


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_gmg_vcycle_is_bitwise_repeatable(cuda, precision):
    """Two multigrid preconditioners built on the same operator, each updated twice, give the same
    V-cycle bits (diagonals, eigenvalue estimates, coarse matrix and its inverse, smoothers and
    transfers have no order-dependent sums)."""
    torch = cuda
    import meltpooldg_amd as M
    orc, gop, _ = make_pair(torch, 3, (16, 16, 12), mode="qp", seed=21, dirichlet=0b000011)
    b = torch.tensor(np.random.default_rng(22).uniform(-1, 1, orc.n_dofs), device="cuda")
    outs = []
    for _ in range(2):
        pc = M.GMGPreconditioner(gop, precision=getattr(M, precision), n_cells_coarse=4, coarse_degree=2)
        for _ in range(2):
            pc.update()
            out = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
            pc.vmult(out, b)
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy().copy())
        del pc
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("prec,solver,degree,nc,mode,mapped", [
    (0, 0, 3, (16, 16, 16), "qp", None), (1, 1, 3, (16, 16, 16), "qp", None),
    (1, 1, 4, (12, 12, 12), "qp", None), (0, 0, 4, (12, 12, 12), "law", None),   # Q4 (the Q4 96^3 solve)
    (1, 1, 3, (12, 12, 12), "const", 0.05), (1, 1, 3, (16, 16, 16), "qp", 0.05),  # mapped Q3 (Q3 96^3 mapped)
])
def test_cg_gmg_solve(cuda, prec, solver, degree, nc, mode, mapped):
    """CG + GMG (FP64 levels) and flexible CG + FP32 GMG reach 1e-10 relative residual in about the
    oracle's FP64 iteration count, with a coarse Q2 level."""
    torch = cuda
    import meltpooldg_amd as M
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, seed=2, mapped=mapped)
    og = O.GMG(orc, nc_coarse=2, coarse_degree=2)
    b = np.random.default_rng(4).uniform(-1, 1, orc.n_dofs)
    _, it_ref, _ = orc.cg(b, rel_tol=1e-10, precond=2, gmg=og, criterion=1)
    pc = M.GMGPreconditioner(gop, precision=prec, n_cells_coarse=2, coarse_degree=2)
    pc.update()
    x = torch.zeros(orc.n_dofs, dtype=torch.float64, device="cuda")
    it = M.LinearSolver.solve(gop, x, torch.tensor(b, device="cuda"),
                              M.LinearSolverData(solver_type=solver, rel_tolerance=1e-10, criterion=M.CRIT_RHS,
                                                 max_iterations=200), pc)
    xs = x.cpu().numpy()
    assert np.linalg.norm(b - orc.vmult(xs)) <= 1.0001e-10 * np.linalg.norm(b)
    assert it <= it_ref + (0 if prec == 0 else 2)


GOLDEN = {
    # applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests/*.mpirun=4.output
    "plain": [5.64198e-03, 5.53556e-03, 5.44694e-03, 5.37273e-03, 5.31056e-03],
    "conv": [1.00000e+02, 9.99999e+01, 9.99999e+01, 9.99998e+01, 9.99998e+01],
    "rad": [9.99998e+01, 9.99997e+01, 9.99995e+01, 9.99994e+01, 9.99992e+01],
    "conv_rad": [9.99998e+01, 9.99996e+01, 9.99994e+01, 9.99992e+01, 9.99990e+01],
}


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_heat_steps_golden_norms(cuda, case):
    """2D Q1 32x32 on [0,0.1]^2 run as a one-cell-thick 3D mesh (the z direction factorises out)."""
    torch = cuda
    import meltpooldg_amd as M
    nc, L = (32, 32, 1), (0.1, 0.1, 1.0)
    if case == "plain":
        rho_cp, k, dt = 1.0, 1.0, 1e-4
    else:
        rho_cp, k, dt = 7850.0 * 460.0, 55.563, 1e-2
    op = M.HeatDiffuseMultiPhaseOperator(1, nc, L, rho_cp=rho_cp)
    op.set_conductivity_constant(k)
    X = O.Operator(1, nc, L).node_coords()
    T0 = X[0].copy() if case == "plain" else np.full(X.shape[1], 1000.0)
    pc = M.JacobiPreconditioner(op)
    heat = M.HeatDiffuseOperation(op, pc, predictor=M.PREDICTOR_LINEAR,
                                  linear=M.LinearSolverData(rel_tolerance=1e-12),
                                  nlsolve=M.NonlinearSolverData(residual_tolerance=5e-9))
    if case != "plain":
        alpha = 25.0 if "conv" in case else 0.0
        eps = 0.98 if "rad" in case else 0.0
        # the temperature the radiation term is linearised about is re-bound each step by the operation
        placeholder = torch.tensor(T0, device="cuda")
        op.set_robin(2, alpha=alpha, T_inf_conv=300.0, emissivity=eps, T_inf_rad=300.0, T=placeholder)
    heat.set_initial_condition(torch.tensor(T0, device="cuda"))
    norms = []
    for _ in range(5):
        heat.solve(dt)
        norms.append(op.l2_norm(heat.solution()))
    assert ["%.5e" % v for v in norms] == ["%.5e" % v for v in GOLDEN[case]]


@pytest.mark.parametrize("precond", ["gmg32", "gmg64"])
def test_linear_step_with_conductivity_law(cuda, precond):
    """Linear laser-scan steps with k(T) frozen at T_n: each step solves
    (c_m M + K_{k(T_n)}) T_{n+1} = c_m M T_n + flux, checked with an independent operator that
    evaluates the law at every quadrature point inside the apply."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    from tests.helpers import tanh_law_params
    k, cells, dt = 3, 6, 2e-3
    law = (M.LAW_TANH, tanh_law_params())
    case = LaserScan(k, cells, dt=dt, law=law, precond=precond, predictor=M.PREDICTOR_LINEAR, rel_tol=1e-11,
                     n_cells_coarse=3)
    X = O.Operator(k, (cells,) * 3, (1.0, 1.0, 1.0)).node_coords()
    T0 = 1.0 + 0.4 * np.sin(2 * np.pi * X[0]) * np.cos(np.pi * X[1]) + 0.1 * X[2]
    case.heat.set_initial_condition(torch.tensor(T0, device="cuda"))
    ref = M.HeatDiffuseMultiPhaseOperator(k, (cells,) * 3, (1.0, 1.0, 1.0), rho_cp=1.0)
    ref.reset_time_increment(dt)
    for _ in range(3):
        case.step()
        T_old, T = case.heat.old_solution().clone(), case.solution().clone()
        ref.set_conductivity_law(law[0], law[1], T_old)
        ref.set_surface_flux(case.peak, case.sigma, case.x0, case.y0, case.speed, case.t)
        b = torch.zeros_like(T)
        ref.create_rhs(b, torch.zeros_like(T), T_old)
        r = torch.empty_like(T)
        ref.vmult(r, T)
        torch.cuda.synchronize()
        assert float(torch.linalg.norm(b - r)) <= 1e-9 * float(torch.linalg.norm(b))
    # the conductivity varies over the field, so a constant-k solve would not pass the check
    kq = np.array([O.kappa_law(law[0], law[1], t)[0] for t in T_old.cpu().numpy()[::7]])
    assert kq.max() - kq.min() > 0.3


@pytest.mark.parametrize("order", [2, 3])
def test_extrapolation_predictor(cuda, order):
    """The step's initial residual is that of the Lagrange extrapolation (order 2, 3) of the recorded
    solutions to the new time, with variable step sizes."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    case = LaserScan(2, 8, dt=1e-3, precond="gmg64", predictor=M.PREDICTOR_EXTRAPOLATION, extrapolation_order=order,
                     n_cells_coarse=2)
    Ts, ts = [case.solution()], [0.0]
    for dt in (1e-3, 1e-3, 1.5e-3, 1e-3, 0.8e-3):
        case.dt = dt
        case.step()
        Ts.append(case.solution())
        ts.append(case.t)
    dt = 1.2e-3
    t_new = ts[-1] + dt
    pts, tau = Ts[-1 - order:][::-1], [t - ts[-1] for t in ts[-1 - order:][::-1]]
    w = [np.prod([(dt - tau[i]) / (tau[j] - tau[i]) for i in range(order + 1) if i != j]) for j in range(order + 1)]
    guess = sum(wj * T for wj, T in zip(w, pts))
    case.dt = dt
    _, _, (_, r0, _) = case.step()
    assert abs(case.t - t_new) < 1e-15
    b = torch.zeros_like(guess)
    case.op.create_rhs(b, torch.zeros_like(guess), Ts[-1])
    Jg = torch.empty_like(guess)
    case.op.vmult(Jg, guess)
    expect = float(torch.linalg.norm(b - Jg) / torch.linalg.norm(b))
    assert abs(r0 - expect) <= 1e-6 * expect


@pytest.mark.parametrize("law", [False, True])
def test_least_squares_projection_predictor(cuda, law):
    """The step's own right-hand side projected on the last 4 solutions (PREDICTOR_LSQ_SOLUTION): the
    step starts from the combination with the smallest residual ||b - J x|| (numpy's dense least squares
    on the same vectors), which is no worse than cubic extrapolation through the same solutions. With a
    conductivity law the projection uses the step's operator J(T_n)."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    from tests.helpers import tanh_law_params
    kw = dict(law=(M.LAW_TANH, tanh_law_params())) if law else {}
    case = LaserScan(2, 8, dt=1e-3, precond="gmg64", predictor=M.PREDICTOR_LSQ_SOLUTION, n_old_solution_vectors=4,
                     n_cells_coarse=2, **kw)
    if law:
        case.heat.set_initial_condition(torch.full_like(case.T0, 0.95))
    Ts = [case.solution()]
    for _ in range(6):
        case.step()
        Ts.append(case.solution())
    _, _, (_, r0, _) = case.step()
    b = torch.zeros_like(Ts[-1])
    case.op.create_rhs(b, torch.zeros_like(b), Ts[-1])
    hist = Ts[-4:][::-1]
    A = []
    for T in hist:
        a = torch.empty_like(T)
        case.op.vmult(a, T)
        A.append(a.cpu().numpy())
    A, bn = np.stack(A, 1), b.cpu().numpy()
    c, *_ = np.linalg.lstsq(A, bn, rcond=None)
    r_min = np.linalg.norm(bn - A @ c) / np.linalg.norm(bn)
    cubic = np.array([4.0, -6.0, 4.0, -1.0])
    r_cubic = np.linalg.norm(bn - A @ cubic) / np.linalg.norm(bn)
    assert r0 <= r_cubic * (1 + 1e-6)
    assert abs(r0 - r_min) <= 1e-3 * r_min


def test_preconditioner_update_interval(cuda):
    """Linear steps with a conductivity law: rebuilding the multigrid only every third step still
    solves every step to the FP64 tolerance, so the trajectory equals the every-step one."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    from tests.helpers import tanh_law_params
    out = []
    for every in (1, 3):
        case = LaserScan(2, 8, dt=2e-3, law=(M.LAW_TANH, tanh_law_params()), precond="gmg32",
                         predictor=M.PREDICTOR_EXTRAPOLATION, extrapolation_order=3, rel_tol=1e-12, n_cells_coarse=2,
                         preconditioner_update_interval=every)
        case.heat.set_initial_condition(torch.full_like(case.T0, 0.95))
        for _ in range(7):
            case.step()
        out.append(case.solution())
    assert float(torch.linalg.norm(out[0] - out[1]) / torch.linalg.norm(out[0])) < 1e-9


def _mgs_projection(A, rhs):
    """numpy restatement of LeastSquaresProjection::vmult (include/meltpooldg/linear_algebra/predictor.hpp:36-77):
    modified Gram-Schmidt on the images without normalisation, stop at P(s,s) < 1e-16 P(0,0), back
    substitution."""
    n_old = A.shape[1]
    V = [None] * n_old
    P = np.zeros((n_old, n_old))
    step = 0
    while step < n_old:
        v = A[:, step].copy()
        P[0, step] = v @ (v if step == 0 else V[0])
        for j in range(step):
            v += -P[j, step] / P[j, j] * V[j]
            P[j + 1, step] = v @ (v if j + 1 == step else V[j + 1])
        V[step] = v
        if P[step, step] < 1e-16 * P[0, 0]:
            break
        step += 1
    c = np.array([V[s] @ rhs for s in range(step)])
    for s in range(step - 1, -1, -1):
        c[s] = (c[s] - sum(c[j] * P[s, j] for j in range(s + 1, step))) / P[s, s]
    return c


def test_least_squares_projection_reference_mode(cuda):
    """least_squares_projection as the reference: right-hand side create_rhs at T = T_old = T_n
    (heat_diffuse_operation.cpp:233-244), images of T_n, T_{n-1}, T_{n-2} orthogonalised by modified
    Gram-Schmidt (the numpy restatement equals numpy's least squares, which fixes it), the previous
    solution for the first three steps (predictor.hpp:147-151). The step's initial residual is that of
    the restated guess."""
    torch = cuda
    import meltpooldg_amd as M
    from meltpooldg_amd.application import LaserScan
    L = 3
    case = LaserScan(2, 8, dt=1e-3, precond="gmg64", predictor=M.PREDICTOR_LSQ, n_old_solution_vectors=L,
                     n_cells_coarse=2)

    def step_rhs_and_images(Tn, hist):
        b = torch.zeros_like(Tn)
        case.op.create_rhs(b, torch.zeros_like(Tn), Tn)
        A = []
        for T in hist:
            a = torch.empty_like(T)
            case.op.vmult(a, T)
            A.append(a.cpu().numpy())
        return b.cpu().numpy(), np.stack(A, 1)

    Ts = [case.solution()]
    for i in range(6):
        _, _, (_, r0, _) = case.step()
        if i < L:  # warm-up: the previous solution is the guess
            bn, A = step_rhs_and_images(Ts[-1], [Ts[-1]])
            expect = np.linalg.norm(bn - A[:, 0]) / np.linalg.norm(bn)
            assert abs(r0 - expect) <= 1e-6 * expect
        Ts.append(case.solution())
    Tn = Ts[-1]
    # the laser at the next step's time, where the step evaluates the predictor's right-hand side
    case.set_source(case.t + case.dt)
    r = torch.zeros_like(Tn)
    case.op.create_rhs(r, Tn, Tn)
    hist = Ts[-L:][::-1]
    bn, A = step_rhs_and_images(Tn, hist)
    rn = r.cpu().numpy()
    c = _mgs_projection(A, rn)
    c_ls, *_ = np.linalg.lstsq(A, rn, rcond=None)
    assert np.allclose(c, c_ls, rtol=1e-6, atol=1e-9 * np.abs(c_ls).max())
    guess = sum(ci * T.cpu().numpy() for ci, T in zip(c, hist))
    Jg = torch.empty_like(Tn)
    case.op.vmult(Jg, torch.tensor(guess, device="cuda"))
    expect = np.linalg.norm(bn - Jg.cpu().numpy()) / np.linalg.norm(bn)
    _, _, (_, r0, _) = case.step()
    assert abs(r0 - expect) <= 1e-6 * expect
