"""Pins the CPU oracle (the checker of the HIP library) without a GPU.

Each function of the oracle is checked against something that fixes it
independently of its own formula: the reference's golden outputs, brute-force
element matrices, closed forms of the mathematics, or a textbook library
routine (scipy) on the assembled matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O


# ---------------------------------------------------------------- 1D tables
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_fe1d_tables(k):
    t = O.fe1d(k)
    n = k + 1
    # Gauss rule with n points integrates x^m exactly for m <= 2n-1 on [0,1]
    for m in range(2 * n):
        assert abs(np.dot(t["qw"], t["qp"] ** m) - 1.0 / (m + 1)) < 1e-14
    # Gauss-Lobatto nodes: end points and roots of P'_k (symmetric, interior)
    assert t["nodes"][0] == 0.0 and t["nodes"][-1] == 1.0
    assert np.allclose(t["nodes"] + t["nodes"][::-1], 1.0, atol=1e-15)
    if k >= 2:
        dP = np.polynomial.legendre.Legendre.basis(k).deriv()
        assert np.allclose(dP(2 * t["nodes"][1:-1] - 1), 0.0, atol=1e-12)
    # partition of unity and exact derivative of x^j
    assert np.allclose(t["S"].sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(t["G"].sum(axis=1), 0.0, atol=1e-12)
    for j in range(1, k + 1):
        assert np.allclose(t["G"] @ t["nodes"] ** j, j * t["qp"] ** (j - 1), atol=1e-12)
    # collocation derivative composed with S equals the shape gradient
    assert np.allclose(t["D"] @ t["S"], t["G"], atol=1e-12)


# ---------------------------------------------------- operator vs brute force
CASES = [
    dict(degree=1, nc=(3, 2, 2)),
    dict(degree=2, nc=(2, 3, 2), kappa=2.5),
    dict(degree=3, nc=(2, 2, 1), kappa_qp=True),
    dict(degree=4, nc=(1, 2, 1), kappa_cell=True),
    dict(degree=2, nc=(2, 2, 2), law="tanh"),
    dict(degree=3, nc=(1, 1, 2), law="mushy"),
    dict(degree=2, nc=(2, 2, 2), mapped=0.05),
    dict(degree=3, nc=(2, 1, 1), mapped=0.08, kappa_qp=True),
    dict(degree=2, nc=(2, 2, 2), dirichlet=0b101101),
    dict(degree=1, nc=(3, 3, 1), robin=True),
]


def build(case, L=(1.0, 0.8, 1.2), seed=0):
    rng = np.random.default_rng(seed)
    k, nc = case["degree"], case["nc"]
    n1 = k + 1
    nq = nc[0] * nc[1] * nc[2] * n1 ** 3
    kw = dict(c_mass=3.0, dirichlet_mask=case.get("dirichlet", 0))
    if case.get("kappa_qp"):
        kw["kappa_qp"] = rng.uniform(0.5, 1.5, nq)
    elif case.get("kappa_cell"):
        kw["kappa_cell"] = rng.uniform(0.5, 1.5, nc[0] * nc[1] * nc[2])
    elif case.get("law"):
        tmp = O.Operator(k, nc, L)
        kw["T"] = rng.uniform(0.0, 2.0, tmp.n_dofs)
        kw["law"] = O.LAW_TANH if case["law"] == "tanh" else O.LAW_MUSHY
        kw["law_p"] = (1.0, 0.5, 1.0, 0.1) if case["law"] == "tanh" else (40.0, 30.0, 0.5, 1.5)
    else:
        kw["kappa"] = case.get("kappa", 1.0)
    if case.get("mapped"):
        kw["mapped_amp"] = case["mapped"]
    if case.get("robin"):
        tmp = O.Operator(k, nc, L)
        kw.update(robin_mask=2 | 16, alpha=3.0, emissivity=0.5, T=rng.uniform(200, 400, tmp.n_dofs))
    return O.Operator(k, nc, L, **kw)


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_vmult_equals_assembled_element_matrices(case):
    op = build(case)
    x = np.random.default_rng(1).standard_normal(op.n_dofs)
    a, b = op.vmult(x), op.vmult_brute(x)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)


def assemble(op):
    n = op.n_dofs
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(op.vmult_brute(e))
    return np.array(cols).T


@pytest.mark.parametrize("case", CASES[:6] + CASES[8:], ids=[str(i) for i in [0, 1, 2, 3, 4, 5, 8, 9]])
def test_diagonal_and_symmetry(case):
    op = build(case)
    A = assemble(op)
    assert np.allclose(A, A.T, rtol=0, atol=1e-12 * np.abs(A).max())
    d = np.diag(A).copy()
    dm = case.get("dirichlet", 0)
    if dm:
        d[np.all(A == 0, axis=1)] = 1.0
    assert np.allclose(op.diagonal(), d, rtol=1e-13, atol=1e-14)
    inv = op.inverse_diagonal()
    assert np.allclose(inv, np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0), rtol=1e-13)
    # positive definite on the unconstrained space
    free = ~np.all(A == 0, axis=1)
    assert np.linalg.eigvalsh(A[np.ix_(free, free)]).min() > 0.0


def test_closed_forms():
    """K 1 = 0 and 1' M 1 = |Omega| (exact quadrature on affine cells), for every degree."""
    L = (0.3, 0.5, 0.7)
    for k in (1, 2, 3, 4):
        nc = (3, 2, 2)
        stiff = O.Operator(k, nc, L, c_mass=0.0, kappa_qp=np.random.default_rng(k).uniform(1, 2, 12 * (k + 1) ** 3))
        one = np.ones(stiff.n_dofs)
        assert np.abs(stiff.vmult(one)).max() < 1e-12
        mass = O.Operator(k, nc, L, c_mass=1.0, kappa=0.0)
        assert abs(one @ mass.vmult(one) - L[0] * L[1] * L[2]) < 1e-14
        # mapped mesh: the transform keeps every column's x-extent, so the volume is unchanged
        mm = O.Operator(k, nc, L, c_mass=1.0, kappa=0.0, mapped_amp=0.04)
        assert abs(one @ mm.vmult(one) - L[0] * L[1] * L[2]) < 1e-12
        assert np.abs(O.Operator(k, nc, L, c_mass=0.0, kappa=1.0, mapped_amp=0.04).vmult(one)).max() < 1e-11


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_exact_polynomial_integrals(k):
    """u' J v equals the analytic integral for u, v in Q_k on the unit-extent box."""
    L = (1.0, 2.0, 0.5)
    op = O.Operator(k, (2, 3, 2), L, c_mass=2.0, kappa=3.0)
    x, y, z = op.node_coords()
    u = x ** k + y * z
    v = x * y ** min(k, 2) + 1.0
    val = u @ op.vmult(v)
    # analytic: 2 * int u v + 3 * int grad u . grad v over [0,1]x[0,2]x[0,0.5]
    import itertools
    g = np.polynomial.legendre.leggauss(8)
    pts = [(0.5 * (g[0] + 1) * Lx, 0.5 * g[1] * Lx) for Lx in L]
    s = 0.0
    m = min(k, 2)
    for (xi, wx), (yi, wy), (zi, wz) in itertools.product(*[list(zip(*p)) for p in pts]):
        uu = xi ** k + yi * zi
        vv = xi * yi ** m + 1.0
        gu = (k * xi ** (k - 1), zi, yi)
        gv = (yi ** m, m * xi * yi ** (m - 1) if m > 0 else 0.0, 0.0)
        s += wx * wy * wz * (2.0 * uu * vv + 3.0 * (gu[0] * gv[0] + gu[1] * gv[1] + gu[2] * gv[2]))
    assert abs(val - s) < 1e-11 * abs(s)


def test_l2_norm_exact():
    for k in (1, 2, 3):
        op = O.Operator(k, (2, 2, 3), (1.0, 1.0, 1.0))
        x, y, z = op.node_coords()
        u = x + 2 * y * z
        # int (x + 2yz)^2 over the unit cube = 1/3 + 2*(1/2)(1/2)*2 ... evaluated exactly:
        exact = 1.0 / 3.0 + 4.0 * (1.0 / 2.0) * (1.0 / 4.0) + 4.0 * (1.0 / 3.0) * (1.0 / 3.0)
        assert abs(op.l2_norm(u) ** 2 - exact) < 1e-13


# ---------------------------------------------------------- material laws
def test_mushy_zone_law_matches_reference_material_output():
    """unit_tests/material.output, row liquid_solid: (T = 20 -> k 31.04, dk/dT 0.096),
    (T = 60 -> k 36.48, dk/dT 0.144) for k_l = 40, k_s = 30, T_s = 0, T_l = 100."""
    p = (40.0, 30.0, 0.0, 100.0)
    k, dk = O.kappa_law(O.LAW_MUSHY, p, 20.0)
    assert abs(k - 31.04) < 1e-12 and abs(dk - 0.096) < 1e-12
    k, dk = O.kappa_law(O.LAW_MUSHY, p, 60.0)
    assert abs(k - 36.48) < 1e-12 and abs(dk - 0.144) < 1e-12
    # the same law on the other columns of that row (rho: l 8000, s 7000 -> 7104, 7648)
    k, _ = O.kappa_law(O.LAW_MUSHY, (8000.0, 7000.0, 0.0, 100.0), 20.0)
    assert abs(k - 7104.0) < 1e-9
    assert O.kappa_law(O.LAW_MUSHY, p, -5.0)[0] == 30.0 and O.kappa_law(O.LAW_MUSHY, p, 150.0)[0] == 40.0


def test_tanh_law_closed_form():
    p = (2.0, 0.5, 1.0, 0.1)
    for T in (0.0, 0.95, 1.0, 1.3, 2.0):
        k, dk = O.kappa_law(O.LAW_TANH, p, T)
        assert abs(k - 2.0 * (1 + 0.5 * np.tanh((T - 1.0) / 0.1))) < 1e-14
        h = 1e-6
        fd = (O.kappa_law(O.LAW_TANH, p, T + h)[0] - O.kappa_law(O.LAW_TANH, p, T - h)[0]) / (2 * h)
        assert abs(dk - fd) < 1e-6


# ------------------------------------------------------------------ solvers
def test_cg_matches_scipy():
    op = build(dict(degree=2, nc=(3, 3, 2), kappa_qp=True))
    A = sp.csr_matrix(assemble(op))
    b = np.random.default_rng(3).uniform(-1, 1, op.n_dofs)
    x, it, _ = op.cg(b, rel_tol=1e-12, precond=0, criterion=1, threads=1)
    xs = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    # scipy's CG (same start, same Krylov space) needs the same number of steps
    count = [0]
    spla.cg(A, b, rtol=1e-12, atol=0.0, maxiter=10000, callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert abs(count[0] - it) <= 1


def test_transfers_exact_on_polynomials():
    """Prolongation reproduces coarse Q_k functions; restriction is its transpose."""
    rng = np.random.default_rng(0)
    for kc, ncc, kf, ncf in [(2, (2, 3, 2), 2, (4, 6, 4)), (3, (2, 2, 2), 3, (4, 4, 4)), (2, (3, 2, 2), 3, (3, 2, 2)),
                             (1, (2, 2, 1), 1, (4, 4, 2))]:
        oc = O.Operator(kc, ncc)
        of = O.Operator(kf, ncf)
        xc, yc, zc = oc.node_coords()
        xf, yf, zf = of.node_coords()
        f = lambda x, y, z: x ** min(kc, 2) + x * y - 2 * z ** kc + 0.5
        assert np.allclose(O.prolongate(kc, ncc, kf, ncf, f(xc, yc, zc)), f(xf, yf, zf), atol=1e-12)
        u, v = rng.standard_normal(oc.n_dofs), rng.standard_normal(of.n_dofs)
        assert abs(v @ O.prolongate(kc, ncc, kf, ncf, u) - u @ O.restrict(kc, ncc, kf, ncf, v)) < 1e-10


def test_chebyshev_is_the_chebyshev_polynomial():
    """One smoothing sweep from x = 0 applies p(D^-1 A) D^-1 with 1 - t p(t) the scaled Chebyshev polynomial."""
    op = build(dict(degree=2, nc=(2, 2, 2), kappa=1.0))
    A = assemble(op)
    inv = op.inverse_diagonal()
    lam = O.kappa_law  # noqa: keep namespace use explicit below
    lmax = op_lmax = float(np.linalg.eigvals(np.diag(inv) @ A).real.max()) * 1.2
    deg, rng_ = 3, 20.0
    b = np.random.default_rng(5).standard_normal(op.n_dofs)
    x = np.zeros(op.n_dofs)
    O.lib().orc_chebyshev(op.ref, O.ptr(inv), lmax, rng_, deg, O.ptr(x), O.ptr(b), 1)
    lmin = lmax / rng_
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    # error propagation E = I - p(DA)DA must equal T_deg((theta - t)/delta)/T_deg(theta/delta)
    Bm = np.diag(inv) @ A
    w, V = np.linalg.eig(Bm)
    Tn = lambda s: np.cos(deg * np.arccos(np.clip(s, -1, 1))) if abs(s) <= 1 else np.cosh(deg * np.arccosh(abs(s))) * (np.sign(s) ** deg)
    Ew = np.array([Tn((theta - t.real) / delta) / Tn(theta / delta) for t in w])
    E = (V @ np.diag(Ew) @ np.linalg.inv(V)).real
    xe = np.linalg.solve(A, b)
    assert np.allclose(xe - x, E @ xe, atol=1e-10 * np.abs(xe).max())


def test_eigenvalue_estimate():
    op = build(dict(degree=2, nc=(3, 2, 2), kappa_qp=True))
    A = assemble(op)
    inv = op.inverse_diagonal()
    exact = np.linalg.eigvals(np.diag(inv) @ A).real.max()
    est = O.lib().orc_estimate_lambda_max(op.ref, O.ptr(inv), 40) / 1.2
    assert abs(est - exact) < 1e-6 * exact


def test_two_grid_contraction():
    """The V-cycle is a good preconditioner: CG + GMG needs few iterations, independent of h."""
    its = []
    for N in (8, 16):
        op = O.Operator(2, (N, N, N), c_mass=10.0, kappa_qp=np.random.default_rng(N).uniform(0.5, 1.5, N ** 3 * 27))
        g = O.GMG(op, nc_coarse=2)
        b = np.random.default_rng(1).uniform(-1, 1, op.n_dofs)
        _, it, _ = op.cg(b, rel_tol=1e-10, precond=2, gmg=g, criterion=1)
        its.append(it)
    assert max(its) <= 15 and abs(its[1] - its[0]) <= 3


# -------------------------------------------------------- golden norm runs
GOLDEN = {
    "plain": [5.64198e-03, 5.53556e-03, 5.44694e-03, 5.37273e-03, 5.31056e-03],
    "conv": [1.00000e+02, 9.99999e+01, 9.99999e+01, 9.99998e+01, 9.99998e+01],
    "rad": [9.99998e+01, 9.99997e+01, 9.99995e+01, 9.99994e+01, 9.99992e+01],
    "conv_rad": [9.99998e+01, 9.99996e+01, 9.99994e+01, 9.99992e+01, 9.99990e+01],
}


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_reference_golden_norm_sequences(case):
    """applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests/
    unidirectional_heat_transfer{,_with_convection,_with_radiation,_with_radiation_and_convection}.mpirun=4.output:
    2D Q1 on [0,0.1]^2 refined 5 times, run here as 32x32x1 cells of thickness 1 (the 3D
    problem restricted to z-constant fields is exactly the 2D one)."""
    nc, L = (32, 32, 1), (0.1, 0.1, 1.0)
    if case == "plain":
        op = O.Operator(1, nc, L, kappa=1.0)
        T0, rc, dt = op.node_coords()[0].copy(), 1.0, 1e-4
    else:
        kw = dict(robin_mask=2, T_inf_conv=300.0)
        if "conv" in case:
            kw["alpha"] = 25.0
        if "rad" in case:
            kw.update(emissivity=0.98, T_inf_rad=300.0)
        op = O.Operator(1, nc, L, kappa=55.563, **kw)
        T0, rc, dt = np.full(op.n_dofs, 1000.0), 7850.0 * 460.0, 1e-2
    rc_, _, norms, _ = O.heat_run(op, rc, dt, 5, T0, predictor=2, nl_res_tol=5e-9)
    assert rc_ == 0
    assert ["%.5e" % v for v in norms] == ["%.5e" % v for v in GOLDEN[case]]


def test_planted_mistake_is_caught():
    """A dropped stiffness term (c_mass only) or a wrong sign must fail the golden sequence."""
    nc, L = (32, 32, 1), (0.1, 0.1, 1.0)
    op = O.Operator(1, nc, L, kappa=0.5)  # wrong conductivity
    T0 = op.node_coords()[0].copy()
    _, _, norms, _ = O.heat_run(op, 1.0, 1e-4, 5, T0, predictor=2, nl_res_tol=5e-9)
    assert ["%.5e" % v for v in norms] != ["%.5e" % v for v in GOLDEN["plain"]]


# ------------------------------------------ full Newton tangent (solidification, velocity)
MUSHY = dict(k=(55.563, 17.0), cp=(460.0, 700.0), rho=(7850.0, 7850.0), T_solidus=1966.0, T_liquidus=1974.0)
TANGENT_GOLDEN = {
    # unidirectional_heat_transfer_with_velocity.mpirun=4.output
    "velocity": [1.00000e+02, 1.00000e+02, 9.99999e+01, 9.99999e+01, 9.99999e+01],
    # unidirectional_heat_transfer_with_velocity_and_solidification.mpirun=4.output
    "velocity_solidification": [1.97002e+02, 1.97003e+02, 1.97004e+02, 1.97005e+02, 1.97006e+02],
}


def tangent_case(case):
    """The reference's 2D Q1 32x32 setups on [0,0.1]^2 (cases/unidirectional_heat_transfer.cpp:194-260 and
    the two .json files), run as a one-cell-thick 3D mesh. Returns (op, T0, Newton residual tolerance)."""
    nc, L = (32, 32, 1), (0.1, 0.1, 1.0)
    if case == "velocity":
        # gas 7850 x 460, k 55.563; convection 25 and radiation 0.98 to 300 K on x+; T0 = 1000; v = -0.01
        dt, rho_cp = 1e-3, 7850.0 * 460.0
        op = O.Operator(1, nc, L, c_mass=rho_cp / dt, kappa=55.563, robin_mask=2, alpha=25.0, emissivity=0.98,
                        T_inf_conv=300.0, T_inf_rad=300.0)
        op.set_tangent(dt, velocity=(-0.01, 0.0, 0.0))
        return op, np.full(op.n_dofs, 1000.0), 1e-6
    op = O.Operator(1, nc, L)
    op.set_tangent(0.01, material=MUSHY, velocity=(-0.005, 0.0, 0.0), do_solidification=True)
    return op, 1960.0 + 20.0 * op.node_coords()[0] / 0.1, 1e-7


@pytest.mark.parametrize("case", sorted(TANGENT_GOLDEN))
def test_tangent_golden_norm_sequences(case):
    op, T0, tol = tangent_case(case)
    rc, _, norms, _ = O.heat_run_tangent(op, 5, T0, nl_res_tol=tol, threads=4)
    assert rc == 0
    assert ["%.5e" % v for v in norms] == ["%.5e" % v for v in TANGENT_GOLDEN[case]]


def test_tangent_planted_mistake_is_caught():
    """Without the velocity the solidification sequence changes in the printed digits."""
    op, T0, tol = tangent_case("velocity_solidification")
    op.s.tg.velocity[0] = 0.0
    _, _, norms, _ = O.heat_run_tangent(op, 5, T0, nl_res_tol=tol, threads=4)
    assert ["%.5e" % v for v in norms] != ["%.5e" % v for v in TANGENT_GOLDEN["velocity_solidification"]]


def test_material_blend_closed_form():
    """rho c_p and k of the mushy zone: cubic blends of the phase values, derivatives by central differences,
    pure-phase values outside [T_s, T_l]; k agrees with the conductivity law pinned by material.output."""
    op = O.Operator(1, (1, 1, 1))
    op.set_tangent(1.0, material=MUSHY)
    for T in (1967.0, 1970.0, 1973.5):
        rc, k, drc, dk = op.material(T)
        fs = (1974.0 - T) / 8.0
        blend = lambda l, s: l + (s - l) * (3 * fs ** 2 - 2 * fs ** 3)
        assert abs(k - blend(55.563, 17.0)) < 1e-12 and abs(rc - 7850.0 * blend(460.0, 700.0)) < 1e-6
        assert abs(k - O.kappa_law(O.LAW_MUSHY, (55.563, 17.0, 1966.0, 1974.0), T)[0]) < 1e-12
        h = 1e-5
        fd = (op.material(T + h) - op.material(T - h)) / (2 * h)
        assert abs(drc - fd[0]) <= 1e-6 * abs(drc) and abs(dk - fd[1]) <= 1e-6 * abs(dk)
    assert tuple(op.material(1960.0)) == (7850.0 * 700.0, 17.0, 0.0, 0.0)
    assert tuple(op.material(1980.0)) == (7850.0 * 460.0, 55.563, 0.0, 0.0)


@pytest.mark.parametrize("solid", [True, False])
def test_tangent_is_the_derivative_of_the_residual(solid):
    """J(T) u = -d/de create_rhs(T + e u) (central differences): the tangent carries every term of the
    residual's derivative; without the d/dT terms (solid = False) the check must fail."""
    rng = np.random.default_rng(5)
    op = O.Operator(2, (3, 3, 2), (0.3, 0.3, 0.2))
    X = op.node_coords()
    T = 1970.0 + 2.0 * np.sin(5 * X[0]) * np.cos(4 * X[1]) + 0.5 * X[2]  # inside the mushy interval
    T_old = T - 0.3 * np.cos(3 * X[0])
    op.set_tangent(0.01, material=MUSHY, velocity=(-0.02, 0.01, 0.005), do_solidification=solid, T_old=T_old)
    op.set_T(T)
    u = rng.uniform(-1, 1, op.n_dofs)
    Ju = op.vmult(u, threads=1)
    eps = 1e-4

    def res(Tx):
        return op.create_rhs(np.zeros(op.n_dofs), np.ascontiguousarray(Tx), T_old)

    fd = -(res(T + eps * u) - res(T - eps * u)) / (2 * eps)
    err = np.linalg.norm(Ju - fd) / np.linalg.norm(fd)
    if solid:
        assert err < 1e-7
    else:
        assert err > 1e-3


def _box_integral(L, c0, cx, cy, cz):
    """Integral of c0 + cx x + cy y + cz z over [0, Lx] x [0, Ly] x [0, Lz]."""
    V = L[0] * L[1] * L[2]
    return V * (c0 + cx * L[0] / 2 + cy * L[1] / 2 + cz * L[2] / 2)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_tangent_velocity_field_closed_forms(k):
    """Velocity DoF field v = (1 + 2y, -0.5 + 3z, 0.25 + x), u = T = 1 + x + 2y + 3z (both reproduced by Q_k):
    since sum_i phi_i = 1 and sum_i x_i phi_i = x, sum_i (J u)_i = int rc/dt u + rc v.grad u,
    sum_i x_i (J u)_i = int (rc/dt u + rc v.grad u) x + k du/dx, and with T_old = T and no source
    sum_i rhs_i = -int rc v.grad T. Components swapped, or the field ignored, change every sum."""
    L = (0.3, 0.2, 0.25)
    dt, cm, kap = 0.1, 3.0, 0.7
    op = O.Operator(k, (2, 3, 2), L, c_mass=cm, kappa=kap)
    x, y, z = op.node_coords()
    vf = np.concatenate([1 + 2 * y, -0.5 + 3 * z, 0.25 + x])
    u = 1 + x + 2 * y + 3 * z
    op.set_T(u)
    op.set_tangent(dt, velocity_field=vf)
    rc = cm * dt
    # v.grad u = 0.75 + 3x + 2y + 6z
    mass = cm * _box_integral(L, 1, 1, 2, 3)
    conv = rc * _box_integral(L, 0.75, 3, 2, 6)
    Ju = op.vmult(u, threads=1)
    assert abs(Ju.sum() - (mass + conv)) <= 1e-12 * abs(mass + conv)
    # int f x for f = c0 + cx x + cy y + cz z
    V = L[0] * L[1] * L[2]

    def xmom(c0, cx, cy, cz):
        return V * (c0 * L[0] / 2 + cx * L[0] ** 2 / 3 + cy * L[0] * L[1] / 4 + cz * L[0] * L[2] / 4)

    xm = cm * xmom(1, 1, 2, 3) + rc * xmom(0.75, 3, 2, 6) + kap * 1.0 * V
    assert abs(x @ Ju - xm) <= 1e-12 * abs(xm)
    rhs = op.create_rhs(np.zeros(op.n_dofs), np.ascontiguousarray(u), np.ascontiguousarray(u))
    assert abs(rhs.sum() + conv) <= 1e-12 * abs(conv)
    # planted mistakes: the x and y components exchanged, the field dropped
    for wrong in (np.concatenate([-0.5 + 3 * z, 1 + 2 * y, 0.25 + x]), np.zeros(3 * op.n_dofs)):
        op.set_tangent(dt, velocity_field=wrong)
        assert abs(op.vmult(u, threads=1).sum() - (mass + conv)) > 1e-3 * abs(conv)


def test_tangent_constant_velocity_field_equals_constant_vector():
    """A velocity field with the same value at every node gives the constant-vector tangent: apply,
    diagonal and residual, with the mushy material and the solidification terms."""
    op = O.Operator(2, (3, 2, 2), (0.3, 0.2, 0.2), dirichlet_mask=1)
    X = op.node_coords()
    T = 1970.0 + 3.0 * np.sin(4 * X[0]) + X[1]
    T_old = T - 0.4 * np.cos(2 * X[2])
    w = (-0.5, 0.2, 0.1)
    out = []
    for field in (None, np.concatenate([np.full(op.n_dofs, c) for c in w])):
        op.set_tangent(0.01, material=MUSHY, velocity=w if field is None else (0.0, 0.0, 0.0),
                       do_solidification=True, T_old=T_old, velocity_field=field)
        op.set_T(T)
        u = np.random.default_rng(3).uniform(-1, 1, op.n_dofs)
        out.append((op.vmult(u, threads=1), op.diagonal(),
                    op.create_rhs(np.zeros(op.n_dofs), np.ascontiguousarray(T), np.ascontiguousarray(T_old))))
    for a, b in zip(out[0], out[1]):
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


def test_tangent_with_velocity_field_is_the_derivative_of_the_residual():
    """J(T) u = -d/de create_rhs(T + e u) with a non-uniform velocity field and the solidification terms."""
    rng = np.random.default_rng(6)
    op = O.Operator(2, (3, 3, 2), (0.3, 0.3, 0.2))
    X = op.node_coords()
    T = 1970.0 + 2.0 * np.sin(5 * X[0]) * np.cos(4 * X[1]) + 0.5 * X[2]
    T_old = T - 0.3 * np.cos(3 * X[0])
    vf = np.concatenate([-0.02 + 0.1 * X[1] ** 2, 0.01 * np.sin(6 * X[0]), 0.005 + 0.05 * X[0] * X[2]])
    op.set_tangent(0.01, material=MUSHY, do_solidification=True, T_old=T_old, velocity_field=vf)
    op.set_T(T)
    u = rng.uniform(-1, 1, op.n_dofs)
    Ju = op.vmult(u, threads=1)
    eps = 1e-4

    def res(Tx):
        return op.create_rhs(np.zeros(op.n_dofs), np.ascontiguousarray(Tx), T_old)

    fd = -(res(T + eps * u) - res(T - eps * u)) / (2 * eps)
    assert np.linalg.norm(Ju - fd) <= 1e-7 * np.linalg.norm(fd)


def test_gmres_matches_direct_solve():
    """Non-symmetric tangent (velocity, solidification terms) on a small mesh: GMRES with and without the
    inverse diagonal against numpy's dense solve; restarted (restart 5) and full cycles agree."""
    op = O.Operator(2, (3, 2, 2), (0.3, 0.2, 0.2))
    X = op.node_coords()
    T = 1970.0 + 3.0 * np.sin(4 * X[0]) + X[1]
    op.set_tangent(0.01, material=MUSHY, velocity=(-0.5, 0.2, 0.0), do_solidification=True, T_old=T - 0.5)
    op.set_T(T)
    A = np.stack([op.vmult(e, threads=1) for e in np.eye(op.n_dofs)], axis=1)
    assert np.linalg.norm(A - A.T) > 1e-3 * np.linalg.norm(A)  # really non-symmetric
    assert np.allclose(np.diag(A), op.diagonal(), rtol=1e-12, atol=0.0)
    b = np.random.default_rng(2).uniform(-1, 1, op.n_dofs)
    xs = np.linalg.solve(A, b)
    for precond in (0, 1):
        for restart in (5, 200):
            x, it, res = op.gmres(b, rel_tol=1e-12, precond=precond, restart=restart, threads=1)
            assert it > 0 and res <= 1e-12 * np.linalg.norm(b)
            assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


# ------------------------------------------------- V-cycle composition (dense)
def _dense(op):
    n = op.n_dofs
    return np.array([op.vmult_brute(e) for e in np.eye(n)]).T


def _chebyshev_matrix(A, lam_max, deg=3, smoothing_range=20.0):
    """C with x = C b for Chebyshev smoothing from x = 0: I - C A = T_deg((theta - D^-1 A)/delta) /
    T_deg(theta/delta), evaluated through the eigen-decomposition of D^-1/2 A D^-1/2."""
    s = 1.0 / np.sqrt(np.diag(A))
    w, V = np.linalg.eigh(s[:, None] * A * s[None, :])
    lmin = lam_max / smoothing_range
    theta, delta = 0.5 * (lam_max + lmin), 0.5 * (lam_max - lmin)

    def T(x):
        return np.cos(deg * np.arccos(x)) if abs(x) <= 1 else np.cosh(deg * np.arccosh(abs(x))) * np.sign(x) ** deg
    f = np.array([T((theta - x) / delta) / T(theta / delta) for x in w])
    E = (s[:, None] * ((V * f) @ V.T)) / s[None, :]
    return (np.eye(len(s)) - E) @ np.linalg.inv(A)


def _n_nodes(k, nc):
    return (k * nc[0] + 1) * (k * nc[1] + 1) * (k * nc[2] + 1)


@pytest.mark.parametrize("k,nc,coarse_degree", [(1, (8, 8, 8), 0), (2, (4, 4, 4), 0), (3, (4, 4, 4), 2)])
def test_vcycle_equals_dense_multigrid_composition(k, nc, coarse_degree):
    """The oracle's V-cycle equals, to rounding, the recursive two-grid composition built from
    assembled matrices: x = C b; x += P B_c P^T (b - A x); x += C (b - A x) with C the Chebyshev
    smoother (closed form), P the prolongation assembled column by column, and B_c the next
    coarser level's V-cycle (the exact inverse on the coarsest level). Dropping the
    post-smoother changes the result by 15-30 %; the smoothers' lambda_max is the oracle's
    20-step estimate, which is checked against the dense eigenvalue here too (a Ritz value within
    2 % below it)."""
    c_mass, kap = 10.0 if k < 3 else 50.0, 1.3
    op = O.Operator(k, nc, c_mass=c_mass, kappa=kap)
    g = O.GMG(op, nc_coarse=2, coarse_degree=coarse_degree, cheb_degree=3, smoothing_range=20.0, eig_iters=20)
    levels, prev = [], None
    for deg, ncl, lam in g.levels():
        A = _dense(O.Operator(deg, ncl, c_mass=c_mass, kappa=kap))
        P = None
        if prev is not None:
            n_c = _n_nodes(*prev)
            P = np.array([O.prolongate(prev[0], prev[1], deg, ncl, e) for e in np.eye(n_c)]).T
            exact = np.linalg.eigvals(A / np.diag(A)[:, None]).real.max()
            # a Ritz value: a lower bound, close to the largest eigenvalue after 20 steps
            assert (1 - 2e-2) * exact <= lam / 1.2 <= (1 + 1e-12) * exact
        levels.append((A, lam, P))
        prev = (deg, ncl)

    def vcycle(l, b, post=True):
        A, lam, P = levels[l]
        if l == 0:
            return np.linalg.solve(A, b)
        C = _chebyshev_matrix(A, lam)
        x = C @ b
        x = x + P @ vcycle(l - 1, P.T @ (b - A @ x), post)
        return x + C @ (b - A @ x) if post else x

    b = np.random.default_rng(3).uniform(-1, 1, op.n_dofs)
    ref = vcycle(len(levels) - 1, b)
    got = g.vcycle(b)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    # the check is sensitive to the post-smoother (a plausible slip)
    assert np.linalg.norm(vcycle(len(levels) - 1, b, post=False) - ref) > 0.05 * np.linalg.norm(ref)


# ----------------------------------------------------- laser surface flux
def _gauss_integrals(c, sig, L):
    """int_0^L g dx and int_0^L x g dx for g = exp(-(x-c)^2/(2 sig^2))."""
    from math import erf, exp, pi, sqrt
    s2 = sig * sqrt(2.0)
    i0 = sig * sqrt(pi / 2.0) * (erf((L - c) / s2) - erf(-c / s2))
    i1 = c * i0 + sig ** 2 * (exp(-c * c / (2 * sig ** 2)) - exp(-(L - c) ** 2 / (2 * sig ** 2)))
    return i0, i1


SHEAR = np.array([[1.1, 0.2, 0.1], [0.05, 0.9, 0.15], [0.1, -0.1, 1.2]])


def _remap(op, fn):
    """Replace the mapped operator's node coordinates (held by the operator) by fn(x, y, z) of the box's."""
    X = op.node_coords()
    op._keep["X"][:] = np.asarray(fn(*X)).reshape(-1)
    return op._keep["X"].reshape(3, -1)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_mapped_robin_faces_give_the_mapped_areas(k):
    """Robin faces on a mapped mesh (face JxW from the mapping, heat_diffuse_operator.cpp:1090-1129): under
    an affine map x -> A x + b, which Q_k reproduces, sum_i (J 1)_i = c_m |det A| V + alpha sum_f |A e_a x A e_b|
    L_a L_b; the box's face areas (the mapping ignored on the faces) miss it by >5 %."""
    L = (0.5, 0.4, 0.3)
    cm, alpha = 2.0, 3.5
    op = O.Operator(k, (2, 3, 2), L, c_mass=cm, kappa=0.7, robin_mask=63, alpha=alpha, mapped_amp=0.0)
    _remap(op, lambda x, y, z: SHEAR @ np.stack([x, y, z]) + np.array([[0.3], [-0.2], [0.1]]))
    V = L[0] * L[1] * L[2]
    E = SHEAR.T  # columns A e_d as rows
    area = sum(2 * np.linalg.norm(np.cross(E[a], E[b])) * L[a] * L[b] for a, b in ((1, 2), (0, 2), (0, 1)))
    exp = cm * abs(np.linalg.det(SHEAR)) * V + alpha * area
    got = op.vmult(np.ones(op.n_dofs), threads=1).sum()
    assert abs(got - exp) <= 1e-12 * exp
    box = cm * abs(np.linalg.det(SHEAR)) * V + alpha * 2 * (L[1] * L[2] + L[0] * L[2] + L[0] * L[1])
    assert abs(box - exp) > 0.05 * alpha * area
    if k <= 2:  # the diagonal (faces included) is the diagonal of the applied matrix
        A = np.stack([op.vmult(e, threads=1) for e in np.eye(op.n_dofs)], axis=1)
        assert np.allclose(op.diagonal(), np.diag(A), rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("k,nc", [(1, (64, 64, 1)), (2, (40, 30, 2)), (3, (32, 32, 1)), (4, (24, 24, 1))])
def test_mapped_laser_flux_integral(k, nc):
    """Surface flux on a mapped top face: x -> x (1 + 0.1 y) (bilinear, reproduced by Q_k) stretches the
    z+ face to x' <= L_x (1 + 0.1 y) with surface element 1 + 0.1 y; sum_i rhs_i = int q dA' =
    int_0^Ly peak g(y) int_0^{Lx(1+0.1y)} g(x') dx' dy (erf inside, adaptive quadrature outside). Taking
    the box's face (the mapping ignored) misses it by >3 %."""
    from scipy.integrate import quad
    from scipy.special import erf
    L = (1.0, 0.8, 0.5)
    peak, sig, x0, y0, vx, t = 1e3, 0.05, 0.57, 0.77, 1.0, 0.4
    xc = x0 + vx * t
    op = O.Operator(k, nc, L, c_mass=1e3, kappa=2.0, laser=(peak, sig, x0, y0, vx, t), mapped_amp=0.0)
    Xm = _remap(op, lambda x, y, z: np.stack([x * (1 + 0.1 * y), y, z]))
    z = np.zeros(op.n_dofs)
    r = op.create_rhs(np.zeros(op.n_dofs), z, z.copy())
    c = sig * np.sqrt(np.pi / 2)
    ix = lambda U: c * (erf((U - xc) / (sig * np.sqrt(2))) - erf(-xc / (sig * np.sqrt(2))))
    gy = lambda y: np.exp(-(y - y0) ** 2 / (2 * sig ** 2))
    exp = peak * quad(lambda y: gy(y) * ix(L[0] * (1 + 0.1 * y)), 0.0, L[1], epsabs=0, epsrel=1e-13, limit=200)[0]
    tol = {1: 1e-5, 2: 1e-6, 3: 1e-8, 4: 1e-9}[k]
    assert abs(r.sum() - exp) <= tol * exp
    # y moment: sum_i y_i rhs_i = int y q dA'
    expy = peak * quad(lambda y: y * gy(y) * ix(L[0] * (1 + 0.1 * y)), 0.0, L[1], epsabs=0, epsrel=1e-13,
                       limit=200)[0]
    assert abs(Xm[1] @ r - expy) <= tol * expy
    box = peak * ix(L[0]) * quad(gy, 0.0, L[1], epsabs=0, epsrel=1e-13)[0]
    assert abs(box / exp - 1) > 0.03


@pytest.mark.parametrize("k,nc", [(1, (64, 64, 1)), (2, (40, 30, 2)), (3, (32, 32, 1)), (4, (24, 24, 1))])
def test_laser_flux_moments_closed_form(k, nc):
    """Neumann flux of the Gauss profile on z+ (heat_diffuse_operator.cpp:761-768,
    laser_intensity_profiles.hpp:134-139 with radius r = 2 sigma, peak = 2P/(pi r^2)):
    with T = T_old = 0 only the flux is left in create_rhs, and since the Q_k basis reproduces
    1, x and y, the sums of rhs_i, x_i rhs_i and y_i rhs_i are the integrals of q, x q and y q
    over the top face, which have erf closed forms. The beam sits near the x+ edge, so a wrong
    width (sqrt 2 sigma) or a wrong direction of travel (x0 - v t) changes the integral by
    38-68 %; the tolerance is set by the face quadrature's error on the resolved Gaussian."""
    L = (1.0, 0.8, 0.5)
    peak, sig, x0, y0, vx, t = 1e3, 0.05, 0.57, 0.77, 1.0, 0.4
    op = O.Operator(k, nc, L, c_mass=1e3, kappa=2.0, laser=(peak, sig, x0, y0, vx, t))
    z = np.zeros(op.n_dofs)
    r = op.create_rhs(np.zeros(op.n_dofs), z, z.copy())
    X = op.node_coords()
    ix0, ix1 = _gauss_integrals(x0 + vx * t, sig, L[0])
    iy0, iy1 = _gauss_integrals(y0, sig, L[1])
    tol = {1: 1e-5, 2: 1e-6, 3: 1e-8, 4: 1e-9}[k]
    for got, exp in [(r.sum(), peak * ix0 * iy0), (X[0] @ r, peak * ix1 * iy0), (X[1] @ r, peak * ix0 * iy1)]:
        assert abs(got - exp) <= tol * abs(exp)
    # r vanishes off the top face
    assert np.abs(r[X[2] < L[2] - 1e-12]).max() == 0.0
    wrong_sigma = _gauss_integrals(x0 + vx * t, sig * np.sqrt(2), L[0])[0] * _gauss_integrals(y0, sig * np.sqrt(2), L[1])[0]
    wrong_dir = _gauss_integrals(x0 - vx * t, sig, L[0])[0] * iy0
    assert abs(peak * wrong_sigma / r.sum() - 1) > 0.3 and abs(peak * wrong_dir / r.sum() - 1) > 0.3
