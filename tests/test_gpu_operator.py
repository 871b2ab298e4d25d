"""GPU parity of the operator apply, diagonal, residual and norm against the CPU oracle.

Bar: FP64 results within 1e-12 relative of the oracle (the sum-factorised
apply differs from the oracle only in summation order); FP32 within 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu

CASES = [
    (1, (8, 8, 8), "const", 0),
    (2, (16, 16, 16), "const", 0),      # the CPU-runnable Q2 16^3 configuration
    (3, (8, 8, 8), "qp", 0),
    (3, (6, 5, 9), "qp", 0),            # ragged: no multiple of the chunk shape
    (4, (5, 4, 6), "law", 0),
    (2, (7, 9, 4), "cell", 0),
    (3, (4, 4, 4), "mushy", 0),
    (2, (6, 6, 6), "const", 0b000011),  # Dirichlet on x faces
    (3, (5, 7, 6), "qp", 0b110100),     # Dirichlet on y-, z-, z+
    (1, (32, 32, 1), "const", 0),       # one cell thick
]


@pytest.mark.parametrize("degree,nc,mode,dmask", CASES)
def test_vmult_matches_oracle(cuda, degree, nc, mode, dmask):
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, L=(1.0, 0.7, 1.3), mode=mode, dirichlet=dmask, seed=degree)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    xd = torch.tensor(x, device="cuda")
    yd = torch.full_like(xd, 3.0)  # vmult must overwrite dst
    gop.vmult(yd, xd)
    torch.cuda.synchronize()
    assert rel_err(yd.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("degree,nc,mode,dmask", [(3, (6, 6, 6), "const", 0), (3, (5, 4, 7), "qp", 0),
                                                 (2, (6, 5, 4), "law", 0b001100), (4, (4, 4, 3), "cell", 0)])
def test_vmult_mapped(cuda, degree, nc, mode, dmask):
    """Mesh mapped by x + 0.05 sin(2 pi y) sin(2 pi z) on all MappingQ_k nodes."""
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, dirichlet=dmask, mapped=0.05, seed=5)
    x = np.random.default_rng(8).uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    y = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    gop.vmult(y, torch.tensor(x, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), ref) < 1e-12
    d = torch.empty_like(y)
    gop.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), orc.inverse_diagonal()) < 1e-12


@pytest.mark.parametrize("degree,nc,mode", [(2, (6, 6, 6), "qp"), (3, (5, 4, 7), "law"), (4, (4, 4, 4), "const"),
                                             (3, (7, 6, 5), "const"), (3, (9, 5, 4), "cell")])
def test_vmult_fp32(cuda, degree, nc, mode):
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, precision=1, seed=3)
    x = np.random.default_rng(1).uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    xd = torch.tensor(x, dtype=torch.float32, device="cuda")
    yd = torch.empty_like(xd)
    gop.vmult(yd, xd)
    torch.cuda.synchronize()
    assert rel_err(yd.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("degree,nc,mode,dmask", [(2, (6, 5, 4), "qp", 0), (3, (4, 4, 4), "law", 3),
                                                 (1, (8, 8, 2), "const", 0)])
def test_inverse_diagonal(cuda, degree, nc, mode, dmask):
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, dirichlet=dmask)
    ref = orc.inverse_diagonal()
    d = torch.zeros(orc.n_dofs, dtype=torch.float64, device="cuda")
    gop.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("degree,nc,mode", [(2, (6, 5, 4), "const"), (3, (4, 5, 3), "law")])
def test_create_rhs(cuda, degree, nc, mode):
    torch = cuda
    orc, gop, data = make_pair(torch, degree, nc, mode=mode)
    rng = np.random.default_rng(11)
    T = orc._keep["T"] if mode == "law" else rng.uniform(0, 1, orc.n_dofs)
    T_old = T + 0.01 * rng.standard_normal(orc.n_dofs)
    q = rng.uniform(0, 2, orc.n_dofs)
    base = rng.standard_normal(orc.n_dofs)
    ref = orc.create_rhs(base.copy(), np.ascontiguousarray(T), T_old, q)
    Td = data["T"] if mode == "law" else torch.tensor(T, device="cuda")
    dst = torch.tensor(base, device="cuda")
    gop.create_rhs(dst, Td, torch.tensor(T_old, device="cuda"), torch.tensor(q, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref) < 1e-11


def test_l2_norm(cuda):
    torch = cuda
    orc, gop, _ = make_pair(torch, 3, (5, 6, 4), L=(0.1, 0.2, 0.3))
    u = np.random.default_rng(2).uniform(-1, 1, orc.n_dofs)
    ref = orc.l2_norm(u)
    val = gop.l2_norm(torch.tensor(u, device="cuda"))
    assert abs(val - ref) / ref < 1e-12


def test_robin_tangent_and_rhs(cuda):
    torch = cuda
    import meltpooldg_amd as M
    nc, L = (6, 5, 2), (0.1, 0.1, 1.0)
    rng = np.random.default_rng(5)
    orc0 = O.Operator(1, nc, L)
    T = 1000.0 + 50.0 * rng.uniform(-1, 1, orc0.n_dofs)
    kw = dict(robin_mask=2 | 8, alpha=25.0, emissivity=0.98, T_inf_conv=300.0, T_inf_rad=310.0)
    orc = O.Operator(1, nc, L, c_mass=3.6e8, kappa=55.563, T=T, **kw)
    gop = M.HeatDiffuseMultiPhaseOperator(1, nc, L, rho_cp=3.6e6)
    gop.reset_time_increment(1e-2)
    gop.set_conductivity_constant(55.563)
    Td = torch.tensor(T, device="cuda")
    gop.set_robin(2 | 8, alpha=25.0, T_inf_conv=300.0, emissivity=0.98, T_inf_rad=310.0, T=Td)
    x = rng.uniform(-1, 1, orc.n_dofs)
    y = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    gop.vmult(y, torch.tensor(x, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), orc.vmult(x)) < 1e-12
    d = torch.empty_like(y)
    gop.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), orc.inverse_diagonal()) < 1e-12
    T_old = T - 1.0
    ref = orc.create_rhs(np.zeros(orc.n_dofs), T, T_old)
    dst = torch.zeros_like(y)
    gop.create_rhs(dst, Td, torch.tensor(T_old, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref) < 1e-11


MAPS = {
    # affine: reproduced exactly, every face tilted
    "shear": lambda x, y, z: np.stack([1.1 * x + 0.2 * y + 0.1 * z + 0.3, 0.05 * x + 0.9 * y + 0.15 * z,
                                       0.1 * x - 0.1 * y + 1.2 * z]),
    # curved faces x+- and a stretched top face
    "curved": lambda x, y, z: np.stack([x * (1 + 0.1 * y) + 0.05 * np.sin(2 * np.pi * y / 0.8) * np.sin(
        2 * np.pi * z / 0.5), y + 0.03 * np.sin(3 * x), z + 0.02 * x * y]),
}


@pytest.mark.parametrize("degree,nc,mapname", [(1, (10, 8, 3), "shear"), (2, (6, 5, 3), "curved"),
                                               (3, (5, 4, 2), "curved"), (4, (3, 3, 2), "shear"),
                                               (3, (20, 3, 2), "curved")])
def test_mapped_robin_faces_and_flux_match_oracle(cuda, degree, nc, mapname):
    """Robin faces (convection + radiation) on x-, x+, y-, y+, z- and the laser flux on z+ of a mapped
    mesh: apply, inverse diagonal and create_rhs against the oracle (whose mapped face geometry is
    pinned by the face areas of affine maps and the flux integral of a stretched face), and the nodal
    flux field at the mapped nodes."""
    torch = cuda
    import meltpooldg_amd as M
    L = (1.0, 0.8, 0.5)
    rng = np.random.default_rng(31)
    X = MAPS[mapname](*O.Operator(degree, nc, L).node_coords())
    T = 1000.0 + 50.0 * rng.uniform(-1, 1, X.shape[1])
    kw = dict(robin_mask=0b011111, alpha=25.0, emissivity=0.98, T_inf_conv=300.0, T_inf_rad=310.0)
    orc = O.Operator(degree, nc, L, c_mass=3.6e8, kappa=55.563, T=T, laser=LASER, mapped_amp=0.0, **kw)
    orc._keep["X"][:] = X.reshape(-1)
    gop = M.HeatDiffuseMultiPhaseOperator(degree, nc, L, rho_cp=3.6e6)
    gop.reset_time_increment(1e-2)
    gop.set_conductivity_constant(55.563)
    Xd = torch.tensor(X.reshape(-1), device="cuda")
    gop.set_mapping(Xd)
    Td = torch.tensor(T, device="cuda")
    gop.set_robin(0b011111, alpha=25.0, T_inf_conv=300.0, emissivity=0.98, T_inf_rad=310.0, T=Td)
    gop.set_surface_flux(*LASER)
    x = rng.uniform(-1, 1, orc.n_dofs)
    y = torch.empty(orc.n_dofs, dtype=torch.float64, device="cuda")
    gop.vmult(y, torch.tensor(x, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), orc.vmult(x)) < 1e-12
    d = torch.empty_like(y)
    gop.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), orc.inverse_diagonal()) < 1e-12
    T_old = T - 1.0
    ref = orc.create_rhs(np.zeros(orc.n_dofs), T, T_old)
    dst = torch.zeros_like(y)
    gop.create_rhs(dst, Td, torch.tensor(T_old, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref) < 1e-11
    out = torch.full((orc.n_dofs,), 7.0, dtype=torch.float64, device="cuda")
    gop.surface_flux_field(out)
    torch.cuda.synchronize()
    peak, sig, x0, y0, vx, t = LASER
    q = peak * np.exp(-((X[0] - x0 - vx * t) ** 2 + (X[1] - y0) ** 2) / (2 * sig ** 2))
    top = np.zeros(orc.n_dofs, dtype=bool)
    top[-(nc[0] * degree + 1) * (nc[1] * degree + 1):] = True
    q[~top] = 0.0
    assert rel_err(out.cpu().numpy(), q) < 1e-14


def test_robin_faces_and_flux_are_bitwise_repeatable(cuda):
    """Robin apply, diagonal and residual with the surface flux, applied repeatedly: identical bits
    (each boundary DoF sums its faces and cells in a fixed order, without atomics)."""
    torch = cuda
    import meltpooldg_amd as M
    nc, L = (24, 20, 12), (1.0, 0.8, 0.5)
    rng = np.random.default_rng(41)
    gop = M.HeatDiffuseMultiPhaseOperator(3, nc, L, rho_cp=3.6e6)
    gop.reset_time_increment(1e-2)
    gop.set_conductivity_constant(55.563)
    n = gop.n_dofs
    Td = torch.tensor(1000.0 + 50.0 * rng.uniform(-1, 1, n), device="cuda")
    gop.set_robin(0b011111, alpha=25.0, T_inf_conv=300.0, emissivity=0.98, T_inf_rad=310.0, T=Td)
    gop.set_surface_flux(*LASER)
    x = torch.tensor(rng.uniform(-1, 1, n), device="cuda")
    outs = []
    for _ in range(3):
        y = torch.empty(n, dtype=torch.float64, device="cuda")
        gop.vmult(y, x)
        d = torch.empty_like(y)
        gop.compute_inverse_diagonal_from_matrixfree(d)
        r = torch.zeros_like(y)
        gop.create_rhs(r, Td, Td - 1.0)
        torch.cuda.synchronize()
        outs.append([v.cpu().numpy().copy() for v in (y, d, r)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


# Meshes with 64 columns of chunks (a multiple of the 8 XCDs): the launch then gives each XCD
# a contiguous range of columns, a path the smaller meshes above do not take.
XCD_CASES = [(1, (8, 32, 32), "const", 0, 0), (2, (14, 24, 16), "qp", 0, 0), (3, (8, 32, 16), "qp", 0b000101, 0),
             (4, (4, 32, 24), "law", 0, 0), (3, (8, 32, 16), "cell", 0, 1), (3, (8, 32, 16), "const", 0b100000, 1),
             (4, (12, 16, 16), "qp", 0, 1)]


@pytest.mark.parametrize("degree,nc,mode,dmask,precision", XCD_CASES)
def test_vmult_xcd_column_launch_matches_oracle(cuda, degree, nc, mode, dmask, precision):
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, L=(0.5, 1.0, 0.8), mode=mode, dirichlet=dmask, seed=11,
                            precision=precision)
    x = np.random.default_rng(12).uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    dtype = torch.float64 if precision == 0 else torch.float32
    xd = torch.tensor(x, dtype=dtype, device="cuda")
    yd = torch.full_like(xd, 3.0)
    gop.vmult(yd, xd)
    torch.cuda.synchronize()
    assert rel_err(yd.cpu().numpy(), ref) < (1e-12 if precision == 0 else 1e-5)


@pytest.mark.parametrize("degree,N,mode", [(3, 128, "qp"), (4, 96, "law")])
def test_vmult_full_size_properties(cuda, degree, N, mode):
    """The benchmark sizes (Q3 on 128^3 cells with per-point conductivity, Q4 on 96^3 cells with
    k(T) from T DoFs; 57,066,625 DoFs each), checked by properties the oracle need not run:
    sum_i (J 1)_i = c_m |Omega| (the mass matrix integrates the constant exactly, the stiffness
    part annihilates constants), symmetry x.(J z) = z.(J x), and linearity."""
    torch = cuda
    import meltpooldg_amd as M
    dt, rho_cp, L = 1e-3, 2.5, (1.0, 0.8, 1.2)
    op = M.HeatDiffuseMultiPhaseOperator(degree, (N, N, N), L, rho_cp=rho_cp)
    op.reset_time_increment(dt)
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    n = op.n_dofs
    assert n == (degree * N + 1) ** 3 == 57066625
    if mode == "qp":
        T = torch.rand(op.n_q_points, generator=g, device="cuda", dtype=torch.float64) * 2.0
        k = 1.0 + 0.5 * torch.tanh((T - 1.0) / 0.1)
        del T
        op.set_conductivity_qp(k)
    else:
        T = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 2.0
        op.set_conductivity_law(M.LAW_TANH, (1.0, 0.5, 1.0, 0.1), T)
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    op.vmult(y, torch.ones(n, dtype=torch.float64, device="cuda"))
    c_m, vol = rho_cp / dt, L[0] * L[1] * L[2]
    assert abs(float(y.sum()) - c_m * vol) <= 1e-10 * c_m * vol
    x = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 0.5
    z = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 0.5
    jx, jz = torch.empty_like(x), torch.empty_like(x)
    op.vmult(jx, x)
    op.vmult(jz, z)
    a, b = float(torch.dot(z, jx)), float(torch.dot(x, jz))
    assert abs(a - b) <= 1e-12 * float(torch.linalg.norm(x)) * float(torch.linalg.norm(jz))
    assert float(torch.dot(x, jx)) > 0.0
    op.vmult(y, 2.0 * x - 3.0 * z)
    ref = 2.0 * jx - 3.0 * jz
    assert float(torch.linalg.norm(y - ref)) <= 1e-13 * float(torch.linalg.norm(ref))


@pytest.mark.parametrize("degree,nc,mode,precision", [(3, (8, 16, 16), "const", 0), (3, (12, 16, 16), "cell", 0),
                                                      (3, (16, 16, 16), "const", 1)])
def test_vmult_mapped_xcd_column_launch_matches_oracle(cuda, degree, nc, mode, precision):
    """Mapped cells take their own chunk shape; meshes with 64 columns of it (XCD-remapped launch)."""
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, mapped=0.05, seed=13, precision=precision)
    x = np.random.default_rng(14).uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    dtype = torch.float64 if precision == 0 else torch.float32
    y = torch.full((orc.n_dofs,), 3.0, dtype=dtype, device="cuda")
    gop.vmult(y, torch.tensor(x, dtype=dtype, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), ref) < (1e-12 if precision == 0 else 1e-5)


# Meshes longer than one chunk along x that are no multiple of it, with ragged y/z columns: the
# x-face carried in LDS between chunks, for every chunk shape (FP64 Q3 8x4x2, FP32 Q3 16x2x2,
# mapped Q3 16x2x2, FP64 Q4 4x4x3, FP32 Q4 12x2x2, FP64 Q2 14x3x2).
WIDE_CASES = [(3, (19, 5, 3), "qp", 0, None), (3, (37, 5, 3), "const", 1, None), (3, (37, 5, 3), "qp", 1, None),
              (3, (37, 5, 3), "law", 1, None), (3, (35, 4, 5), "const", 0, 0.05), (3, (35, 4, 5), "qp", 0, 0.05),
              (3, (35, 4, 5), "const", 1, 0.05), (4, (9, 5, 4), "qp", 0, None), (4, (27, 3, 5), "const", 1, None),
              (4, (27, 3, 5), "law", 1, None), (2, (31, 4, 3), "cell", 0, None), (4, (9, 5, 4), "const", 0, 0.05)]


@pytest.mark.parametrize("degree,nc,mode,precision,mapped", WIDE_CASES)
def test_vmult_chunks_along_x_match_oracle(cuda, degree, nc, mode, precision, mapped):
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, L=(1.7, 0.6, 0.9), mode=mode, seed=17, precision=precision,
                            mapped=mapped)
    x = np.random.default_rng(18).uniform(-1.0, 1.0, orc.n_dofs)
    ref = orc.vmult(x)
    dtype = torch.float64 if precision == 0 else torch.float32
    y = torch.full((orc.n_dofs,), 3.0, dtype=dtype, device="cuda")
    gop.vmult(y, torch.tensor(x, dtype=dtype, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), ref) < (1e-12 if precision == 0 else 1e-5)


LASER = (1e3, 0.05, 0.57, 0.77, 1.0, 0.4)  # peak, sigma, x0, y0, vx, t: beam centre near the x+ edge


@pytest.mark.parametrize("degree,nc,mode", [(2, (12, 10, 3), "const"), (3, (10, 8, 2), "law"), (1, (24, 20, 2), "qp")])
def test_create_rhs_with_surface_flux_matches_oracle(cuda, degree, nc, mode):
    """create_rhs with the Gaussian z+ flux (rhs_boundary_loop, heat_diffuse_operator.cpp:761-768) on
    top of the cell residual, against the oracle's laser term (pinned by its closed-form moments)."""
    torch = cuda
    L = (1.0, 0.8, 0.5)
    orc0, gop, data = make_pair(torch, degree, nc, L=L, mode=mode, seed=6)
    kw = dict(c_mass=orc0.s.c_mass, laser=LASER)
    if mode == "law":
        kw.update(law=O.LAW_TANH, law_p=(1.0, 0.5, 1.0, 0.1), T=orc0._keep["T"])
    elif mode == "qp":
        kw.update(kappa_qp=orc0._keep["k"])
    orc = O.Operator(degree, nc, L, **kw)
    gop.set_surface_flux(*LASER)
    rng = np.random.default_rng(12)
    T = orc0._keep["T"] if mode == "law" else rng.uniform(0, 1, orc.n_dofs)
    T_old = T + 0.01 * rng.standard_normal(orc.n_dofs)
    ref = orc.create_rhs(np.zeros(orc.n_dofs), np.ascontiguousarray(T), T_old)
    Td = data["T"] if mode == "law" else torch.tensor(T, device="cuda")
    dst = torch.zeros(orc.n_dofs, dtype=torch.float64, device="cuda")
    gop.create_rhs(dst, Td, torch.tensor(T_old, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref) < 1e-11
    # the flux alone (T = T_old = 0) against the oracle
    z = np.zeros(orc.n_dofs)
    ref0 = orc.create_rhs(np.zeros(orc.n_dofs), z, z.copy()) if mode != "law" else None
    if ref0 is not None:
        d0 = torch.zeros_like(dst)
        zd = torch.zeros_like(dst)
        gop.create_rhs(d0, zd, zd)
        torch.cuda.synchronize()
        assert rel_err(d0.cpu().numpy(), ref0) < 1e-12


def test_surface_flux_field_is_the_profile_at_the_nodes(cuda):
    torch = cuda
    import meltpooldg_amd as M
    k, nc, L = 3, (8, 6, 4), (1.0, 0.8, 0.5)
    op = M.HeatDiffuseMultiPhaseOperator(k, nc, L)
    op.set_surface_flux(*LASER)
    out = torch.full((op.n_dofs,), 7.0, dtype=torch.float64, device="cuda")
    op.surface_flux_field(out)
    torch.cuda.synchronize()
    X = O.Operator(k, nc, L).node_coords()
    peak, sig, x0, y0, vx, t = LASER
    q = peak * np.exp(-((X[0] - x0 - vx * t) ** 2 + (X[1] - y0) ** 2) / (2 * sig ** 2))
    q[X[2] < L[2] - 1e-12] = 0.0
    assert rel_err(out.cpu().numpy(), q) < 1e-14


def test_flux_and_robin_on_one_face_are_refused(cuda):
    """The reference throws when a face has a Neumann flux and radiation/convection
    (heat_diffuse_operator.cpp:739-743); both orders of setting are refused here."""
    torch = cuda
    import meltpooldg_amd as M
    op = M.HeatDiffuseMultiPhaseOperator(1, (4, 4, 2))
    op.set_robin(32, alpha=2.0, T_inf_conv=1.0)
    with pytest.raises(M.MpdgError):
        op.set_surface_flux(*LASER)
    op.set_robin(16, alpha=2.0, T_inf_conv=1.0)  # z- only: the flux on z+ is allowed
    op.set_surface_flux(*LASER)
    with pytest.raises(M.MpdgError):
        op.set_robin(32 | 2, alpha=2.0, T_inf_conv=1.0)
    with pytest.raises(M.MpdgError):
        op.set_surface_flux(1.0, 0.0, 0.5, 0.5, 0.0, 0.0)  # sigma <= 0


@pytest.mark.parametrize("degree,nc,mode,precision,mapped,dmask", [
    (3, (8, 32, 16), "qp", 0, None, 0), (2, (14, 24, 16), "const", 0, None, 0), (4, (4, 32, 24), "law", 0, None, 0),
    (3, (16, 16, 16), "qp", 1, None, 0), (3, (12, 18, 10), "const", 0, 0.05, 0), (3, (9, 13, 11), "qp", 0, None, 0b110101),
    (2, (10, 21, 9), "cell", 1, 0.05, 0)])
def test_vmult_is_deterministic(cuda, degree, nc, mode, precision, mapped, dmask):
    """vmult on meshes with many columns in y and z, repeated: bitwise identical results, and equal
    to the oracle. DoF rows shared by neighbouring columns must not depend on the order in which
    the columns finish; rows where four columns meet take the edge buffer and its fix-up pass."""
    torch = cuda
    orc, gop, _ = make_pair(torch, degree, nc, mode=mode, seed=23, precision=precision, mapped=mapped,
                            dirichlet=dmask)
    dtype = torch.float64 if precision == 0 else torch.float32
    x = np.random.default_rng(24).uniform(-1.0, 1.0, orc.n_dofs)
    xd = torch.tensor(x, dtype=dtype, device="cuda")
    outs = []
    for _ in range(5):
        y = torch.full_like(xd, 3.0)
        gop.vmult(y, xd)
        outs.append(y)
    torch.cuda.synchronize()
    for y in outs[1:]:
        assert torch.equal(y, outs[0])
    assert rel_err(outs[0].cpu().numpy(), orc.vmult(x)) < (1e-12 if precision == 0 else 1e-5)


@pytest.mark.parametrize("degree,nc,amp", [(3, (4, 3, 5), 0.05), (2, (3, 4, 2), 0.0), (4, (2, 3, 2), 0.05)])
def test_node_coordinates_match_the_oracle_mapping(cuda, degree, nc, amp):
    """M.node_coordinates (the benchmark's curved box) equals the oracle's MappingQ_k nodes."""
    import meltpooldg_amd as M
    L = (1.0, 0.9, 1.2)
    orc = O.Operator(degree, nc, L, mapped_amp=amp if amp else None)
    X = M.node_coordinates(degree, nc, L, mapping_amplitude=amp).cpu().numpy()
    ref = orc.coords if amp else np.concatenate(orc.node_coords())
    assert np.max(np.abs(X - ref)) < 1e-14
