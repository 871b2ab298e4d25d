"""Shared helpers of the parity tests: build matching oracle / GPU operators."""
import numpy as np

from oracle import oracle as O


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tanh_law_params(k0=1.0):
    return (k0, 0.5, 1.0, 0.1)


def make_pair(torch, degree, nc, L=(1.0, 1.0, 1.0), rho_cp=1.0, dt=1e-3, mode="const", seed=0,
              dirichlet=0, kappa=1.0, precision=0, mapped=None):
    """Oracle operator and GPU operator with the same data; returns (orc, gpu, data)."""
    import meltpooldg_amd as M
    rng = np.random.default_rng(seed)
    gop = M.HeatDiffuseMultiPhaseOperator(degree, nc, L, rho_cp=rho_cp, dirichlet_faces=dirichlet,
                                          precision=precision)
    gop.reset_time_increment(dt)
    dtype = torch.float64 if precision == 0 else torch.float32
    data = {}
    kw = dict(c_mass=rho_cp / dt, dirichlet_mask=dirichlet, mapped_amp=mapped)
    n1 = degree + 1
    ncells = nc[0] * nc[1] * nc[2]
    if mode == "const":
        gop.set_conductivity_constant(kappa)
        orc = O.Operator(degree, nc, L, kappa=kappa, **kw)
    elif mode == "qp":
        Tq = rng.uniform(0.0, 2.0, ncells * n1 ** 3)
        kq = 1.0 * (1.0 + 0.5 * np.tanh((Tq - 1.0) / 0.1))
        data["k"] = torch.tensor(kq, dtype=dtype, device="cuda")
        gop.set_conductivity_qp(data["k"])
        orc = O.Operator(degree, nc, L, kappa_qp=kq, **kw)
    elif mode == "cell":
        kc = rng.uniform(0.5, 1.5, ncells)
        data["k"] = torch.tensor(kc, dtype=dtype, device="cuda")
        gop.set_conductivity_cell(data["k"])
        orc = O.Operator(degree, nc, L, kappa_cell=kc, **kw)
    elif mode in ("law", "mushy"):
        tmp = O.Operator(degree, nc, L)
        X = tmp.node_coords()
        T = 1.0 + 0.8 * np.sin(3.0 * X[0]) * np.cos(2.0 * X[1]) + 0.3 * X[2] + 0.05 * rng.standard_normal(tmp.n_dofs)
        if mode == "law":
            law, p = O.LAW_TANH, tanh_law_params()
        else:
            law, p = O.LAW_MUSHY, (40.0, 30.0, 0.5, 1.5)
        data["T"] = torch.tensor(T, dtype=dtype, device="cuda")
        gop.set_conductivity_law(law, p, data["T"])
        orc = O.Operator(degree, nc, L, law=law, law_p=p, T=T, **kw)
    else:
        raise ValueError(mode)
    if mapped is not None:
        data["X"] = torch.tensor(orc.coords, dtype=torch.float64, device="cuda")
        gop.set_mapping(data["X"])
    return orc, gop, data
