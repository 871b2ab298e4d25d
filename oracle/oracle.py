"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/libmpdg_oracle.so (built by oracle/Makefile) and wraps the
operator description (``orc_op``) plus the entry points declared in
oracle/mpdg_oracle.h. Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use this module; the product package
``meltpooldg_amd`` never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpdg_oracle.so")

KAPPA_CONST, KAPPA_QP, KAPPA_CELL, KAPPA_LAW = 0, 1, 2, 3
LAW_TANH, LAW_MUSHY = 1, 2
_dp = C.POINTER(C.c_double)


class OrcTangent(C.Structure):
    _fields_ = [
        ("on", C.c_int),
        ("dt", C.c_double),
        ("transition", C.c_int),
        ("k", C.c_double * 2),
        ("cp", C.c_double * 2),
        ("rho", C.c_double * 2),
        ("T_solidus", C.c_double),
        ("T_liquidus", C.c_double),
        ("do_solidification", C.c_int),
        ("velocity", C.c_double * 3),
        ("T_old", _dp),
        ("velocity_field", _dp),
    ]


class OrcOp(C.Structure):
    _fields_ = [
        ("degree", C.c_int),
        ("nc", C.c_int * 3),
        ("L", C.c_double * 3),
        ("coords", _dp),
        ("geom_kind", C.c_int),
        ("geom_amp", C.c_double),
        ("c_mass", C.c_double),
        ("kappa_mode", C.c_int),
        ("kappa", C.c_double),
        ("kappa_arr", _dp),
        ("law", C.c_int),
        ("law_p", C.c_double * 4),
        ("T", _dp),
        ("dirichlet_mask", C.c_int),
        ("robin_mask", C.c_int),
        ("alpha", C.c_double),
        ("emissivity", C.c_double),
        ("T_inf_conv", C.c_double),
        ("T_inf_rad", C.c_double),
        ("laser_on", C.c_int),
        ("laser_p", C.c_double * 6),
        ("tg", OrcTangent),
    ]


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s"], cwd=HERE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OrcOp)
        L.orc_fe1d.argtypes = [C.c_int] + [_dp] * 6
        L.orc_n_dofs.argtypes = [P]
        L.orc_n_dofs.restype = C.c_long
        L.orc_n_cells.argtypes = [P]
        L.orc_n_cells.restype = C.c_long
        L.orc_make_coords.argtypes = [P, C.c_int, C.c_double, _dp]
        L.orc_vmult.argtypes = [P, _dp, _dp, C.c_int]
        L.orc_vmult_brute.argtypes = [P, _dp, _dp]
        L.orc_diagonal.argtypes = [P, _dp]
        L.orc_inverse_diagonal.argtypes = [P, _dp]
        L.orc_create_rhs.argtypes = [P, _dp, _dp, _dp, _dp]
        L.orc_l2_norm.argtypes = [P, _dp]
        L.orc_l2_norm.restype = C.c_double
        L.orc_kappa_law.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.orc_interpolate_qp.argtypes = [P, _dp, _dp]
        L.orc_cg.argtypes = [P, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p,
                             C.c_int, _dp, C.c_int]
        L.orc_cg.restype = C.c_int
        L.orc_gmg_create.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_gmg_create.restype = C.c_void_p
        L.orc_gmg_vcycle.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_gmg_n_levels.argtypes = [C.c_void_p]
        L.orc_gmg_level_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), _dp]
        L.orc_gmg_destroy.argtypes = [C.c_void_p]
        L.orc_prolongate.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), _dp, _dp]
        L.orc_restrict.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), _dp, _dp]
        L.orc_chebyshev.argtypes = [P, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, C.c_int]
        L.orc_estimate_lambda_max.argtypes = [P, _dp, C.c_int]
        L.orc_estimate_lambda_max.restype = C.c_double
        L.orc_heat_run.argtypes = [P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, _dp, _dp,
                                   C.POINTER(C.c_int), C.c_int]
        L.orc_heat_run.restype = C.c_int
        L.orc_gmres.argtypes = [P, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                C.c_int]
        L.orc_gmres.restype = C.c_int
        L.orc_material.argtypes = [C.POINTER(OrcTangent), C.c_double, _dp]
        L.orc_heat_run_tangent.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                           _dp, C.POINTER(C.c_int), C.c_int]
        L.orc_heat_run_tangent.restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def fe1d(k):
    n = k + 1
    out = [np.zeros(n) for _ in range(3)] + [np.zeros((n, n)) for _ in range(3)]
    lib().orc_fe1d(k, *[ptr(a) for a in out])
    return dict(zip(["nodes", "qp", "qw", "S", "G", "D"], out))


class Operator:
    """Heat operator J = c_mass*M + K_kappa (+ Robin faces) on a structured Q_k box mesh.

    Arrays handed over are kept alive by this object.
    """

    def __init__(self, degree, nc, L=(1.0, 1.0, 1.0), c_mass=1.0, kappa=1.0, kappa_qp=None,
                 kappa_cell=None, law=None, law_p=None, T=None, mapped_amp=None,
                 dirichlet_mask=0, robin_mask=0, alpha=0.0, emissivity=0.0, T_inf_conv=0.0,
                 T_inf_rad=0.0, laser=None):
        s = OrcOp()
        s.degree = degree
        s.nc[:] = list(nc)
        s.L[:] = list(L)
        s.c_mass = c_mass
        s.kappa = kappa
        s.kappa_mode = KAPPA_CONST
        self._keep = {}
        if kappa_qp is not None:
            self._keep["k"] = np.ascontiguousarray(kappa_qp, dtype=np.float64)
            s.kappa_mode = KAPPA_QP
            s.kappa_arr = ptr(self._keep["k"])
        elif kappa_cell is not None:
            self._keep["k"] = np.ascontiguousarray(kappa_cell, dtype=np.float64)
            s.kappa_mode = KAPPA_CELL
            s.kappa_arr = ptr(self._keep["k"])
        elif law is not None:
            s.kappa_mode = KAPPA_LAW
            s.law = law
            s.law_p[:] = list(law_p)
        if T is not None:
            self._keep["T"] = np.ascontiguousarray(T, dtype=np.float64)
            s.T = ptr(self._keep["T"])
        s.dirichlet_mask = dirichlet_mask
        s.robin_mask = robin_mask
        s.alpha = alpha
        s.emissivity = emissivity
        s.T_inf_conv = T_inf_conv
        s.T_inf_rad = T_inf_rad
        if laser is not None:
            s.laser_on = 1
            s.laser_p[:] = list(laser)
        self.s = s
        self.n_dofs = lib().orc_n_dofs(C.byref(s))
        self.n_cells = lib().orc_n_cells(C.byref(s))
        if mapped_amp is not None:
            self._keep["X"] = np.zeros(3 * self.n_dofs)
            lib().orc_make_coords(C.byref(s), 1, mapped_amp, ptr(self._keep["X"]))
            s.coords = ptr(self._keep["X"])
            s.geom_kind = 1
            s.geom_amp = mapped_amp

    @property
    def ref(self):
        return C.byref(self.s)

    @property
    def coords(self):
        return self._keep.get("X")

    def set_T(self, T):
        self._keep["T"] = np.ascontiguousarray(T, dtype=np.float64)
        self.s.T = ptr(self._keep["T"])

    def node_coords(self):
        """(x, y, z) of every DoF of the affine box, each of length n_dofs."""
        X = np.zeros(3 * self.n_dofs)
        lib().orc_make_coords(self.ref, 0, 0.0, ptr(X))
        return X.reshape(3, -1)

    def vmult(self, src, threads=os.cpu_count()):
        dst = np.zeros(self.n_dofs)
        lib().orc_vmult(self.ref, ptr(dst), ptr(np.ascontiguousarray(src, dtype=np.float64)), threads)
        return dst

    def vmult_brute(self, src):
        dst = np.zeros(self.n_dofs)
        lib().orc_vmult_brute(self.ref, ptr(dst), ptr(np.ascontiguousarray(src, dtype=np.float64)))
        return dst

    def diagonal(self):
        d = np.zeros(self.n_dofs)
        lib().orc_diagonal(self.ref, ptr(d))
        return d

    def inverse_diagonal(self):
        d = np.zeros(self.n_dofs)
        lib().orc_inverse_diagonal(self.ref, ptr(d))
        return d

    def create_rhs(self, dst, T, T_old, q=None):
        lib().orc_create_rhs(self.ref, ptr(dst), ptr(T), ptr(T_old), ptr(q))
        return dst

    def l2_norm(self, u):
        return lib().orc_l2_norm(self.ref, ptr(np.ascontiguousarray(u, dtype=np.float64)))

    def interpolate_qp(self, u):
        n1 = self.s.degree + 1
        out = np.zeros(self.n_cells * n1 ** 3)
        lib().orc_interpolate_qp(self.ref, ptr(np.ascontiguousarray(u, dtype=np.float64)), ptr(out))
        return out

    def set_tangent(self, dt, material=None, velocity=(0.0, 0.0, 0.0), do_solidification=False, T_old=None,
                    velocity_field=None):
        """Switch to the full Newton tangent. material: None (constant: rho_cp = c_mass*dt, k = kappa)
        or dict(k=(liquid, solid), cp=(..), rho=(..), T_solidus=.., T_liquidus=..) for the mushy zone.
        velocity_field: None (the constant `velocity`) or 3*n_dofs nodal values (x, y, z blocks)."""
        t = self.s.tg
        t.on = 1
        t.dt = dt
        if material is None:
            t.transition = 0
            t.k[:] = [self.s.kappa, self.s.kappa]
            t.cp[:] = [self.s.c_mass * dt] * 2
            t.rho[:] = [1.0, 1.0]
        else:
            t.transition = 1
            t.k[:] = list(material["k"])
            t.cp[:] = list(material["cp"])
            t.rho[:] = list(material["rho"])
            t.T_solidus, t.T_liquidus = material["T_solidus"], material["T_liquidus"]
        t.do_solidification = 1 if do_solidification else 0
        t.velocity[:] = list(velocity)
        if velocity_field is not None:
            self._keep["velocity_field"] = np.ascontiguousarray(velocity_field, dtype=np.float64).reshape(-1)
            assert self._keep["velocity_field"].size == 3 * self.n_dofs
            t.velocity_field = ptr(self._keep["velocity_field"])
        else:
            t.velocity_field = None
        if T_old is not None:
            self._keep["T_old"] = np.ascontiguousarray(T_old, dtype=np.float64)
            t.T_old = ptr(self._keep["T_old"])

    def material(self, T):
        out = np.zeros(4)
        lib().orc_material(C.byref(self.s.tg), float(T), ptr(out))
        return out

    def gmres(self, b, x=None, rel_tol=1e-10, abs_tol=1e-20, max_it=10000, restart=28, precond=1, criterion=1,
              threads=os.cpu_count()):
        x = np.zeros(self.n_dofs) if x is None else np.array(x, dtype=np.float64)
        res = C.c_double(0)
        it = lib().orc_gmres(self.ref, ptr(np.ascontiguousarray(b)), ptr(x), rel_tol, abs_tol, max_it, restart,
                             precond, criterion, C.byref(res), threads)
        return x, it, res.value

    def cg(self, b, x=None, rel_tol=1e-10, abs_tol=1e-20, max_it=10000, precond=1, gmg=None,
           criterion=1, threads=os.cpu_count()):
        x = np.zeros(self.n_dofs) if x is None else np.array(x, dtype=np.float64)
        res = C.c_double(0)
        it = lib().orc_cg(self.ref, ptr(np.ascontiguousarray(b)), ptr(x), rel_tol, abs_tol, max_it,
                          precond, gmg.h if gmg is not None else None, criterion, C.byref(res),
                          threads)
        return x, it, res.value


class GMG:
    def __init__(self, op, nc_coarse=4, coarse_degree=0, cheb_degree=3, smoothing_range=20.0,
                 eig_iters=20, threads=os.cpu_count()):
        self.op = op
        self.h = lib().orc_gmg_create(op.ref, nc_coarse, coarse_degree, cheb_degree,
                                      smoothing_range, eig_iters, threads)

    def vcycle(self, b):
        x = np.zeros(self.op.n_dofs)
        lib().orc_gmg_vcycle(self.h, ptr(x), ptr(np.ascontiguousarray(b)))
        return x

    def levels(self):
        out = []
        for l in range(lib().orc_gmg_n_levels(self.h)):
            deg = C.c_int()
            nc = (C.c_int * 3)()
            lam = C.c_double()
            lib().orc_gmg_level_info(self.h, l, C.byref(deg), nc, C.byref(lam))
            out.append((deg.value, tuple(nc), lam.value))
        return out

    def __del__(self):
        try:
            lib().orc_gmg_destroy(self.h)
        except Exception:
            pass


def kappa_law(law, p, T):
    out = np.zeros(2)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    lib().orc_kappa_law(law, ptr(pp), float(T), ptr(out))
    return out


def prolongate(kc, ncc, kf, ncf, coarse):
    fine = np.zeros((kf * ncf[0] + 1) * (kf * ncf[1] + 1) * (kf * ncf[2] + 1))
    lib().orc_prolongate(kc, (C.c_int * 3)(*ncc), kf, (C.c_int * 3)(*ncf),
                         ptr(np.ascontiguousarray(coarse)), ptr(fine))
    return fine


def restrict(kc, ncc, kf, ncf, fine):
    coarse = np.zeros((kc * ncc[0] + 1) * (kc * ncc[1] + 1) * (kc * ncc[2] + 1))
    lib().orc_restrict(kc, (C.c_int * 3)(*ncc), kf, (C.c_int * 3)(*ncf),
                       ptr(np.ascontiguousarray(fine)), ptr(coarse))
    return coarse


def heat_run(op, rho_cp, dt, n_steps, T0, predictor=2, lin_rel_tol=1e-12, nl_res_tol=1e-9,
             nl_corr_tol=1e-10, nl_max_it=10, threads=os.cpu_count()):
    T = np.array(T0, dtype=np.float64)
    norms = np.zeros(n_steps)
    its = (C.c_int * n_steps)()
    rc = lib().orc_heat_run(op.ref, rho_cp, dt, n_steps, predictor, lin_rel_tol, nl_res_tol,
                            nl_corr_tol, nl_max_it, ptr(T), ptr(norms), its, threads)
    return rc, T, norms, list(its)


def heat_run_tangent(op, n_steps, T0, predictor=2, lin_rel_tol=1e-12, nl_res_tol=1e-9, nl_corr_tol=1e-10,
                     nl_max_it=10, threads=os.cpu_count()):
    """Time loop with the full tangent (op.set_tangent first): Newton + GMRES + inverse diagonal."""
    T = np.array(T0, dtype=np.float64)
    norms = np.zeros(n_steps)
    its = (C.c_int * n_steps)()
    rc = lib().orc_heat_run_tangent(op.ref, n_steps, predictor, lin_rel_tol, nl_res_tol, nl_corr_tol, nl_max_it,
                                    ptr(T), ptr(norms), its, threads)
    return rc, T, norms, list(its)
