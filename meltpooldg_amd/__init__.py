"""MI355X-native matrix-free heat operator path of MeltPoolDG.

Python mirror of the reference's operator / preconditioner / solver interfaces,
calling the HIP library ``libmpdg_heat.so`` through its C ABI
(include/mpdg_heat.h). Vectors are torch tensors on the GPU (device memory and
streams only); every computation happens in the library's HIP kernels. There is
no CPU fallback: if the library or a GPU is missing, construction raises.

Reference interfaces mirrored (paths relative to the MeltPoolDG repository):
  HeatDiffuseMultiPhaseOperator  source/heat/heat_diffuse_operator.cpp:181-806
  OperatorMatrixFree             include/meltpooldg/core/operator_base.hpp:189-282
  JacobiPreconditioner           include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:17-89
  LinearSolver / LinearSolverData include/meltpooldg/linear_algebra/linear_solver.hpp:21-134,
                                 linear_solver_data.hpp:7-60
  HeatDiffuseOperation + Newton  source/heat/heat_diffuse_operation.cpp:215-298,
                                 source/linear_algebra/newton_raphson_solver.cpp:24-152
  Predictor                      include/meltpooldg/linear_algebra/predictor.hpp:118-160
"""
import ctypes as C
import os
from dataclasses import dataclass, field

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpdg_heat.so")

MPDG_OK = 0
ERRORS = {-1: "invalid argument", -2: "out of device memory", -3: "no convergence",
          -4: "non-finite value", -5: "HIP runtime error", -6: "zero time increment",
          -7: "mesh beyond the fixed scratch maximum"}
FP64, FP32 = 0, 1
KAPPA_CONST, KAPPA_QP, KAPPA_CELL, KAPPA_LAW = 0, 1, 2, 3
LAW_TANH, LAW_MUSHY = 1, 2
PRECOND_IDENTITY, PRECOND_DIAGONAL, PRECOND_GMG = 0, 1, 2
SOLVER_CG, SOLVER_FCG, SOLVER_GMRES = 0, 1, 2
TRANSITION_NONE, TRANSITION_MUSHY = 0, 1
CRIT_REDUCTION, CRIT_RHS = 0, 1
PREDICTOR_NONE, PREDICTOR_ZERO, PREDICTOR_LINEAR, PREDICTOR_LEARNED, PREDICTOR_EXTRAPOLATION, PREDICTOR_LSQ = \
    0, 1, 2, 3, 4, 5
PREDICTOR_LSQ_SOLUTION = 6


class MpdgError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed: %s (%d)" % (what, ERRORS.get(code, "error"), code))
        self.code = code


class NoConvergence(MpdgError):
    """SolverControl::NoConvergence / ExcNewtonDidNotConverge."""


class HeatDesc(C.Structure):
    _fields_ = [("degree", C.c_int), ("n_cells", C.c_int * 3), ("extent", C.c_double * 3),
                ("precision", C.c_int), ("dirichlet_faces", C.c_int), ("rho_cp", C.c_double)]


class GMGParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("n_cells_coarse", C.c_int), ("coarse_degree", C.c_int),
                ("smoother_degree", C.c_int), ("smoothing_range", C.c_double),
                ("eig_cg_n_iterations", C.c_int), ("pre_smoother_degree", C.c_int), ("vcycle_graph", C.c_int)]


class SolverParams(C.Structure):
    _fields_ = [("solver_type", C.c_int), ("max_iterations", C.c_int), ("rel_tolerance", C.c_double),
                ("abs_tolerance", C.c_double), ("criterion", C.c_int), ("gmres_restart", C.c_int)]


class Material(C.Structure):
    """mpdg_heat_material: liquid/solid values (index 0 liquid, 1 solid) and the mushy interval."""
    _fields_ = [("transition", C.c_int), ("conductivity", C.c_double * 2), ("specific_heat", C.c_double * 2),
                ("density", C.c_double * 2), ("solidus", C.c_double), ("liquidus", C.c_double)]


class OperationParams(C.Structure):
    _fields_ = [("predictor", C.c_int), ("linear", SolverParams), ("max_nonlinear_iterations", C.c_int),
                ("residual_tolerance", C.c_double), ("field_correction_tolerance", C.c_double),
                ("step_mode", C.c_int), ("guess_check", C.c_int), ("guess_smoothing", C.c_int),
                ("extrapolation_order", C.c_int), ("n_old_solution_vectors", C.c_int),
                ("preconditioner_update_interval", C.c_int), ("guess_pause_after", C.c_int),
                ("guess_pause_steps", C.c_int)]


STEP_NEWTON, STEP_LINEAR = 0, 1
GUESS_USED, GUESS_REJECTED, GUESS_PAUSED = 0, 1, 2


_V, _I, _D, _L = C.c_void_p, C.c_int, C.c_double, C.c_long
_SIGNATURES = {
    "mpdg_heat_op_create": ([C.POINTER(HeatDesc), C.POINTER(_V)], _I),
    "mpdg_heat_op_set_tangent": ([_V, C.POINTER(Material), _I, C.POINTER(_D)], _I),
    "mpdg_io_write_vtu": ([C.c_char_p, _I, C.POINTER(_I), C.POINTER(_D), _I, C.POINTER(C.c_char_p),
                           C.POINTER(C.POINTER(_D))], _I),
    "mpdg_io_write_pvd": ([C.c_char_p, _I, C.POINTER(_D), C.POINTER(C.c_char_p)], _I),
    "mpdg_io_dof_permutation": ([_I, C.POINTER(_I), C.POINTER(_D), C.POINTER(_D), _L, C.POINTER(_L)], _I),
    "mpdg_io_permute": ([_V, _V, _V, _L, _I, _V], _I),
    "mpdg_heat_op_set_velocity_field": ([_V, _V], _I),
    "mpdg_heat_op_clear_tangent": ([_V], _I),
    "mpdg_heat_op_bind_temperature": ([_V, _V, _V, _V], _I),
    "mpdg_heat_op_destroy": ([_V], _I),
    "mpdg_heat_op_n_dofs": ([_V], _L),
    "mpdg_heat_op_n_cells": ([_V], _L),
    "mpdg_heat_op_reset_time_increment": ([_V, _D], _I),
    "mpdg_heat_op_set_conductivity_constant": ([_V, _D], _I),
    "mpdg_heat_op_set_conductivity_qp": ([_V, _V], _I),
    "mpdg_heat_op_set_conductivity_cell": ([_V, _V], _I),
    "mpdg_heat_op_set_conductivity_law": ([_V, _I, C.POINTER(_D), _V], _I),
    "mpdg_heat_op_set_robin": ([_V, _I, _D, _D, _D, _D, _V], _I),
    "mpdg_heat_op_set_surface_flux": ([_V, C.POINTER(_D)], _I),
    "mpdg_heat_op_set_mapping": ([_V, _V], _I),
    "mpdg_heat_op_vmult": ([_V, _V, _V, _V], _I),
    "mpdg_heat_op_vmult_f32": ([_V, _V, _V, _V], _I),
    "mpdg_heat_op_vmult_noalloc": ([_V, _V, _V, _V], _I),
    "mpdg_heat_op_compute_inverse_diagonal": ([_V, _V, _V], _I),
    "mpdg_heat_op_compute_diagonal": ([_V, _V, _V], _I),
    "mpdg_heat_op_create_rhs": ([_V, _V, _V, _V, _V, _V], _I),
    "mpdg_heat_op_l2_norm": ([_V, _V, C.POINTER(_D), _V], _I),
    "mpdg_heat_op_set_timing": ([_V, _I], _I),
    "mpdg_heat_op_get_timing": ([_V, C.POINTER(_D), C.POINTER(_I)], _I),
    "mpdg_precond_create": ([_V, _I, C.POINTER(GMGParams), C.POINTER(_V)], _I),
    "mpdg_precond_update": ([_V, _V], _I),
    "mpdg_precond_vmult": ([_V, _V, _V, _V], _I),
    "mpdg_precond_destroy": ([_V], _I),
    "mpdg_gmg_n_levels": ([_V], _I),
    "mpdg_gmg_level_info": ([_V, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)], _I),
    "mpdg_linear_solve": ([_V, _V, C.POINTER(SolverParams), _V, _V, C.POINTER(_I), C.POINTER(_D), _V], _I),
    "mpdg_heat_operation_create": ([_V, _V, C.POINTER(OperationParams), C.POINTER(_V)], _I),
    "mpdg_heat_operation_destroy": ([_V], _I),
    "mpdg_heat_operation_set_initial": ([_V, _V, _V], _I),
    "mpdg_heat_operation_solve_step": ([_V, _D, C.POINTER(_I), C.POINTER(_I), _V], _I),
    "mpdg_heat_operation_solution": ([_V], _V),
    "mpdg_heat_operation_old_solution": ([_V], _V),
    "mpdg_heat_operation_set_learned": ([_V, _V, _V], _I),
    "mpdg_heat_operation_record_guess_feature": ([_V, _V, _V], _I),
    "mpdg_heat_operation_last_step_info": ([_V, C.POINTER(_I), C.POINTER(_D), C.POINTER(_D)], _I),
    "mpdg_heat_op_surface_flux_field": ([_V, _V, _V], _I),
    "mpdg_stream_probe": ([_I, _V, _V, _L, _V], _I),
    "mpdg_learned_create": ([C.POINTER(_I), _I, _I, _D, C.c_ulonglong, C.POINTER(_V)], _I),
    "mpdg_learned_grid": ([_V, C.POINTER(_I), C.POINTER(_I)], _I),
    "mpdg_learned_get_output_scale": ([_V, C.POINTER(_D)], _I),
    "mpdg_learned_set_output_scale": ([_V, _D], _I),
    "mpdg_learned_destroy": ([_V], _I),
    "mpdg_learned_n_params": ([_V], _I),
    "mpdg_learned_get_weights": ([_V, _V], _I),
    "mpdg_learned_set_weights": ([_V, _V], _I),
    "mpdg_learned_predict": ([_V, _V, _V, _V, _V, _V, _V, _V], _I),
    "mpdg_learned_train_step": ([_V, _V, _V, _V, _V, _V, _V, _D, C.POINTER(_D), _V], _I),
    "mpdg_learned_gradients": ([_V, _V, _V, _V, _V, _V, _V, _V, C.POINTER(_D), _V], _I),
    "mpdg_learned_activation": ([_V, _I, _V], _I),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libmpdg_heat.so and bind every exported entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError("libmpdg_heat.so not built: run __graft_entry__.build() (%s)" % path)
    # torch first: its HIP runtime (soname libamdhip64.so.7) is then the one the library binds to.
    # Loaded the other way round, the library pulls in the system runtime and torch later loads
    # its own copy, and device work of the two does not mix (the first sync reports an error).
    import torch  # noqa: F401
    L = C.CDLL(path)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def node_coordinates(degree, n_cells, extent=(1.0, 1.0, 1.0), mapping_amplitude=0.0):
    """Coordinates of the Q_degree nodes (Gauss-Lobatto points in every cell, lexicographic) as one
    device FP64 vector (X block, Y block, Z block), the layout set_mapping takes. With
    mapping_amplitude a, every node is moved to x + a sin(2 pi y / L_y) sin(2 pi z / L_z) (on the
    unit cube the curved box of the mapped benchmark mesh, x + a sin(2 pi y) sin(2 pi z))."""
    import numpy as np
    torch = _torch()
    if degree == 1:
        gl = np.array([-1.0, 1.0])
    else:
        P = np.polynomial.legendre.Legendre.basis(degree).deriv()
        gl = np.concatenate([[-1.0], np.sort(P.roots().real), [1.0]])
    gl = 0.5 * (gl + 1.0)
    axes = []
    for d in range(3):
        h = extent[d] / n_cells[d]
        pts = np.concatenate([c * h + gl[:-1] * h for c in range(n_cells[d])] + [np.array([extent[d]])])
        axes.append(torch.tensor(pts, dtype=torch.float64, device="cuda"))
    Z, Y, X = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    if mapping_amplitude:
        X = X + mapping_amplitude * torch.sin(2 * np.pi * Y / extent[1]) * torch.sin(2 * np.pi * Z / extent[2])
    return torch.cat([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)]).contiguous()


def stream_probe(kind, dst, src=None):
    """Bandwidth probe on device FP64 vectors: 0 copy src -> dst, 1 read src (dst >= 4096
    entries receives partial sums), 2 write dst. Returns nothing; time it with events."""
    L = load_library()
    n = src.numel() if kind != 2 else dst.numel()
    _check(L.mpdg_stream_probe(kind, _ptr(dst), _ptr(src), n, _stream()), "stream probe")


def exported_symbols():
    return sorted(_SIGNATURES)


def _check(rc, what):
    if rc != MPDG_OK:
        if rc == -3:
            raise NoConvergence(rc, what)
        raise MpdgError(rc, what)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("meltpooldg_amd needs a GPU: torch.cuda.is_available() is False")
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t, dtype=None):
    if t is None:
        return None
    if dtype is not None and t.dtype != dtype:
        raise TypeError("expected %s, got %s" % (dtype, t.dtype))
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous GPU tensor")
    return C.c_void_p(t.data_ptr())


class HeatDiffuseMultiPhaseOperator:
    """J = (rho c_p / dt) M + K_k (+ Robin faces) on a structured Q_k hex mesh.

    Mirrors OperatorMatrixFree<3,double>: vmult, create_rhs,
    compute_inverse_diagonal_from_matrixfree, reset_time_increment.
    Attached arrays (conductivity, temperature) stay owned by the caller; this
    object keeps a reference so they outlive the operator.
    """

    def __init__(self, degree, n_cells, extent=(1.0, 1.0, 1.0), rho_cp=1.0, precision=FP64,
                 dirichlet_faces=0):
        self.lib = load_library()
        _torch()
        d = HeatDesc()
        d.degree = degree
        d.n_cells[:] = list(n_cells)
        d.extent[:] = list(extent)
        d.precision = precision
        d.dirichlet_faces = dirichlet_faces
        d.rho_cp = rho_cp
        h = C.c_void_p()
        _check(self.lib.mpdg_heat_op_create(C.byref(d), C.byref(h)), "mpdg_heat_op_create")
        self.h = h
        self.degree, self.n_cells, self.extent, self.precision = degree, tuple(n_cells), tuple(extent), precision
        self.n_dofs = self.lib.mpdg_heat_op_n_dofs(h)
        self.n_cells_total = self.lib.mpdg_heat_op_n_cells(h)
        self.n_q_points = self.n_cells_total * (degree + 1) ** 3
        self._keep = {}

    @property
    def dtype(self):
        torch = _torch()
        return torch.float64 if self.precision == FP64 else torch.float32

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.mpdg_heat_op_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def reset_time_increment(self, dt):
        _check(self.lib.mpdg_heat_op_reset_time_increment(self.h, dt), "reset_time_increment")

    def set_conductivity_constant(self, k):
        _check(self.lib.mpdg_heat_op_set_conductivity_constant(self.h, k), "set_conductivity")

    def set_conductivity_qp(self, k_qp):
        assert k_qp.numel() == self.n_q_points
        self._keep["k"] = k_qp
        _check(self.lib.mpdg_heat_op_set_conductivity_qp(self.h, _ptr(k_qp, self.dtype)), "set_conductivity_qp")

    def set_conductivity_cell(self, k_cell):
        assert k_cell.numel() == self.n_cells_total
        self._keep["k"] = k_cell
        _check(self.lib.mpdg_heat_op_set_conductivity_cell(self.h, _ptr(k_cell, self.dtype)),
               "set_conductivity_cell")

    def set_conductivity_law(self, law, params, T):
        assert T.numel() == self.n_dofs
        self._keep["T"] = T
        p = (C.c_double * 4)(*params)
        _check(self.lib.mpdg_heat_op_set_conductivity_law(self.h, law, p, _ptr(T, self.dtype)),
               "set_conductivity_law")

    def set_robin(self, face_mask, alpha=0.0, T_inf_conv=0.0, emissivity=0.0, T_inf_rad=0.0, T=None):
        self._keep["Trobin"] = T
        _check(self.lib.mpdg_heat_op_set_robin(self.h, face_mask, alpha, T_inf_conv, emissivity, T_inf_rad,
                                               _ptr(T)), "set_robin")

    def set_tangent(self, material=None, velocity=(0.0, 0.0, 0.0), do_solidification=False):
        """Full Newton tangent (heat_diffuse_operator.cpp:1028-1056) with a prescribed velocity.
        material: dict(conductivity=(l, s), specific_heat=(l, s), density=(l, s), solidus=.., liquidus=..)
        for the liquid/solid mushy zone, or dict(conductivity=k, specific_heat=c, density=rho) constant."""
        m = Material()
        mat = material or {}
        def pair(v):
            return (v, v) if not isinstance(v, (tuple, list)) else tuple(v)
        m.transition = TRANSITION_MUSHY if "solidus" in mat else TRANSITION_NONE
        m.conductivity[:] = pair(mat.get("conductivity", 1.0))
        m.specific_heat[:] = pair(mat.get("specific_heat", 1.0))
        m.density[:] = pair(mat.get("density", 1.0))
        m.solidus, m.liquidus = mat.get("solidus", 0.0), mat.get("liquidus", 1.0)
        v = (C.c_double * 3)(*velocity)
        _check(self.lib.mpdg_heat_op_set_tangent(self.h, C.byref(m), 1 if do_solidification else 0, v), "set_tangent")

    def set_velocity_field(self, velocity):
        """Velocity DoF field for the tangent (heat_diffuse_operator.cpp:975-978): a float64 GPU tensor of
        3 n_dofs values (x, y, z blocks) on the temperature's nodes, or None for the constant vector of
        set_tangent. Read by every bind_temperature and create_rhs; kept alive by this object."""
        if velocity is not None and velocity.numel() != 3 * self.n_dofs:
            raise ValueError("velocity field: expected 3 * n_dofs values")
        self._keep["velocity"] = velocity
        _check(self.lib.mpdg_heat_op_set_velocity_field(self.h, _ptr(velocity, _torch().float64)),
               "set_velocity_field")

    def clear_tangent(self):
        self._keep.pop("velocity", None)
        _check(self.lib.mpdg_heat_op_clear_tangent(self.h), "clear_tangent")

    def bind_temperature(self, T, T_old=None):
        """Evaluate the tangent at T (and T_old) -- call after T changes, before vmult / diagonal."""
        _check(self.lib.mpdg_heat_op_bind_temperature(self.h, _ptr(T), _ptr(T_old) if T_old is not None else None,
                                                      _stream()), "bind_temperature")

    def set_mapping(self, coords):
        """Mapped mesh: node coordinates (device FP64, X block, Y block, Z block)."""
        assert coords.numel() == 3 * self.n_dofs
        self._keep["X"] = coords
        _check(self.lib.mpdg_heat_op_set_mapping(self.h, _ptr(coords, _torch().float64)), "set_mapping")

    def set_surface_flux(self, peak, sigma, x0, y0, vx, t):
        p = (C.c_double * 6)(peak, sigma, x0, y0, vx, t)
        _check(self.lib.mpdg_heat_op_set_surface_flux(self.h, p), "set_surface_flux")

    def vmult(self, dst, src):
        if self.precision == FP64:
            _check(self.lib.mpdg_heat_op_vmult(self.h, _ptr(dst, self.dtype), _ptr(src, self.dtype), _stream()),
                   "vmult")
        else:
            _check(self.lib.mpdg_heat_op_vmult_f32(self.h, _ptr(dst, self.dtype), _ptr(src, self.dtype),
                                                   _stream()), "vmult_f32")

    def compute_inverse_diagonal_from_matrixfree(self, diag):
        _check(self.lib.mpdg_heat_op_compute_inverse_diagonal(self.h, _ptr(diag), _stream()),
               "compute_inverse_diagonal")

    def compute_diagonal(self, diag):
        _check(self.lib.mpdg_heat_op_compute_diagonal(self.h, _ptr(diag, self.dtype), _stream()),
               "compute_diagonal")

    def create_rhs(self, dst, T, T_old, q=None):
        _check(self.lib.mpdg_heat_op_create_rhs(self.h, _ptr(dst), _ptr(T), _ptr(T_old), _ptr(q), _stream()),
               "create_rhs")

    def l2_norm(self, u):
        out = C.c_double()
        _check(self.lib.mpdg_heat_op_l2_norm(self.h, _ptr(u), C.byref(out), _stream()), "l2_norm")
        return out.value

    def set_timing(self, enable=True):
        """Record HIP events around every cell-apply kernel launch of this operator."""
        _check(self.lib.mpdg_heat_op_set_timing(self.h, 1 if enable else 0), "set_timing")

    def get_timing(self):
        """(summed kernel milliseconds, number of launches) since set_timing."""
        ms, n = C.c_double(), C.c_int()
        _check(self.lib.mpdg_heat_op_get_timing(self.h, C.byref(ms), C.byref(n)), "get_timing")
        return ms.value, n.value

    def surface_flux_field(self, out):
        """Nodal values of the z+ surface flux set by set_surface_flux (zero off that face)."""
        _check(self.lib.mpdg_heat_op_surface_flux_field(self.h, _ptr(out, _torch().float64), _stream()),
               "surface_flux_field")

    def new_vector(self):
        torch = _torch()
        return torch.zeros(self.n_dofs, dtype=self.dtype, device="cuda")


class Preconditioner:
    """Type-erased preconditioner (include/meltpooldg/linear_algebra/preconditioner.hpp:41-203)."""

    def __init__(self, op, type_, gmg=None):
        self.lib = op.lib
        self.op = op
        h = C.c_void_p()
        _check(self.lib.mpdg_precond_create(op.h, type_, C.byref(gmg) if gmg is not None else None, C.byref(h)),
               "mpdg_precond_create")
        self.h = h
        self.type = type_

    def update(self):
        _check(self.lib.mpdg_precond_update(self.h, _stream()), "precond update")

    def vmult(self, dst, src):
        _check(self.lib.mpdg_precond_vmult(self.h, _ptr(dst), _ptr(src), _stream()), "precond vmult")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.mpdg_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


class JacobiPreconditioner(Preconditioner):
    def __init__(self, op):
        super().__init__(op, PRECOND_DIAGONAL)


class GMGPreconditioner(Preconditioner):
    """Geometric multigrid V-cycle (recipe of unit_tests/gmg.cc) with FP32 or FP64 levels."""

    def __init__(self, op, precision=FP32, n_cells_coarse=4, coarse_degree=0, smoother_degree=3,
                 smoothing_range=20.0, eig_cg_n_iterations=20, pre_smoother_degree=0, vcycle_graph=True):
        p = GMGParams()
        p.precision, p.n_cells_coarse, p.coarse_degree = precision, n_cells_coarse, coarse_degree
        p.smoother_degree, p.smoothing_range, p.eig_cg_n_iterations = smoother_degree, smoothing_range, \
            eig_cg_n_iterations
        p.pre_smoother_degree = pre_smoother_degree
        p.vcycle_graph = 1 if vcycle_graph else 0
        super().__init__(op, PRECOND_GMG, p)

    def levels(self):
        out = []
        for l in range(self.lib.mpdg_gmg_n_levels(self.h)):
            deg, nc, lam = C.c_int(), (C.c_int * 3)(), C.c_double()
            _check(self.lib.mpdg_gmg_level_info(self.h, l, C.byref(deg), nc, C.byref(lam)), "level_info")
            out.append((deg.value, tuple(nc), lam.value))
        return out


@dataclass
class LinearSolverData:
    """include/meltpooldg/linear_algebra/linear_solver_data.hpp:36-56 (CG on this path)."""
    solver_type: int = SOLVER_CG
    max_iterations: int = 10000
    rel_tolerance: float = 1e-12
    abs_tolerance: float = 1e-20
    criterion: int = CRIT_REDUCTION
    gmres_restart: int = 0

    def c(self):
        p = SolverParams()
        p.solver_type, p.max_iterations = self.solver_type, self.max_iterations
        p.rel_tolerance, p.abs_tolerance, p.criterion = self.rel_tolerance, self.abs_tolerance, self.criterion
        p.gmres_restart = self.gmres_restart
        return p


class LinearSolver:
    @staticmethod
    def solve(op, solution, rhs, data=None, preconditioner=None):
        """Returns the iteration count (solver_control.last_step()); raises NoConvergence."""
        data = data or LinearSolverData()
        it, res = C.c_int(), C.c_double()
        rc = op.lib.mpdg_linear_solve(op.h, preconditioner.h if preconditioner is not None else None,
                                      C.byref(data.c()), _ptr(rhs), _ptr(solution), C.byref(it), C.byref(res),
                                      _stream())
        LinearSolver.last_residual = res.value
        _check(rc, "LinearSolver::solve")
        return it.value


@dataclass
class NonlinearSolverData:
    """include/meltpooldg/linear_algebra/nonlinear_solver_data.hpp:13-19"""
    max_nonlinear_iterations: int = 10
    field_correction_tolerance: float = 1e-10
    residual_tolerance: float = 1e-9


class HeatDiffuseOperation:
    """Implicit heat time step: predictor, preconditioner update, Newton-Raphson with CG."""

    def __init__(self, op, preconditioner=None, predictor=PREDICTOR_LINEAR, linear=None, nlsolve=None,
                 step_mode=STEP_NEWTON, learned=None, q_field=None, guess_check=True, guess_smoothing=0,
                 extrapolation_order=2, n_old_solution_vectors=2, preconditioner_update_interval=1,
                 guess_pause_after=0, guess_pause_steps=0):
        self.op = op
        self.lib = op.lib
        self.pc = preconditioner
        self.learned, self.q_field = learned, q_field
        linear = linear or LinearSolverData()
        nlsolve = nlsolve or NonlinearSolverData()
        p = OperationParams()
        p.predictor = predictor
        p.linear = linear.c()
        p.max_nonlinear_iterations = nlsolve.max_nonlinear_iterations
        p.residual_tolerance = nlsolve.residual_tolerance
        p.field_correction_tolerance = nlsolve.field_correction_tolerance
        p.step_mode = step_mode
        p.guess_check = 1 if guess_check else 0
        p.guess_smoothing = guess_smoothing
        p.extrapolation_order = extrapolation_order
        p.n_old_solution_vectors = n_old_solution_vectors
        p.preconditioner_update_interval = preconditioner_update_interval
        p.guess_pause_after = guess_pause_after
        p.guess_pause_steps = guess_pause_steps
        h = C.c_void_p()
        _check(self.lib.mpdg_heat_operation_create(op.h, preconditioner.h if preconditioner else None,
                                                   C.byref(p), C.byref(h)), "heat_operation_create")
        self.h = h
        if learned is not None:
            _check(self.lib.mpdg_heat_operation_set_learned(h, learned.h, _ptr(q_field)), "set_learned")

    def record_guess_feature(self, z, base=None):
        """From the next step on, write each step's network residual input D^-1 (b - J T_b) into
        the device vector z and the base guess T_b (the predictor's extrapolation) into base
        (None stops / skips); training pairs need them."""
        self.z_record = (z, base)
        _check(self.lib.mpdg_heat_operation_record_guess_feature(self.h, _ptr(z), _ptr(base)),
               "record_guess_feature")

    def last_step_info(self):
        """(guess status, initial and final ||b - J T||/||b|| of the step); the status is
        GUESS_USED (0), GUESS_REJECTED (1: the learned guess lost the guess check and its base
        guess was used) or GUESS_PAUSED (2: the network was not evaluated after repeated
        rejections, see guess_pause_after)"""
        fb, r0, r = C.c_int(), C.c_double(), C.c_double()
        _check(self.lib.mpdg_heat_operation_last_step_info(self.h, C.byref(fb), C.byref(r0), C.byref(r)), "info")
        return fb.value, r0.value, r.value

    def old_solution(self):
        torch = _torch()
        out = torch.empty(self.op.n_dofs, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _copy_device(out, self.lib.mpdg_heat_operation_old_solution(self.h), self.op.n_dofs)
        return out

    def set_initial_condition(self, T0):
        _check(self.lib.mpdg_heat_operation_set_initial(self.h, _ptr(T0), _stream()), "set_initial")

    def solve(self, dt):
        li, ni = C.c_int(), C.c_int()
        _check(self.lib.mpdg_heat_operation_solve_step(self.h, dt, C.byref(li), C.byref(ni), _stream()),
               "HeatDiffuseOperation::solve")
        return li.value, ni.value

    def solution(self):
        """A copy of the current temperature (device)."""
        torch = _torch()
        out = torch.empty(self.op.n_dofs, dtype=torch.float64, device="cuda")
        ptr = self.lib.mpdg_heat_operation_solution(self.h)
        torch.cuda.synchronize()
        _copy_device(out, ptr, self.op.n_dofs)
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.mpdg_heat_operation_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LearnedPredictor:
    """Learned initial guess: 3D CNN predicting the correction of a base guess T_b (the linear
    or a higher-order extrapolation of the history) for T_{n+1} from (T_{n-1}, T_n, q) and
    z = D^-1 (b - J T_b), the scaled residual of the base guess for the step.

    Fields are DoF vectors of Q_degree on n_cells cells; the network works on the DoF grid of
    Q_degree on n_cells / coarsening cells (full weighting there, FE prolongation back).
    A new predictor type for the slot of include/meltpooldg/linear_algebra/predictor.hpp:118-160;
    training (Adam, FP32 masters) and inference run in the library's BF16 MFMA kernels.
    """

    def __init__(self, n_cells, degree=1, coarsening=1, qscale=1.0, seed=1234):
        self.lib = load_library()
        _torch()
        h = C.c_void_p()
        nc = (C.c_int * 3)(*n_cells)
        _check(self.lib.mpdg_learned_create(nc, degree, coarsening, qscale, seed, C.byref(h)),
               "mpdg_learned_create")
        self.h = h
        fd, nd = (C.c_int * 3)(), (C.c_int * 3)()
        _check(self.lib.mpdg_learned_grid(h, fd, nd), "mpdg_learned_grid")
        self.dof_dims, self.dims = tuple(fd), tuple(nd)
        self.n_params = self.lib.mpdg_learned_n_params(h)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.mpdg_learned_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def predict(self, T_nm1, T_n, q, z, out, base=None):
        """out = T_b + rms(z) y with T_b = base, or 2 T_n - T_{n-1} (device FP64 vectors; q, z,
        base may be None)."""
        _check(self.lib.mpdg_learned_predict(self.h, _ptr(T_nm1), _ptr(T_n), _ptr(q), _ptr(z), _ptr(base),
                                             _ptr(out), _stream()), "learned predict")

    def train_step(self, T_nm1, T_n, q, z, T_np1, lr, base=None):
        loss = C.c_double()
        _check(self.lib.mpdg_learned_train_step(self.h, _ptr(T_nm1), _ptr(T_n), _ptr(q), _ptr(z), _ptr(base),
                                                _ptr(T_np1), lr, C.byref(loss), _stream()), "learned train_step")
        return loss.value

    def gradients(self, T_nm1, T_n, q, z, T_np1, base=None):
        import numpy as np
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        _check(self.lib.mpdg_learned_gradients(self.h, _ptr(T_nm1), _ptr(T_n), _ptr(q), _ptr(z), _ptr(base),
                                               _ptr(T_np1), g.ctypes.data_as(C.c_void_p), C.byref(loss),
                                               _stream()), "learned gradients")
        return loss.value, g

    def get_weights(self):
        import numpy as np
        w = np.zeros(self.n_params, dtype=np.float32)
        _check(self.lib.mpdg_learned_get_weights(self.h, w.ctypes.data_as(C.c_void_p)), "get_weights")
        return w

    def set_weights(self, w):
        import numpy as np
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.n_params
        _check(self.lib.mpdg_learned_set_weights(self.h, w.ctypes.data_as(C.c_void_p)), "set_weights")

    @property
    def output_scale(self):
        v = C.c_double()
        _check(self.lib.mpdg_learned_get_output_scale(self.h, C.byref(v)), "get_output_scale")
        return v.value

    @output_scale.setter
    def output_scale(self, v):
        _check(self.lib.mpdg_learned_set_output_scale(self.h, float(v)), "set_output_scale")

    def save(self, path):
        """Checkpoint: the FP32 weights then the FP64 output scale, raw little-endian."""
        import numpy as np
        with open(path, "wb") as f:
            f.write(self.get_weights().tobytes())
            f.write(np.float64(self.output_scale).tobytes())

    def load(self, path):
        import numpy as np
        raw = np.fromfile(path, dtype=np.uint8)
        nw = 4 * self.n_params
        self.set_weights(raw[:nw].view(np.float32))
        self.output_scale = float(raw[nw:nw + 8].view(np.float64)[0])


def _copy_device(dst, src_ptr, n):
    """dst (torch, fp64) <- n doubles at the device address src_ptr."""
    import torch
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = hip.hipMemcpy(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr), n * 8, 3)
    if rc != 0:
        raise MpdgError(-5, "hipMemcpy")
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# Data formats: Paraview output (Postprocessor::do_paraview_output,
# source/post_processing/postprocessor.cpp:195-270) and the DoF numbering of a
# DoFHandler vs. the lexicographic node grid.
# ---------------------------------------------------------------------------
def write_vtu(filename, degree, n_cells, extent, fields):
    """fields: {name: lexicographic DoF values (numpy array or tensor, copied to the host)}."""
    import numpy as np
    lib = load_library()
    names = list(fields)
    arrays = []
    for nm in names:
        v = fields[nm]
        if hasattr(v, "detach"):
            v = v.detach().cpu().numpy()
        arrays.append(np.ascontiguousarray(v, dtype=np.float64))
    c_names = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    c_fields = (C.POINTER(_D) * max(1, len(arrays)))(*[a.ctypes.data_as(C.POINTER(_D)) for a in arrays])
    _check(lib.mpdg_io_write_vtu(str(filename).encode(), degree, (_I * 3)(*n_cells), (_D * 3)(*extent), len(names),
                                 c_names, c_fields), "write_vtu")


def write_pvd(filename, times, vtu_names):
    lib = load_library()
    n = len(times)
    _check(lib.mpdg_io_write_pvd(str(filename).encode(), n, (_D * max(1, n))(*times),
                                 (C.c_char_p * max(1, n))(*[str(v).encode() for v in vtu_names])), "write_pvd")


class ParaviewOutput:
    """One .vtu per output step and a .pvd record (the reference's paraview output: file names
    <filename>_<step with n_digits>.vtu, fields temperature / temperature_old / heat_source)."""

    def __init__(self, directory, op, filename="solution", n_digits=4):
        import os
        os.makedirs(directory, exist_ok=True)
        self.dir, self.op, self.filename, self.n_digits = directory, op, filename, n_digits
        self.records = []

    def write(self, step, time, fields):
        import os
        name = "%s_%0*d.vtu" % (self.filename, self.n_digits, step)
        write_vtu(os.path.join(self.dir, name), self.op.degree, self.op.n_cells, self.op.extent, fields)
        if not self.records or self.records[-1][1] != name:
            self.records.append((time, name))
        write_pvd(os.path.join(self.dir, self.filename + ".pvd"), [t for t, _ in self.records],
                  [n for _, n in self.records])
        return name


def dof_permutation(degree, n_cells, extent, points):
    """lex_index[i] = lexicographic node of DoF i, from a DoFHandler's support points (n x 3)."""
    import numpy as np
    lib = load_library()
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(pts), dtype=np.int64)
    _check(lib.mpdg_io_dof_permutation(degree, (_I * 3)(*n_cells), (_D * 3)(*extent), pts.ctypes.data_as(C.POINTER(_D)),
                                       len(pts), out.ctypes.data_as(C.POINTER(_L))), "dof_permutation")
    return out


def permute(lex_index, src, dst, to_lex=True):
    """Device vectors: to_lex: dst[lex_index[i]] = src[i]; else dst[i] = src[lex_index[i]]."""
    lib = load_library()
    _check(lib.mpdg_io_permute(_ptr(lex_index, _torch().int64), _ptr(src), _ptr(dst), src.numel(),
                               1 if to_lex else 0, _stream()), "permute")
