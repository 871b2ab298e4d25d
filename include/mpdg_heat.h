/*
 * mpdg_heat.h -- C ABI of the MI355X-native matrix-free heat operator path.
 *
 * Drop-in for MeltPoolDG's implicit heat operator, its preconditioners and the
 * Krylov/Newton/predictor stack that drives it (citations are relative to the
 * MeltPoolDG repository root). Plain C: device pointers are `void*`/typed
 * pointers to device memory, streams are hipStream_t passed as `void*`
 * (NULL = the default stream). Every call returns MPDG_OK (0) or a negative
 * MPDG_ERR_* code; no exception crosses this boundary.
 *
 * Discretisation: structured hexahedral mesh of n_cells[0] x n_cells[1] x
 * n_cells[2] cells on [0,extent[0]] x [0,extent[1]] x [0,extent[2]],
 * continuous Q_k (k = 1..4) on Gauss-Lobatto nodes, QGauss(k+1). DoF vectors
 * are lexicographic: i = ix + nx*(iy + ny*iz), nx = k*n_cells[0]+1. Per-qp
 * arrays are cell-major: cell*(k+1)^3 + qx + (k+1)*(qy + (k+1)*qz) with
 * cell = cx + n_cells[0]*(cy + n_cells[1]*cz). Faces: bit 0 x-, 1 x+, 2 y-,
 * 3 y+, 4 z-, 5 z+ (deal.II's colorized hyper_rectangle boundary ids).
 * Handles are not internally synchronised: use one handle per stream.
 */
#ifndef MPDG_HEAT_H
#define MPDG_HEAT_H

#ifdef __cplusplus
extern "C" {
#endif

#define MPDG_OK 0
#define MPDG_ERR_INVALID (-1)          /* invalid argument / unsupported configuration */
#define MPDG_ERR_NOMEM (-2)            /* device allocation failed */
#define MPDG_ERR_NO_CONVERGENCE (-3)   /* SolverControl::NoConvergence / ExcNewtonDidNotConverge */
#define MPDG_ERR_NONFINITE (-4)        /* NaN/Inf met in a residual */
#define MPDG_ERR_DEVICE (-5)           /* HIP runtime error */
#define MPDG_ERR_ZERO_TIME_INCREMENT (-6) /* ExcZeroTimeIncrement (heat_diffuse_operator.cpp:788) */
#define MPDG_ERR_TOO_LARGE (-7)        /* mesh beyond a fixed maximum (mpdg_heat_op_vmult_noalloc) */

#define MPDG_FP64 0
#define MPDG_FP32 1

#define MPDG_KAPPA_CONST 0 /* one conductivity for the whole mesh */
#define MPDG_KAPPA_QP 1    /* one value per quadrature point */
#define MPDG_KAPPA_CELL 2  /* one value per cell */
#define MPDG_KAPPA_LAW 3   /* k(T) from a temperature DoF vector at every quadrature point */

#define MPDG_LAW_TANH 1  /* k0*(1 + a*tanh((T - T0)/w)), params {k0, a, T0, w} */
#define MPDG_LAW_MUSHY 2 /* liquid_solid mushy zone, params {k_liquid, k_solid, T_solidus, T_liquidus}
                            (include/meltpooldg/core/material.templates.hpp:296-330, :546-598, :765-778) */

typedef struct mpdg_heat_desc {
  int degree;          /* k, 1..4 */
  int n_cells[3];
  double extent[3];
  int precision;       /* MPDG_FP64 or MPDG_FP32 (the number type of the operator's vectors) */
  int dirichlet_faces; /* homogeneous Dirichlet faces (bitmask); constrained rows of J are 0 */
  double rho_cp;       /* volumetric heat capacity rho*c_p */
} mpdg_heat_desc;

typedef struct mpdg_heat_op mpdg_heat_op;

/* --- operator: HeatDiffuseMultiPhaseOperator<3,double> as an OperatorMatrixFree
       (include/meltpooldg/core/operator_base.hpp:189-282) --- */
int mpdg_heat_op_create(const mpdg_heat_desc *desc, mpdg_heat_op **op);
int mpdg_heat_op_destroy(mpdg_heat_op *op);
long mpdg_heat_op_n_dofs(const mpdg_heat_op *op);
long mpdg_heat_op_n_cells(const mpdg_heat_op *op);
/* OperatorBase::reset_time_increment (operator_base.hpp:42-46): mass coefficient rho_cp/dt */
int mpdg_heat_op_reset_time_increment(mpdg_heat_op *op, double dt);
/* conductivity data (Material::compute_parameters, heat_diffuse_operator.cpp:1172-1210);
   arrays are device memory of the operator's precision and stay owned by the caller */
int mpdg_heat_op_set_conductivity_constant(mpdg_heat_op *op, double k);
/* k_qp: [cell][qz][qy][qx] per quadrature point, 16-byte aligned (MPDG_ERR_INVALID otherwise) */
int mpdg_heat_op_set_conductivity_qp(mpdg_heat_op *op, const void *k_qp);
int mpdg_heat_op_set_conductivity_cell(mpdg_heat_op *op, const void *k_cell);
int mpdg_heat_op_set_conductivity_law(mpdg_heat_op *op, int law, const double *params4,
                                      const void *T_dofs);
/* convection/radiation boundary faces (heat_diffuse_operator.cpp:1090-1129, :721-781);
   T (device, FP64) is the temperature the radiation term is linearised about */
int mpdg_heat_op_set_robin(mpdg_heat_op *op, int face_mask, double alpha, double T_inf_conv,
                           double emissivity, double T_inf_rad, const double *T);
/* mapped geometry: node coordinates of MappingQ(k) on the FE's support points (device FP64,
   3*n_dofs: X block, Y block, Z block; read when called and by multigrid set-up, kept by the
   caller). Stores the per-quadrature-point metric JxW and JxW J^-1 J^-T like MatrixFree does for
   non-affine cells (scratch_data.cpp:192-266). Robin faces and the surface flux take their face
   geometry (surface JxW, the flux's (x, y)) from the same coordinates, so the caller keeps them
   alive while the operator is used. */
int mpdg_heat_op_set_mapping(mpdg_heat_op *op, const double *coords);
/* Gaussian surface flux on face z+ used by create_rhs: q = peak*exp(-((x-x0-vx*t)^2+(y-y0)^2)/(2 sigma^2)),
   p = {peak, sigma, x0, y0, vx, t}; peak = 0 switches it off. The reference's
   GaussProjectionIntensityProfile (include/meltpooldg/heat/laser_intensity_profiles.hpp:98-165,
   2P/(pi r^2) exp(-2 d^2/r^2)) is sigma = r/2, peak = 2P/(pi r^2). MPDG_ERR_INVALID for sigma <= 0
   or when face z+ carries a Robin term (and set_robin refuses z+ while a flux is set): the
   reference forbids Neumann and radiation/convection on one face (heat_diffuse_operator.cpp:739-743) */
int mpdg_heat_op_set_surface_flux(mpdg_heat_op *op, const double *p6);

/* Full Newton tangent with a temperature-dependent material and a prescribed constant
   velocity (tangent_local_cell_operation with do_solidification and a velocity field,
   heat_diffuse_operator.cpp:1028-1056; residual :467-527), Cartesian cells. While it is set,
   vmult applies the tangent at the temperatures last bound by mpdg_heat_op_bind_temperature,
   create_rhs evaluates the residual with this material and velocity, and the diagonal is the
   tangent's; the constant rho_cp and conductivity settings are not used. The tangent is
   non-symmetric: solve with MPDG_SOLVER_GMRES (the reference's default heat solver,
   source/heat/heat_data.cpp:11-12) and the diagonal preconditioner; the multigrid
   preconditioner and the FP32 apply refuse it (MPDG_ERR_INVALID). */
#define MPDG_TRANSITION_NONE 0  /* constant material: the liquid values */
#define MPDG_TRANSITION_MUSHY 1 /* liquid/solid cubic blend in the solid fraction
                                   clamp((liquidus - T)/(liquidus - solidus)), rho_cp = rho*c_p
                                   (include/meltpooldg/core/material.templates.hpp:296-380,
                                   :546-598, :765-778) */
typedef struct mpdg_heat_material {
  int transition;
  double conductivity[2];  /* liquid, solid */
  double specific_heat[2]; /* liquid, solid */
  double density[2];       /* liquid, solid */
  double solidus, liquidus;
} mpdg_heat_material;
/* do_solidification: add the d/dT terms of the material to the tangent (:1044-1056) */
int mpdg_heat_op_set_tangent(mpdg_heat_op *op, const mpdg_heat_material *mat, int do_solidification,
                             const double *velocity3);
/* Velocity as a DoF field instead of the constant vector of mpdg_heat_op_set_tangent: the
   reference's `velocity` vector of the operator (HeatDiffuseMultiPhaseOperator constructor,
   heat_diffuse_operator.cpp:51-65, read with vel_eval at :449-452 and :975-978). velocity: device
   FP64, 3 * n_dofs values (x block, y block, z block) on the temperature's Q_k nodes (the
   reference's default, where the velocity's FE_Q degree is the heat degree), or NULL to go back
   to the constant vector. The library keeps the pointer: the field is read by every
   mpdg_heat_op_bind_temperature and create_rhs until it is replaced, so the caller keeps it alive
   and may update its values between steps. Requires mpdg_heat_op_set_tangent (MPDG_ERR_INVALID
   otherwise); clear_tangent drops it; the tangent must be bound again afterwards. */
int mpdg_heat_op_set_velocity_field(mpdg_heat_op *op, const double *velocity);
int mpdg_heat_op_clear_tangent(mpdg_heat_op *op);
/* evaluates the tangent's per-quadrature-point data at T (and T_old for the solidification
   term; both device FP64 DoF vectors, read now) -- the reference reads them inside every
   tangent apply (do_reinit_cells, :1001-1015) */
int mpdg_heat_op_bind_temperature(mpdg_heat_op *op, const double *T, const double *T_old, void *stream);

/* vmult (heat_diffuse_operator.cpp:181-197): dst = J src, dst is overwritten */
int mpdg_heat_op_vmult(const mpdg_heat_op *op, double *dst, const double *src, void *stream);
int mpdg_heat_op_vmult_f32(const mpdg_heat_op *op, float *dst, const float *src, void *stream);
/* vmult that allocates nothing (no hipMalloc/hipFree, no allocating library call). No
   counterpart in the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid it). Same result as
   mpdg_heat_op_vmult, bit for bit (same kernels). The apply's scratch -- the contributions to
   the DoF rows where four columns of the launch meet -- is a fixed per-GPU array of 2^20
   doubles (8 MiB) instead of the operator's lazily allocated buffer (entries needed: about
   0.75 M for Q3 on 128^3 cells, 0.64 M for mapped Q3 on 96^3 cells, 0.27 M for Q4 on 96^3
   cells); meshes that need more get MPDG_ERR_TOO_LARGE. FP64
   operators only; a k(T) law on a mapped mesh needs one earlier mpdg_heat_op_vmult (which
   sizes its rescaled metric), else MPDG_ERR_INVALID. All calls on one GPU share the array:
   calls that overlap on one GPU race. */
int mpdg_heat_op_vmult_noalloc(const mpdg_heat_op *op, double *dst, const double *src, void *stream);
/* compute_inverse_diagonal_from_matrixfree (heat_diffuse_operator.cpp:273-284):
   inv = |d| > 1e-10 ? 1/d : 1, constrained rows 1 */
int mpdg_heat_op_compute_inverse_diagonal(const mpdg_heat_op *op, double *inv_diag, void *stream);
/* diagonal (not inverted) in the operator's precision */
int mpdg_heat_op_compute_diagonal(const mpdg_heat_op *op, void *diag, void *stream);
/* create_rhs (heat_diffuse_operator.cpp:785-806): dst += -R(T) with T_old and an optional
   volumetric source q (DoF vector, may be NULL); conductivity laws are evaluated at T */
int mpdg_heat_op_create_rhs(const mpdg_heat_op *op, double *dst, const double *T,
                            const double *T_old, const double *q, void *stream);
/* VectorTools::compute_norm (source/utilities/vector_tools.cpp:170-200): L2 norm of the FE function */
int mpdg_heat_op_l2_norm(const mpdg_heat_op *op, const double *u, double *norm, void *stream);

/* HIP-event timing of the cell-apply kernels launched by this operator (for
   benchmarks): enable resets the record; get returns the summed elapsed time
   of the recorded launches (waits for them) and clears the record */
int mpdg_heat_op_set_timing(mpdg_heat_op *op, int enable);
int mpdg_heat_op_get_timing(mpdg_heat_op *op, double *total_ms, int *count);

/* --- preconditioners: the Preconditioner concept update()/vmult()/reinit()
       (include/meltpooldg/linear_algebra/preconditioner.hpp:19-37) --- */
typedef struct mpdg_precond mpdg_precond;
#define MPDG_PRECOND_IDENTITY 0
#define MPDG_PRECOND_DIAGONAL 1 /* JacobiPreconditioner (preconditioner_jacobi.hpp:17-89) */
#define MPDG_PRECOND_GMG 2      /* geometric multigrid V-cycle (recipe unit_tests/gmg.cc:343-383, :447-635) */

typedef struct mpdg_gmg_params {
  int precision;          /* MPDG_FP32 (mixed precision) or MPDG_FP64 level operators */
  int n_cells_coarse;     /* stop h-coarsening at this many cells in x */
  int coarse_degree;      /* p-coarsening to this degree on the coarsest mesh (0: keep) */
  int smoother_degree;    /* Chebyshev degree */
  double smoothing_range; /* Chebyshev smoothing range */
  int eig_cg_n_iterations;/* CG steps of the eigenvalue estimate */
  int pre_smoother_degree;/* Chebyshev degree before the coarse correction (0: smoother_degree);
                             the first step from a zero guess needs no operator apply */
  int vcycle_graph;       /* 1: the V-cycle's kernels are captured into a HIP graph after each
                             update and replayed by vmult (one launch instead of one per kernel);
                             0: launched one by one. Same results either way. */
} mpdg_gmg_params;

int mpdg_precond_create(mpdg_heat_op *op, int type, const mpdg_gmg_params *gmg, mpdg_precond **p);
int mpdg_precond_update(mpdg_precond *p, void *stream);
int mpdg_precond_vmult(mpdg_precond *p, double *dst, const double *src, void *stream);
int mpdg_precond_destroy(mpdg_precond *p);
int mpdg_gmg_n_levels(const mpdg_precond *p);
int mpdg_gmg_level_info(const mpdg_precond *p, int level, int *degree, int *n_cells,
                        double *lambda_max);

/* --- LinearSolver::solve (include/meltpooldg/linear_algebra/linear_solver.hpp:21-134) --- */
#define MPDG_SOLVER_CG 0
#define MPDG_SOLVER_FCG 1 /* flexible CG (Polak-Ribiere beta) for variable preconditioners */
#define MPDG_SOLVER_GMRES 2 /* restarted GMRES, right preconditioning (linear_solver.hpp:112-118) */
#define MPDG_CRIT_REDUCTION 0 /* ReductionControl: ||r|| <= max(abs_tol, rel_tol*||r0||) */
#define MPDG_CRIT_RHS 1       /* ||b - J x|| <= rel_tol*||b|| */
typedef struct mpdg_solver_params {
  int solver_type;
  int max_iterations;
  double rel_tolerance;
  double abs_tolerance;
  int criterion;
  int gmres_restart; /* GMRES basis vectors per cycle, 1..31 (0: 28, deal.II's default
                        max_n_tmp_vectors 30 minus its two work vectors) */
} mpdg_solver_params;
/* x holds the initial guess on entry; iterations = last_step(); res = final ||r|| */
int mpdg_linear_solve(mpdg_heat_op *op, mpdg_precond *p, const mpdg_solver_params *params,
                      const double *b, double *x, int *iterations, double *res, void *stream);

/* --- HeatDiffuseOperation: predictor + Newton-Raphson + linear solver per time step
       (source/heat/heat_diffuse_operation.cpp:215-298, newton_raphson_solver.cpp:24-152,
        include/meltpooldg/linear_algebra/predictor.hpp:118-160) --- */
typedef struct mpdg_heat_operation mpdg_heat_operation;
#define MPDG_PREDICTOR_NONE 0
#define MPDG_PREDICTOR_ZERO 1
#define MPDG_PREDICTOR_LINEAR 2
#define MPDG_PREDICTOR_LEARNED 3
#define MPDG_PREDICTOR_EXTRAPOLATION 4 /* polynomial extrapolation of order extrapolation_order
                                          (1: linear_extrapolation) through the last solutions */
#define MPDG_PREDICTOR_LSQ 5 /* least_squares_projection as the reference (predictor.hpp:19-86,
                                heat_diffuse_operation.cpp:233-244): the combination of the last
                                n_old_solution_vectors solutions T_n, T_{n-1}, ... minimising
                                ||r - J x|| with r = create_rhs at T = T_old = T_n, J orthogonalised
                                images by modified Gram-Schmidt; the previous solution until as many
                                steps as solutions have run */
#define MPDG_PREDICTOR_LSQ_SOLUTION 6 /* the same projection onto the step's own right-hand side
                                b (create_rhs at T = 0, T_old = T_n), so that the combination is the
                                best solution guess in that span (no counterpart in the reference) */
#define MPDG_STEP_NEWTON 0 /* Newton-Raphson as the reference (>= 2 linear solves per step) */
#define MPDG_STEP_LINEAR 1 /* one linear solve J(T_n) T_{n+1} = b per step, coefficient frozen at T_n,
                              stopped at ||b - J T|| <= rel_tolerance ||b|| (no radiation faces) */
typedef struct mpdg_heat_operation_params {
  int predictor;
  mpdg_solver_params linear;
  int max_nonlinear_iterations;
  double residual_tolerance;
  double field_correction_tolerance;
  int step_mode;          /* MPDG_STEP_NEWTON or MPDG_STEP_LINEAR */
  int guess_check;        /* learned predictor: fall back to its base guess (the extrapolation it
                             corrects) when the learned guess's residual is not smaller */
  int guess_smoothing;    /* learned predictor (with guess_check, linear step mode): smoothing
                             steps on J T = b applied to the network's guess before the check
                             (multigrid: Chebyshev steps of its finest level; 0: none) */
  int extrapolation_order;/* MPDG_PREDICTOR_EXTRAPOLATION, and the base guess that
                             MPDG_PREDICTOR_LEARNED corrects: 1..6 (the order drops to what the
                             history holds: after set_initial the first steps use fewer points) */
  int n_old_solution_vectors; /* MPDG_PREDICTOR_LSQ(_SOLUTION): 1..7 solutions (predictor_data.hpp:27-28;
                                 until the history holds them, the previous solution is used) */
  int preconditioner_update_interval; /* linear step mode with a conductivity law: the
                                 preconditioner is rebuilt every this many steps (0 or 1: every
                                 step, as the reference; a change of dt always rebuilds it) */
  int guess_pause_after;  /* learned predictor with guess_check in linear step mode: after this many
                             consecutive steps whose learned guess the check rejected, the network
                             is not evaluated for guess_pause_steps steps (the base guess is used
                             as is), then tried again (0: never pause) */
  int guess_pause_steps;
} mpdg_heat_operation_params;
int mpdg_heat_operation_create(mpdg_heat_op *op, mpdg_precond *p,
                               const mpdg_heat_operation_params *params,
                               mpdg_heat_operation **h);
int mpdg_heat_operation_destroy(mpdg_heat_operation *h);
/* initial condition: T_{n} = T_{n-1} = T0 (device FP64) */
int mpdg_heat_operation_set_initial(mpdg_heat_operation *h, const double *T0, void *stream);
/* one implicit time step of size dt; lin_iters = accumulated Krylov iterations,
   nl_iters = Newton iterations */
int mpdg_heat_operation_solve_step(mpdg_heat_operation *h, double dt, int *lin_iters,
                                   int *nl_iters, void *stream);
/* learned predictor (MPDG_PREDICTOR_LEARNED): the network and the device DoF field of the heat
   source it reads (the caller refreshes the field before each step) */
typedef struct mpdg_learned mpdg_learned;
int mpdg_heat_operation_set_learned(mpdg_heat_operation *h, mpdg_learned *net, const double *q_field);
/* record, from the next step on, the network's inputs of every step: into the device DoF
   vector z the residual of the predictor's guess T_b (the linear extrapolation
   T_n + (dt/dt_old)(T_n - T_{n-1}), or the polynomial extrapolation of order
   extrapolation_order) scaled by the inverse operator diagonal, z = D^-1 (b - J T_b) (in
   MPDG_STEP_NEWTON b - J T_b is the Newton residual at T_b), and into base (may be NULL) T_b
   itself; z = NULL stops recording. Training pairs for mpdg_learned_train_step come from a run
   with this set. */
int mpdg_heat_operation_record_guess_feature(mpdg_heat_operation *h, double *z, double *base);
/* statistics of the last step: guess status (0: the predictor's guess was used; 1: the learned
   guess was rejected by the guess check and replaced by its base guess; 2: the learned predictor
   was paused after repeated rejections and the base guess used without evaluating the network),
   initial and final relative residuals ||b - J T||/||b|| of the step's linear solve
   (MPDG_STEP_LINEAR) */
int mpdg_heat_operation_last_step_info(const mpdg_heat_operation *h, int *fell_back, double *rel_res0,
                                       double *rel_res);
/* device pointers to the current / previous solution (owned by the handle) */
const double *mpdg_heat_operation_solution(const mpdg_heat_operation *h);
const double *mpdg_heat_operation_old_solution(const mpdg_heat_operation *h);

/* nodal field of the z+ surface flux of create_rhs (zero off the z+ face; on mapped meshes at the
   nodes' mapped (x, y)), for the network input */
int mpdg_heat_op_surface_flux_field(const mpdg_heat_op *op, double *out, void *stream);

/* Bandwidth probe for measurements (no counterpart in the reference): kind 0 copies src to dst,
   1 reads src (dst receives 4096 partial sums), 2 writes dst; n doubles (even), 16-byte aligned
   device vectors, 16-byte accesses. Gives the attainable HBM rates next to the device's spec. */
int mpdg_stream_probe(int kind, double *dst, const double *src, long n, void *stream);

/* --- data formats on either side of the path --- */
/* Paraview output as Postprocessor::do_paraview_output writes it
   (source/post_processing/postprocessor.cpp:195-270, DataOut::build_patches(mapping, k)): every
   cell subdivided into k^3 equal hexahedra whose vertices carry the FE fields' values (shared
   between neighbouring cells), base64 binary VTU. fields: n_fields host arrays of n_dofs
   lexicographic values, e.g. "temperature", "temperature_old", "heat_source"
   (heat_diffuse_operation.cpp:365-405). */
int mpdg_io_write_vtu(const char *filename, int degree, const int *n_cells, const double *extent, int n_fields,
                      const char *const *names, const double *const *fields);
/* .pvd record relating the VTU files to times (DataOutBase::write_pvd_record) */
int mpdg_io_write_pvd(const char *filename, int n, const double *times, const char *const *vtu_names);
/* DoF numbering of a DoFHandler on the same box: points (host, n x 3) are its support points
   (DoFTools::map_dofs_to_support_points); lex_index[i] (host) = lexicographic node of DoF i.
   MPDG_ERR_INVALID when a point is off the node grid or two DoFs share a node. */
int mpdg_io_dof_permutation(int degree, const int *n_cells, const double *extent, const double *points, long n,
                            long *lex_index);
/* device vectors: to_lex: dst[lex_index[i]] = src[i]; else dst[i] = src[lex_index[i]]
   (lex_index: a device copy of mpdg_io_dof_permutation's output; src != dst) */
int mpdg_io_permute(const long *lex_index, const double *src, double *dst, long n, int to_lex, void *stream);

/* --- learned initial guess: 3D CNN (4 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3 taps, ReLU) on the
       DoF grid, BF16 MFMA kernels, FP32 master weights, Adam. Enters the predictor slot of
       include/meltpooldg/linear_algebra/predictor.hpp:118-160 as a new predictor type. --- */
/* network for the DoF grid of Q_degree on n_cells[3] cells; it works on the DoF grid of
   Q_degree on n_cells / coarsening cells (coarsening 1, 2, 4 or 8, dividing n_cells; with
   coarsening > 1 at most 1024 DoFs along x): fields are carried there by full weighting and the
   correction back by the FE prolongation. qscale multiplies the heat-source channel (e.g.
   dt / (rho c_p h)). MPDG_ERR_INVALID otherwise. */
int mpdg_learned_create(const int *n_cells, int degree, int coarsening, double qscale,
                        unsigned long long seed, mpdg_learned **net);
/* output scale: the correction is oscale * rms(z) * y; the first train step sets it from its
   sample (target rms 1) unless set before. Part of a checkpoint next to the weights. */
int mpdg_learned_get_output_scale(const mpdg_learned *net, double *oscale);
int mpdg_learned_set_output_scale(mpdg_learned *net, double oscale);
/* DoF grid of the fields it takes and its working grid (either pointer may be NULL) */
int mpdg_learned_grid(const mpdg_learned *net, int *dof_dims, int *net_dims);
int mpdg_learned_destroy(mpdg_learned *net);
int mpdg_learned_n_params(const mpdg_learned *net);
/* weights as one FP32 host array: per layer W[co][ci][tap] then b[co] */
int mpdg_learned_get_weights(const mpdg_learned *net, float *host);
int mpdg_learned_set_weights(mpdg_learned *net, const float *host);
/* out = T_b + P (oscale sz y(T_{n-1}, T_n, q, z)), sz = rms(z) on the network grid: the base
   guess T_b (base; NULL: the linear extrapolation 2 T_n - T_{n-1}, equal steps) plus the learned
   correction. Device FP64 DoF vectors; z (the residual input of T_b) and base are recorded by
   mpdg_heat_operation_record_guess_feature; q and z may be NULL (channel zero; without z,
   sz = rms(T_n - T_{n-1})); out may be base */
int mpdg_learned_predict(mpdg_learned *net, const double *T_nm1, const double *T_n, const double *q,
                         const double *z, const double *base, double *out, void *stream);
/* one Adam step on the sample (T_{n-1}, T_n, q, z, T_b) -> T_{n+1}; loss = mean squared error of
   y against W(T_{n+1} - T_b)/(oscale sz) before the step (W = full weighting to the network
   grid; base NULL: T_b = 2 T_n - T_{n-1}) */
int mpdg_learned_train_step(mpdg_learned *net, const double *T_nm1, const double *T_n, const double *q,
                            const double *z, const double *base, const double *T_np1, double lr, double *loss,
                            void *stream);
/* gradients of the loss (host FP32, weight layout) without an update */
int mpdg_learned_gradients(mpdg_learned *net, const double *T_nm1, const double *T_n, const double *q,
                           const double *z, const double *base, const double *T_np1, float *host_grad,
                           double *loss, void *stream);
/* layer activations of the last forward pass as bf16 bit patterns [voxel][16] on the network grid
   (layer 0 = input) */
int mpdg_learned_activation(const mpdg_learned *net, int layer, unsigned short *host);

#ifdef __cplusplus
}
#endif
#endif
