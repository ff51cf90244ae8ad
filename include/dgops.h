/*
 * dgops: MI355X (gfx950) matrix-free discontinuous-Galerkin operators and
 * time integrators for linear advection (+ SIPG diffusion, inviscid
 * Burgers), C ABI.
 *
 * Every entry point returns DG_OK (0) or an error code; dg_last_error()
 * describes the most recent failure of the calling thread. Vectors are
 * device pointers owned by the caller, in the dof numbering of the
 * reference's DoFHandler: cells in Triangulation order (coarse cells
 * lexicographic, refine_global children in Morton order), (k+1)^dim
 * Gauss-Lobatto nodal values per cell, lexicographic (x fastest); vector
 * valued fields (Burgers) store their components per cell one after the
 * other; block vectors (IRK stages) store their blocks one after the other.
 * Operators with precision DG_FP32 take float vectors, all others double.
 * One handle is used by one thread; work is queued on the handle's stream
 * and the calls that return host scalars synchronize that stream.
 *
 * Reference interfaces replaced (file:line in kronbichler/advection_miniapp):
 *   AdvectionOperation (explicit)   advection_solver.cc:330-421
 *   AdvectionOperation (implicit)   advection_solver_implicit_real.cc:816-909
 *   AdvectionOperation (IRK)        advection_solver_irk_real.cc:1227-1331
 *   LowStorageRungeKuttaIntegrator  advection_solver.cc:823-944
 *   DiagonallyImplicitRungeKutta... advection_solver_implicit_real.cc:1561-1748
 *   BurgersOperation                burgers_solver.cc:120-225
 *   SIPG advection-diffusion        advection_diffusion_solver.cc:360-613
 */
#ifndef DGOPS_H
#define DGOPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dg_status
{
  DG_OK              = 0,
  DG_ERR_ARGUMENT    = 1, /* invalid argument (null pointer, size, ...)  */
  DG_ERR_UNSUPPORTED = 2, /* combination not built (degree, dim, ...)    */
  DG_ERR_HIP         = 3, /* HIP runtime error                           */
  DG_ERR_NO_CONVERGENCE = 4, /* SolverControl::NoConvergence            */
  DG_ERR_STATE       = 5  /* call not valid for this operator kind       */
};

enum dg_precision
{
  DG_FP64 = 0,
  DG_FP32 = 1
};

enum dg_operator_kind
{
  DG_OP_ADVECTION           = 0, /* advection_solver.cc, _variable.cc        */
  DG_OP_ADVECTION_DIFFUSION = 1, /* advection + SIPG, advection_diffusion_... */
  DG_OP_BURGERS             = 2  /* burgers_solver.cc                        */
};

enum dg_velocity_kind
{
  DG_VEL_CONSTANT    = 0, /* velocity[]                                   */
  DG_VEL_CELLULAR_3D = 1, /* (sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy) */
  DG_VEL_VORTEX      = 2, /* advection_solver_variable.cc:153-183          */
  DG_VEL_VORTEX_PERT = 3  /* advection_solver_implicit_real.cc:159-190     */
};

enum dg_solution_kind
{
  DG_SOL_ZERO         = 0,
  DG_SOL_WAVE         = 1, /* advection_solver.cc:127-183                   */
  DG_SOL_SINE_PRODUCT = 2, /* prod sin(p0 (x_d - a_d t)) exp(-p1 dim p0^2 t) */
  DG_SOL_GAUSSIAN     = 3, /* exp(-|x - (p0,p1,p2)|^2 / p3^2)               */
  DG_SOL_POLYNOMIAL   = 4, /* (1 + sum_d (d+1)(x_d - a_d t))^p0             */
  DG_SOL_TWO_PULSES   = 5, /* advection_solver_variable.cc:141-147          */
  DG_SOL_BURGERS_SINE = 6  /* burgers_solver.cc:109-117                     */
};

enum dg_lsrk_scheme
{
  /* advection_solver.cc:83-90, coefficients :828-897 */
  DG_LSRK_3_3 = 0,
  DG_LSRK_5_4 = 1,
  DG_LSRK_7_4 = 2,
  DG_LSRK_9_5 = 3
};

enum dg_time_factor_kind
{
  DG_TIME_FACTOR_ONE = 0, /* velocity independent of time                  */
  DG_TIME_FACTOR_COS = 1  /* cos(pi t / final_time), the vortex files     */
};

typedef struct dg_config
{
  int      dim;           /* 2 or 3                                       */
  int      degree;        /* polynomial degree k                          */
  int      n_components;  /* 1, or dim for Burgers                        */
  int      operator_kind; /* dg_operator_kind                             */
  int      precision;     /* dg_precision of the operator arithmetic      */
  /* mesh: n_coarse[d] lexicographic coarse cells, refine_global(refine)  */
  int      n_coarse[3];
  int      refine;
  double   lo[3], hi[3];
  int      periodic[3];
  double   perturbation;  /* interior vertex offset, fraction of h        */
  uint64_t seed;          /* mt19937_64 seed of the perturbation          */
  /* transport */
  int      velocity_kind; /* dg_velocity_kind                             */
  double   velocity[3];
  int      time_factor;   /* dg_time_factor_kind                          */
  double   final_time;    /* T in cos(pi t / T)                           */
  double   flux_alpha;    /* 0 central, 1 upwind                          */
  double   factor_skew;   /* 0 conservative, 0.5 skew-symmetric           */
  /* diffusion (DG_OP_ADVECTION_DIFFUSION)                                */
  double   viscosity;
  double   penalty_factor; /* sigma = penalty_factor * (1/h- + 1/h+) / 2  */
  /* analytic solution: initial data and Dirichlet boundary values        */
  int      solution_kind;
  double   solution_param[8];
  /* Radau IIA stages of the IRK entry points (1..4; dg_irk_coefficients
     alone also tabulates 5)                                               */
  int      irk_stages;
  /* quadrature switches of advection_solver.cc:110-116 (:440-445):
     use_gl_quad integrates the operator and the mass matrix with
     QGaussLobatto(k+1) instead of QGauss(k+1); use_gl_quad_mass only the
     mass matrix (inverse mass, and the apply and stages through it).
     Not for the Burgers operator.                                         */
  int      use_gl_quad;
  int      use_gl_quad_mass;
} dg_config;

typedef struct dg_info
{
  int64_t n_cells;
  int64_t n_dofs;          /* per component and block                     */
  int64_t vector_size;     /* n_dofs * n_components                       */
  int     dofs_per_cell;
  int     cells_per_direction[3];
  double  min_vertex_distance;
  double  h[3];
} dg_info;

typedef struct dg_operator dg_operator;

const char *dg_last_error(void);
void        dg_config_default(dg_config *cfg);

/* Builds mesh, dof numbering, geometry and 1D tables, uploads them to the
 * current device; work is queued on `stream` (hipStream_t, 0 = null).
 * replaces AdvectionOperation::reinit            advection_solver.cc:435-465 */
int dg_create(const dg_config *cfg, void *stream, dg_operator **op);
int dg_destroy(dg_operator *op);
int dg_get_info(const dg_operator *op, dg_info *info);
int dg_set_stream(dg_operator *op, void *stream);

/* ---- tile shapes of the operator kernels (per-degree autotuner) -------
 * The explicit right-hand side kernels of 3D operators (apply and
 * perform_stage, advection and SIPG, FP64 and FP32) are built for three
 * tile shapes: one, two or four wavefronts of cell lines per workgroup.
 * By default the first launch of such a variant on a device (meshes of
 * 4096 cells and more) times every shape with HIP events on the handle's
 * stream, into scratch vectors, and keeps the fastest for the process.
 * The shape changes the time, never the result.
 *                  replaces MatrixFree's batch setup advection_solver.cc:436-465 */
/* force the cells per workgroup (one of the shapes of the degree); 0
 * restores the autotuner                                                */
int dg_set_tile_shape(dg_operator *op, int cells_per_workgroup);
/* the shape of the handle's last operator launch and its tuned time (0
 * when not tuned); candidates (may be null) receives 3 pairs {cells per
 * workgroup, tuned ms}, ms -1 where a shape was not built or not timed  */
int dg_get_tile_shape(const dg_operator *op, int *cells_per_workgroup,
                      double *tuned_ms, double *candidates);

/* ---- explicit operator ------------------------------------------------ */

/* dst = M^-1 (-L(t) src)            AdvectionOperation::apply
 *                                    advection_solver.cc:374-377, :661-691 */
int dg_apply(dg_operator *op, const void *src, void *dst, double current_time);

/* k = M^-1(-L(t) current_ri); solution += b_dt k; next_ri = solution_old +
 * a_dt k (skipped for a_dt == 0). Any of the vectors may alias as in the
 * reference's integrator; vec_ki is scratch.
 *                     AdvectionOperation::perform_stage advection_solver.cc:379-386,
 *                     :695-756 */
int dg_perform_stage(dg_operator *op,
                     double       current_time,
                     double       b_dt,
                     double       a_dt,
                     const void  *current_ri,
                     void        *vec_ki,
                     void        *solution,
                     void        *next_ri);

/* one step of LowStorageRungeKuttaIntegrator::perform_time_step
 * (advection_solver.cc:906-939): result in `solution`, vec_ri and vec_ki
 * are scratch registers of the same size.                                 */
int dg_lsrk_time_step(dg_operator *op,
                      int          scheme,
                      double       current_time,
                      double       time_step,
                      void        *solution,
                      void        *vec_ri,
                      void        *vec_ki);
/* n_steps steps of dg_lsrk_time_step from current_time (the time loop of
 * AdvectionProblem::run, advection_solver.cc:1264-1299), with the same
 * result bit for bit. The first step is launched on the stream. When the
 * operator does not depend on time (constant speed factor, no boundary
 * faces), runs of DG_GRAPH_STEPS later steps may replay one captured HIP
 * graph, cached per (scheme, time step, registers): the first call with
 * at least 3 DG_GRAPH_STEPS further steps times both launch forms on
 * them and keeps the faster (dg_get_step_launch). Other steps are
 * launched one by one.                                                   */
#define DG_GRAPH_STEPS 8
int dg_lsrk_time_steps(dg_operator *op,
                       int          scheme,
                       double       current_time,
                       double       time_step,
                       int          n_steps,
                       void        *solution,
                       void        *vec_ri,
                       void        *vec_ki);
/* launch form of dg_lsrk_time_steps: 0 measured on the first long call
 * (default), 1 stream launches only, 2 graph replay wherever allowed. No
 * counterpart in the reference (its time loop calls perform_time_step
 * per step, advection_solver.cc:1264-1299); the form changes the time,
 * never the result.                                                      */
int dg_set_step_launch(dg_operator *op, int mode);
/* the form in use (1 or 2; 0 before it was measured) and the measured
 * time of DG_GRAPH_STEPS steps in each form (ms, 0 when not measured)    */
int dg_get_step_launch(const dg_operator *op, int *mode, double *eager_ms,
                       double *graph_ms);
int dg_lsrk_n_stages(int scheme);
/* the scheme's b_i and a_i (n_stages values each, a_i of the last stage
 * 0), for integrators written on top of dg_perform_stage; returns the
 * number of stages or -1          LowStorageRungeKuttaIntegrator
 *                                  advection_solver.cc:828-897          */
int dg_lsrk_coefficients(int scheme, double *bi, double *ai);

/* L2 projection of the analytic solution at time t with QGauss(k+1)
 *                      AdvectionOperation::project_initial advection_solver.cc:760-784 */
int dg_project_initial(dg_operator *op, void *dst, double t);

/* (mass, energy, H1 seminorm)   compute_mass_and_energy advection_solver.cc:787-818 */
int dg_compute_mass_and_energy(dg_operator *op, const void *vec, double out[3]);

/* out[0] = ||u_h - u(t)||_L2 with QGauss(k+2), out[1] = ||u(t)||_L2 with
 * QGauss(k+1)                    output_results advection_solver.cc:1183-1215 */
int dg_l2_error(dg_operator *op, const void *vec, double t, double out[2]);

/* VTU file of a solution at time t: fields "solution" and
 * "analytic_solution" (the exact solution interpolated into the DG space),
 * one VTK Lagrange cell of the operator's degree per cell with points at
 * equidistant reference positions (DataOut with write_higher_order_cells
 * and build_patches(mapping, degree)); vec is a device vector of the
 * operator's precision        output_results advection_solver.cc:1217-1241 */
int dg_write_vtu(dg_operator *op, const void *vec, double t,
                 const char *filename);
/* the same file from a host vector (FP64, no GPU needed)                 */
int dg_host_write_vtu(const dg_config *cfg, const double *vec, double t,
                      const char *filename);

/* ---- implicit single stage operator ----------------------------------- */

/* advection_solver_implicit_real.cc:869-873                               */
int dg_set_time(dg_operator *op, double current_time, double time_step);
/* dst = (M/dt + L(t+dt)) src, skew form    vmult  advection_solver_implicit_real.cc:876-891 */
int dg_vmult(dg_operator *op, void *dst, const void *src);
/* dst = M^-1 dst            apply_inverse_mass_matrix :894-909             */
int dg_apply_inverse_mass_matrix(dg_operator *op, void *dst);
/* cellwise fast-diagonalization inverse of the separable part of
 * M/dt + L, computed in `prec_precision`     precondition_block_jacobi
 *                                             advection_solver_implicit_real.cc:1470-1537 */
int dg_precondition_block_jacobi(dg_operator *op, void *dst, const void *src);

/* ---- implicit Runge-Kutta (Radau IIA) --------------------------------- */

/* dst[s] = -L(t + c_s dt) src / dt          compute_rhs advection_solver_irk_real.cc:1266-1282 */
int dg_irk_compute_rhs(dg_operator *op, void *dst_blocks, const void *src);
/* dst = (A^-1 (x) M)/dt src + (I (x) L) src   vmult :1310-1326               */
int dg_irk_vmult(dg_operator *op, void *dst_blocks, const void *src_blocks);
/* solution += dt sum_c b_c sum_b A^-1_cb stage_b  update_solution :1284-1308 */
int dg_irk_update_solution(dg_operator *op, void *solution, const void *stage_blocks);
/* stage-diagonalized cellwise FDM              precondition_block_jacobi :2061-2136 */
int dg_irk_precondition_block_jacobi(dg_operator *op, void *dst_blocks, const void *src_blocks);
/* Radau IIA coefficients: A (s*s), b (s), c (s)      class IRK :344-454 */
int dg_irk_coefficients(int stages, double *A, double *b, double *c);

/* ---- solvers ------------------------------------------------------------ */

typedef struct dg_solver_control
{
  int    max_iterations;  /* SolverControl(max_it, tol)                  */
  double abs_tolerance;   /* converged when |r| <= max(abs_tolerance,    */
  double rel_tolerance;   /*   rel_tolerance * |rhs|)                     */
  int    restart;         /* max_basis_size                              */
  int    flexible;        /* 1: FGMRES, 0: right-preconditioned GMRES    */
  int    prec_precision;  /* dg_precision of the FDM preconditioner       */
  int    inner_iterations; /* 1..16: inner GMRES of that many steps as
                             preconditioner, in prec_precision; 0: none */
  int    cell_solver;     /* 1: the block-Jacobi step is the batched cell
                             solver (do_batched_solver, one GMRES step per
                             cell in the collocation basis) instead of the
                             FDM alone       advection_solver_irk_real.cc:2108-2126 */
  /* outputs */
  int    iterations;
  double initial_residual;
  double final_residual;
} dg_solver_control;

/* Block-Jacobi step with the batched cell solver: per cell one
 * right-preconditioned GMRES step on the cell-local stage operator in the
 * Gauss-point collocation basis, preconditioned by the FDM of that basis
 *         precondition_block_jacobi, do_batched_solver = true,
 *                                  advection_solver_irk_real.cc:2079-2126 */
int dg_irk_precondition_cell_solver(dg_operator *op, void *dst,
                                    const void *src);
/* The cell-local stage operator of that solver: values at the Gauss
 * points of every cell (lexicographic) in, the weak form tested with the
 * Lagrange functions of the Gauss points out, neighbour traces zero
 *              CellwiseOperator::vmult advection_solver_irk_real.cc:545-711 */
int dg_irk_cell_operator(dg_operator *op, void *dst, const void *src);

/* Solve the IRK stage system (A^-1 (x) M)/dt Z + (I (x) L) Z = rhs with
 * (F)GMRES and the FDM preconditioner; `stage_blocks` holds the initial
 * guess on entry.   SolverFGMRES::solve      advection_solver_irk_real.cc:2725-2757 */
int dg_irk_solve(dg_operator *op, void *stage_blocks, const void *rhs_blocks,
                 dg_solver_control *control);

/* The preconditioner that dg_irk_solve applies for `control` (FP64 block
 * vectors): the cellwise FDM block-Jacobi step in control->prec_precision
 * (FP32 arithmetic on the FP64 vectors for DG_FP32), the batched cell
 * solver (cell_solver), or an inner GMRES of inner_iterations steps around
 * either; the outputs of `control` are not touched.
 *                          BlockJacobi::vmult advection_solver_irk_real.cc:2462-2479 */
int dg_irk_block_jacobi_vmult(dg_operator *op,
                              const dg_solver_control *control, void *dst,
                              const void *src);

/* One Radau IIA step u_{n+1} = u_n + dt b^T A^-1 Z, with a zero initial
 * guess.                               run() loop :2533-2779 without projection */
int dg_irk_time_step(dg_operator *op, double current_time, double time_step,
                     void *solution, dg_solver_control *control);

/* Radau IIA integrator that keeps the last four stage solutions and starts
 * every solve from their least-squares extrapolation (fused Gram matrix,
 * Cholesky, refresh when the initial residual doubles or after 20 steps)
 *                                   run() advection_solver_irk_real.cc:2525-2773 */
typedef struct dg_irk_integrator dg_irk_integrator;
int dg_irk_integrator_create(dg_operator *op, dg_irk_integrator **it);
int dg_irk_integrator_destroy(dg_irk_integrator *it);
/* one step of size dt from t; `project_sol` (4 values, may be null)
 * receives the extrapolation weights                                    */
int dg_irk_integrator_step(dg_irk_integrator *it, double current_time,
                           double time_step, void *solution,
                           dg_solver_control *control, double *project_sol);

/* One DIRK step with per-stage (F)GMRES solves    perform_time_step
 *                                                 advection_solver_implicit_real.cc:1653-1731 */
int dg_dirk_time_step(dg_operator *op, int n_stages, double current_time,
                      double time_step, void *solution,
                      dg_solver_control *control);

/* ---- vector operations on the operator's stream ------------------------ */

/* precision of the vectors: DG_FP64 or DG_FP32                          */
int dg_vector_dot(dg_operator *op, int precision, int64_t n, const void *x,
                  const void *y, double *result);
/* result[j] = <x_j, y> for j < m, x_j = x + j * n                       */
int dg_vector_multi_dot(dg_operator *op, int precision, int64_t n, int m,
                        const void *x, const void *y, double *result);
/* dg_vector_multi_dot without any device allocation: no counterpart in
 * the reference. For callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). The block partials and the sums live in a static device
 * array of 131,584 bytes per GPU, held until the process ends; m <= 64
 * (larger m is rejected with DG_ERR_ARGUMENT). Every call on one GPU
 * shares that array: calls that overlap on one GPU race.                  */
int dg_vector_multi_dot_noalloc(dg_operator *op, int precision, int64_t n,
                                int m, const void *x, const void *y,
                                double *result);
/* y += sum_j alpha[j] x_j                                                */
int dg_vector_multi_axpy(dg_operator *op, int precision, int64_t n, int m,
                         const double *alpha, const void *x, void *y);
/* y = a x + b y                                                          */
int dg_vector_axpby(dg_operator *op, int precision, int64_t n, double a,
                    const void *x, double b, void *y);
int dg_vector_copy_convert(dg_operator *op, int64_t n, int src_precision,
                           const void *src, int dst_precision, void *dst);

/* ---- index data (deal.II numbering) ---------------------------------- */

/* integer cell coordinates [n_cells][3] in dof order                     */
int dg_export_cell_coordinates(const dg_operator *op, int32_t *ijk);
/* neighbours [n_cells][2 dim], face f = 2 d + side, -1 at a boundary     */
int dg_export_neighbors(const dg_operator *op, int32_t *neighbors);
/* corners [n_cells][2^dim][dim]                                          */
int dg_export_cell_vertices(const dg_operator *op, double *vertices);
/* Host-only index map (no GPU needed): the number of cells, the centre of
 * every cell (mean of its corners, [cell][dim]) and the offset of its
 * first dof in the vectors. Null arrays are skipped, so a first call with
 * both null returns n_cells. The numbering is the one of the reference's
 * DoFHandler on the refine_global mesh (advection_solver.cc:1002-1174),
 * printed by tests/cpp/index_map for comparison with a deal.II build.  */
int dg_host_index_map(const dg_config *cfg, int64_t *n_cells,
                      double *cell_centres, int64_t *dof_offsets);

/* ---- serving: many independent explicit solves on one GPU -------------
 * No counterpart in the reference (its drivers run one solve per process);
 * this is the request loop of the serving workload: 3D periodic unit
 * cubes with a constant transport speed and Gaussian initial data, LSRK(5,4)
 * to final_time with the time step rule of advection_solver.cc:1163-1167.
 * Requests are admitted first come, first served while their device
 * memory fits the budget; dg_server_step advances all admitted requests
 * time step by time step, with one kernel launch per LSRK stage for all
 * requests of one degree (continuous batching: requests join and leave at
 * step boundaries). A request's result is bit-identical to the same solve
 * run alone with dg_lsrk_time_step.                                      */
typedef struct dg_serve_request
{
  int    degree;              /* 1..7                                       */
  int    cells_per_direction; /* c 2^r with c in {1, 3, 5, 7}               */
  double velocity[3];         /* constant transport speed                  */
  double center[3], width;    /* exp(-|x - center|^2 / width^2)           */
  double final_time;
  double courant;             /* dt = C h / |a| / k^1.5, then shortened    */
} dg_serve_request;

enum dg_request_state
{
  DG_REQUEST_QUEUED  = 0,
  DG_REQUEST_ACTIVE  = 1,
  DG_REQUEST_DONE    = 2,
  DG_REQUEST_UNKNOWN = 3,
  DG_REQUEST_FAILED  = 4  /* its operator or vectors could not be set up;
                             the step call that admitted it reported why */
};

typedef struct dg_server dg_server;

/* the operator configuration, time step and step count of a request (host
 * only; the same solve can be run alone through dg_create)              */
int dg_serve_request_config(const dg_serve_request *r, dg_config *cfg,
                            double *time_step, int *n_steps);
int dg_server_create(int64_t hbm_budget_bytes, void *stream,
                     dg_server **server);
int dg_server_destroy(dg_server *s);
int dg_server_submit(dg_server *s, const dg_serve_request *r, int64_t *id);
/* admit what fits, then advance every active request by up to max_steps
 * time steps (fewer when a request finishes earlier: the call returns at
 * the first completion) without synchronizing in between, and retire the
 * finished requests; n_finished counts them                              */
int dg_server_step(dg_server *s, int max_steps, int *n_active, int *n_queued,
                   int *n_finished);
int dg_server_request_info(dg_server *s, int64_t id, int *state,
                           int64_t *n_dofs, int *steps_done,
                           int *steps_total);
/* copy a finished request's solution (n_dofs doubles, device pointer dst)
 * and release it; a FAILED request is released without a copy and the
 * call returns the error of its setup                                     */
int dg_server_take_result(dg_server *s, int64_t id, void *dst);
/* device memory held by admitted and finished requests                  */
int dg_server_memory(dg_server *s, int64_t *bytes_in_use);

#ifdef __cplusplus
}
#endif

#endif
