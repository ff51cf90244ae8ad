/*
 * CPU oracle: test infrastructure only. A plain restatement of the
 * reference's DG algorithms (deal.II FEEvaluation semantics) used as the
 * expected result in the parity tests of the GPU library and as the CPU
 * baseline of bench.py.
 * Nothing in the product links or calls this library.
 *
 * Parity unpinned: the mesh's cell order and DoF numbering
 * (or_cell_coordinates, or_neighbors, or_cell_vertices: the refine_global
 * child order, lexicographic DoFs per cell) follow a restated rule that the
 * tests check, but nothing here can compare them with deal.II's DoFHandler
 * itself. Every other function is checked
 * against closed forms, known answers, invariants or convergence orders.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_problem
{
  int      dim, degree, n_components;
  int      operator_kind; /* 0 advection, 1 advection + SIPG, 2 Burgers */
  int      n_coarse[3], refine;
  double   lo[3], hi[3];
  int      periodic[3];
  double   perturbation;
  uint64_t seed;
  int      velocity_kind;
  double   velocity[3];
  int      time_factor; /* 0: 1, 1: cos(pi t / final_time) */
  double   final_time;
  double   flux_alpha, factor_skew;
  double   viscosity, penalty_factor;
  int      solution_kind;
  double   solution_param[8];
  int      irk_stages;
  /* QGaussLobatto(k+1) for the operator and the mass matrix, or for the
     mass matrix only (advection_solver.cc:110-116, :440-445) */
  int      use_gl_quad, use_gl_quad_mass;
} or_problem;

typedef struct or_handle or_handle;

or_handle *or_create(const or_problem *p);
void       or_destroy(or_handle *h);
int64_t    or_n_cells(const or_handle *h);
int64_t    or_n_dofs(const or_handle *h);
void       or_set_threads(int n);

/* mesh export, same layout as the library's */
void or_cell_coordinates(const or_handle *h, int32_t *ijk);
void or_neighbors(const or_handle *h, int32_t *nb);
void or_cell_vertices(const or_handle *h, double *v);
double or_min_vertex_distance(const or_handle *h);

/* 1D data: gauss points/weights on [0,1], Gauss-Lobatto nodes */
void or_basis_1d(int n, double *gauss_pts, double *gauss_wts, double *gl_pts);

/* explicit operator  dst = M^-1(-L(t) src)  (advection_solver.cc:661-691) */
void or_apply(or_handle *h, const double *src, double *dst, double t);
/* operator loop only (no inverse mass), sign convention of the explicit
 * files: dst = -L(t) src  (advection_solver.cc:671-679)                  */
void or_loop(or_handle *h, const double *src, double *dst, double t);
void or_inverse_mass(or_handle *h, const double *src, double *dst);
void or_mass(or_handle *h, const double *src, double *dst);
/* perform_stage with the reference's two passes (advection_solver.cc:695-756) */
void or_perform_stage(or_handle *h, double t, double b_dt, double a_dt,
                      const double *current_ri, double *vec_ki,
                      double *solution, double *next_ri);
/* one LSRK time step, reference recursion (advection_solver.cc:906-939) */
void or_lsrk_time_step(or_handle *h, int scheme, double t, double dt,
                       double *solution, double *vec_ri, double *vec_ki);
void or_project(or_handle *h, double *dst, double t);
void or_mass_energy(or_handle *h, const double *vec, double out[3]);
void or_l2_error(or_handle *h, const double *vec, double t, double out[2]);

/* implicit single stage: dst = (M/dt + L(t+dt)) src
 * (advection_solver_implicit_real.cc:876-891, :988-1058)                  */
void or_vmult(or_handle *h, double t, double dt, const double *src,
              double *dst);
/* IRK (advection_solver_irk_real.cc:1266-1326, :1561-1703, :1763-1951)  */
void or_irk_coefficients(int s, double *A, double *b, double *c);
void or_irk_compute_rhs(or_handle *h, double t, double dt, const double *src,
                        double *dst);
void or_irk_vmult(or_handle *h, double t, double dt, const double *src,
                  double *dst);
/* Burgers right-hand side M^-1(-L(u)) (burgers_solver.cc:259-410)       */
void or_burgers_apply(or_handle *h, const double *src, double *dst);

void or_irk_update_solution(or_handle *h, double dt, double *solution,
                            const double *stage);

#ifdef __cplusplus
}
#endif
#endif
