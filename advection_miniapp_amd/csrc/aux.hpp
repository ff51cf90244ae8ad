// Host wrappers of the auxiliary kernels: projection, integrals, vector
// operations and pointwise time-integrator updates.
#pragma once

#include "internal.hpp"

namespace dgops
{
  void aux_project(const Operator &op, void *dst, double t);
  void aux_mass_energy(const Operator &op, const void *vec, double out[3]);
  void aux_l2_error(const Operator &op, const void *vec, double t,
                    double out[2]);

  // per cell: volume average of the reference-space velocity J^-1 a(x)
  // (at time factor 1) and det(J^-1) at the first quadrature point
  // (advection_solver_irk_real.cc:1543-1556, :2102)
  void aux_fdm_cell_data(const Operator &op, double *velocity,
                         double *invdet);

  // vec = M^-1 vec in place, precision p
  // (CellwiseInverseMassMatrix::apply, advection_solver_implicit_real.cc:894-909)
  void aux_inverse_mass(const Operator &op, int p, void *vec);
  // dst += scale * (SIPG diffusion terms of src, integrated form, sign
  // P.sign) on cells of any shape (FP64)
  void aux_sipg_general(Operator &op, const OpParams &P,
                        const void *src, void *dst, double scale);

  // VTU file of a host solution vector (vtu.cpp)
  void write_vtu_host(const dg_config &cfg, const StructuredMesh &mesh,
                      const double *values, double time, const char *filename);

  // boundary face index and face point tables (every geometry mode)
  void aux_boundary_tables(Operator &op);
  // geometry tables of the operator kernels in precision p (see OpParams);
  // includes the boundary tables
  void aux_geometry_tables(Operator &op, int p);
  // analytic Dirichlet data (P.sol_kind at P.sol_time) on every boundary
  // face point, read by the operator kernels (no-op without boundary
  // faces)
  void aux_boundary_values(Operator &op, const OpParams &P);

  // vector operations, precision p (DG_FP64 / DG_FP32)
  void vec_axpby(Operator &op, int p, int64_t n, double a, const void *x,
                 double b, void *y);
  void vec_multi_axpy(Operator &op, int p, int64_t n, int m,
                      const double *alpha, const void *x, int64_t x_stride,
                      void *y);
  // y = (y + sum_j alpha_j x_j) * scale
  void vec_multi_axpy_scaled(Operator &op, int p, int64_t n, int m,
                             const double *alpha, const void *x,
                             int64_t x_stride, void *y, double scale);
  void vec_multi_dot(Operator &op, int p, int64_t n, int m, const void *x,
                     int64_t x_stride, const void *y, double *result);
  // y += sum_j alpha_j x_j, then result[j] = <x_j, y> (j < m) and
  // result[m] = <y, y> of the updated y, in one pass for m <= 32
  void vec_multi_axpy_dot(Operator &op, int p, int64_t n, int m,
                          const double *alpha, const void *x,
                          int64_t x_stride, void *y, double *result);
  void vec_copy_convert(Operator &op, int64_t n, int ps, const void *src,
                        int pd, void *dst);

  // Arnoldi steps with their scalars in device memory (d_ pointers), for
  // a Krylov method without host round trips; m <= 16 everywhere.
  // d_out[j] = <x_j, y>, j < m
  // vec_multi_dot with its partials and sums in a static device array
  // instead of allocated buffers (m <= 64): no allocation at all
  void vec_multi_dot_noalloc(Operator &op, int p, int64_t n, int m,
                             const void *x, int64_t xs, const void *y,
                             double *result);
  void vec_multi_dot_dev(Operator &op, int p, int64_t n, int m,
                         const void *x, int64_t x_stride, const void *y,
                         double *d_out);
  // x *= 1 / sqrt(*d_s2)
  void vec_scale_inv_sqrt(Operator &op, int p, int64_t n, const double *d_s2,
                          void *x);
  // w = (w - sum_{k<m} h_k V_k) / sqrt(h_m - sum h_k^2), with d_h[0..m-1]
  // = <V_k, w> and d_h[m] = <w, w> (one classical Gram-Schmidt sweep)
  void vec_arnoldi_update(Operator &op, int p, int64_t n, int m,
                          const double *d_h, const void *V, int64_t v_stride,
                          void *w);
  // d_y = argmin |sqrt(*d_beta2) e1 - H y|, column j of the Hessenberg
  // matrix at d_h + j h_stride as written by vec_multi_dot_dev for
  // vec_arnoldi_update
  void vec_arnoldi_lsq(Operator &op, int m, const double *d_h,
                       int64_t h_stride, const double *d_beta2, double *d_y);
  // FP64 x = sum_{k<m} d_y[k] Z_k, Z in precision p
  void vec_combine(Operator &op, int p, int64_t n, int m, const double *d_y,
                   const void *Z, int64_t z_stride, double *x);
  // one pass over four vectors and a right-hand side: the lower triangle of
  // their Gram matrix (10 values, row by row) and their products with rhs
  // (4 values), as the fused loop of advection_solver_irk_real.cc:2551-2645
  void vec_gram4(Operator &op, int64_t n, const double *const v[4],
                 const double *rhs, double out[14]);
  // sol_out = sol_in + b k; ri_out = sol_in + a k (ri_out may be null)
  void vec_lsrk_update(Operator &op, int p, int64_t n, double b, double a,
                       const void *k, const void *sol_in, void *sol_out,
                       void *ri_out);
} // namespace dgops
